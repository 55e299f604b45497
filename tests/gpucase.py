"""Shared set-up of GPU contexts and oracle contexts on the reference case."""
import numpy as np

import nim_amd
import oracle as O
import refdata as R

GEO_FIELDS = ("volume", "xi_x", "gij", "gij_faces", "G1", "G2", "G3", "faces_area", "faces_normal", "faces_mid",
              "bc_points", "bc_a")


def ref_geometry_for_gpu(geo):
    return {n: geo["bc_points"] if n == "bc_points" else geo[n] for n in GEO_FIELDS}


def state_from_steps(st, pre):
    m = {"faces_" + f: st[pre + "_faces_" + f] for f in "uvwpT"}
    for f in "uvwT":
        m[f + "xyz"] = st[pre + "_%sxyz" % f]
        m[f + "xyz_old"] = st[pre + "_%sxyz_old" % f]
        m[f + "xyzt"] = st[pre + "_%sxyzt" % f]
    m["meshes_p"] = st[pre + "_p"]
    m["meshes_dila"] = st[pre + "_dila"]
    return m


def gpu_context(geo, params=None):
    ctx = nim_amd.Context(*R.REF_DIMS, params or O.reference_params())
    ctx.upload_geometry(ref_geometry_for_gpu(geo))
    ctx.build_pressure_coeffs()
    return ctx


def oracle_context(geo, params=None):
    o = O.Oracle(*R.REF_DIMS, params or O.reference_params())
    O.load_reference_geometry(o, geo)
    return o


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def rel_max(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)
