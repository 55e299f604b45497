"""Shared set-up of GPU contexts and oracle contexts on the reference case
and on run-time sized cases."""
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


def faces_bc(case):
    """faces_BC of the reference numbering (create_geometry.f90:626-664):
    interior faces 1, then bottom 2, top 3, hexagon 4, cylinder 5."""
    nt, nr, nz = case.dims
    N = nt * nr * nz
    NF = 3 * N + nt * nz + nt * nr
    bc = np.ones(NF, dtype=np.int32)
    o3 = NF - 2 * nt * nr - 2 * nt * nz
    bc[o3:o3 + nt * nr] = 2
    bc[o3 + nt * nr:o3 + 2 * nt * nr] = 3
    bc[o3 + 2 * nt * nr:o3 + 2 * nt * nr + nt * nz] = 4
    bc[o3 + 2 * nt * nr + nt * nz:] = 5
    return bc


def oracle_for_case(case, geo, st):
    """An oracle context of a host.Case with its geometry (host or device
    metrics, bitwise the same) and state loaded, pressure constants built."""
    p = O.Params()
    for n, _ in O.Params._fields_:
        setattr(p, n, getattr(case, n))
    o = O.Oracle(*case.dims, p)
    g = dict(geo)
    g["faces_BC"] = faces_bc(case)
    O.load_reference_geometry(o, g)
    o.pressure_constants()
    for k, v in st.items():
        o.set({"meshes_dila": "dila"}.get(k, k), v)
    return o


def gpu_for_case(case, geo, st):
    """A GPU context of a host.Case with its geometry and state uploaded."""
    ctx = nim_amd.Context(*case.dims, case)
    ctx.upload_geometry(geo)
    ctx.build_pressure_coeffs()
    ctx.upload_state(st)
    return ctx
