"""Loaders for the reference golden fixtures in tests/golden/."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_DIMS = (36, 6, 12)  # parameters.f90:87-89 (theta, r, z)
REF_TOL = np.array([1e-6, 1e-6, 1e-6, 5e-3, 1e-6])  # parameters.f90:146-150


def geometry():
    return np.load(os.path.join(GOLDEN, "ref_geometry.npz"))


def steps():
    return np.load(os.path.join(GOLDEN, "ref_steps.npz"))


def iteration():
    return np.load(os.path.join(GOLDEN, "ref_iteration.npz"))


def load_state(o, st, pre):
    """Copy a dumped reference state (prefix s0/s1/s3) into an object with a
    set(name, values) method using the oracle's field names."""
    for f in "uvwpT":
        o.set("faces_" + f, st[pre + "_faces_" + f])
    for f in "uvwT":
        o.set(f + "xyz", st[pre + "_%sxyz" % f])
        o.set(f + "xyz_old", st[pre + "_%sxyz_old" % f])
        o.set(f + "xyzt", st[pre + "_%sxyzt" % f])
    o.set("meshes_p", st[pre + "_p"])
    o.set("dila", st[pre + "_dila"])
