"""Fortran host setup (libnim_host.so) against the compiled reference.

On the reference's own case the restated mesh generator, metrics and
boundary tables must reproduce the reference's arrays bit for bit. For other
cases the mathematics pins them: the O-grid without the hexagon deformation
is an annulus whose cell volumes integrate exactly with the 5-point rule.
"""
import numpy as np
import pytest

import refdata as R
from nim_amd import host


@pytest.fixture(scope="module")
def built():
    return host.build_geometry(host.reference_case())


INT_KEYS = ("meshes_faces", "meshes_neighbors", "bc_faces", "bc_position", "bc_mesh", "bc_points")
REAL_KEYS = ("node_xyz", "volume", "xi_x", "gij", "gij_faces", "G1", "G2", "G3", "faces_area", "faces_normal",
             "faces_tangential", "faces_bitangential", "faces_mid", "bc_a")


@pytest.mark.parametrize("key", INT_KEYS + REAL_KEYS)
def test_reference_case_bitwise(built, key):
    g = R.geometry()
    assert np.array_equal(built[key], g[key]), key


def test_undeformed_annulus_volume_exact():
    # no grading (exponent 1), blend radii beyond the domain: the cells are
    # annular sectors with straight-chord faces, whose trilinear volume the
    # Gauss rule integrates exactly; the total is the polygonal annulus volume
    nt, nr, nz = 24, 5, 4
    c = host.Case((nt, nr, nz), 0.5, 1.5, 2.0, Ra=1e4, Pr=0.7, DT=1e-3, blend=(0.0, 0.0), grading=1.0)
    g = host.build_geometry(c)
    poly = 0.5 * nt * np.sin(2 * np.pi / nt) * (1.5 ** 2 - 0.5 ** 2) * 2.0
    assert abs(g["volume"].sum() - poly) < 1e-12 * poly
    assert np.all(g["volume"] > 0)


def test_baseline_case_sizes_and_positive_volumes():
    c = host.baseline_case("C1")
    g = host.build_geometry(c)
    nt, nr, nz = c.dims
    assert g["volume"].size == nt * nr * nz
    assert np.all(g["volume"] > 0)
    # the deformed domain: hexagon of circumradius 0.3 inside the unit cylinder
    hexa = 1.5 * np.sqrt(3) * 0.3 ** 2
    cyl_poly = 0.5 * nt * np.sin(2 * np.pi / nt)
    np.testing.assert_allclose(g["volume"].sum(), (cyl_poly - hexa) * 1.0, rtol=2e-3)


def test_noise_matches_oracle_generator():
    import oracle as O
    L = host.lib()
    idx = np.arange(0, 5000, 7)
    got = np.array([L.nimh_noise(20221, 3, int(i)) for i in idx])
    assert np.array_equal(got, O.ic_noise(20221, 3, idx))


def test_initial_state_matches_reference_driver():
    st = R.steps()
    s = host.initial_state(host.reference_case(), amp=st["meta"][1], seed=int(st["meta"][2]))
    for f in "uvwpT":
        assert np.array_equal(s["faces_" + f], st["s0_faces_" + f]), f
    for f in "uvwT":
        assert np.array_equal(s[f + "xyz"], st["s0_%sxyz" % f]), f
