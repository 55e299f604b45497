"""The CPU oracle against the compiled reference's golden vectors.

The fixtures come from the reference's own modules (tests/golden/make_golden.py)
on its built-in 36 x 6 x 12 case. The oracle follows the Fortran expression
order, so every stage is required to agree to the last bit.
"""
import numpy as np
import pytest

import oracle as O
import refdata as R


@pytest.fixture(scope="module")
def geo():
    return R.geometry()


def fresh(geo):
    o = O.Oracle(*R.REF_DIMS, O.reference_params())
    O.load_reference_geometry(o, geo)
    return o


def test_pressure_constants_bitwise(geo):
    o = fresh(geo)
    for n in ("AA_px", "AA_py", "AA_pz", "bc_ap"):
        o.set(n, np.zeros_like(geo[n]))
    o.pressure_constants()  # nim.f90:556-743
    for n in ("AA_px", "AA_py", "AA_pz", "bc_ap"):
        assert np.array_equal(o.get(n), geo[n]), n


def test_one_iteration_stage_by_stage_bitwise(geo):
    st, it = R.steps(), R.iteration()
    o = fresh(geo)
    R.load_state(o, st, "s3")
    for f in "uvwpT":
        o.set("faces_%s_old" % f, it["it0_faces_%s_old" % f])
    for f in "uvwT":
        o.set("%sxyz_old_con" % f, it["it0_%sxyz_old_con" % f])

    def eq(name, key):
        assert np.array_equal(o.get(name), it[key]), (name, key)

    o.velocity_constants()
    for k in ("M_F1", "M_F2", "M_F3", "M_F4", "meshes_u", "meshes_v", "meshes_w"):
        eq(k, "it_" + k)
    o.velocity_sources()
    for k in ("R_B_u", "R_B_v", "R_B_w", "R_C_u", "R_C_v", "R_C_w", "R_D_u", "R_D_v", "R_D_w"):
        eq(k, "it_" + k)
    o.velocity_sweep(O.NATURAL)
    for f in "uvw":
        eq("faces_" + f, "it_vel_faces_" + f)
        eq(f + "xyz", "it_vel_%sxyz" % f)
    o.pressure_sources()
    eq("R_B_p", "it_R_b_p")
    eq("R_D_p", "it_R_D_p")
    eq("dila", "it_dila")
    o.pressure_bc()
    eq("faces_p", "it_pbc_faces_p")
    o.pressure_sweep()
    eq("faces_p", "it_p_faces_p")
    o.temperature_constants()
    for k in ("MT_F1", "MT_F2", "MT_F3", "MT_F4"):
        eq(k, "it_" + k)
    o.zero_flux()
    eq("faces_T", "it_zf_faces_T")
    o.temperature_sources()
    eq("R_D_T", "it_R_D_T")
    o.temperature_sweep(O.NATURAL)
    eq("faces_T", "it_T_faces_T")
    eq("Txyz", "it_T_Txyz")
    e = o.find_error()
    assert np.array_equal(e, it["it_err"][:5])


def test_three_step_trajectory_bitwise(geo):
    st = R.steps()
    o = fresh(geo)
    R.load_state(o, st, "s0")
    printed = st["printed_steps"]
    printed_nu = st["printed_nu"]
    for k in range(1, 4):
        rc, scal, nu = o.time_step(R.REF_TOL, O.NATURAL)
        assert rc == 0
        assert int(scal[0]) == int(printed[k - 1][1])  # "# iters." nse.f90:169
        # printed with 5 significant digits (nse.f90:2792-2796)
        np.testing.assert_allclose(nu, printed_nu[k - 1], rtol=6e-5)
        if k in (1, 3):
            pre = "s%d" % k
            for f in "uvwpT":
                assert np.array_equal(o.get("faces_" + f), st[pre + "_faces_" + f]), (k, f)
            for f in "uvwT":
                for suf in ("xyz", "xyz_old", "xyzt"):
                    assert np.array_equal(o.get(f + suf), st["%s_%s%s" % (pre, f, suf)]), (k, f, suf)
            assert np.array_equal(o.get("meshes_p"), st[pre + "_p"])
            np.testing.assert_array_equal(scal[6:9], st[pre + "_scalars"][:3])


def test_noise_generator_matches_fixture():
    st = R.steps()
    n = np.prod(R.REF_DIMS)
    amp = st["meta"][1]
    seed = int(st["meta"][2])
    got = amp * O.ic_noise(seed, 3, np.arange(n))
    assert np.array_equal(got, st["s0_Txyz"])
