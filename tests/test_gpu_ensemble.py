"""Ensembles: independent cases of different mesh sizes and parameters
solved together, each kernel of a pass launched once over all cases. Every
case must get exactly (bitwise) the result and pass count of its own solve,
in the FP64 loop and in the mixed loop."""
import numpy as np
import pytest

import nim_amd
from nim_amd import host

pytestmark = pytest.mark.gpu

CASES = [
    # (n_theta, n_r, n_z), Ra, Pr, hexagon circumradius
    ((24, 8, 6), 1e4, 0.71, 0.30),
    ((32, 6, 8), 1e5, 7.0, 0.25),
    ((20, 10, 5), 3e3, 0.71, 0.35),
]
NAMES = ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]
TOL = np.array([1e-9, 1e-9, 1e-9, 1e-5, 1e-9])


@pytest.fixture(scope="module")
def setups():
    out = []
    for q, (dims, Ra, Pr, Ri) in enumerate(CASES):
        case = host.Case(dims, Ri, 1.0, 1.0, Ra, Pr, 1e-3)
        out.append((case, host.build_geometry(case), host.initial_state(case, amp=1e-3, seed=100 + q)))
    return out


def contexts(setups):
    cs = []
    for case, geo, st in setups:
        c = nim_amd.Context(*case.dims, case)
        c.upload_geometry(geo)
        c.build_pressure_coeffs()
        c.upload_state(st)
        cs.append(c)
    return cs


@pytest.mark.parametrize("inner", [0, 4])
def test_ensemble_matches_individual_solves(setups, inner):
    alone = contexts(setups)
    it_alone = []
    for c in alone:
        its = []
        for step in range(2):
            c.build_momentum_coeffs()
            if inner:
                rc, n, _ = c.inner_solve_mixed(TOL, inner=inner)
            else:
                rc, n, _ = c.inner_solve(TOL, min_iter=2 if step == 0 else 0)
            its.append(n)
            c.end_step()
            c.advance_time()
        it_alone.append(its)
    together = contexts(setups)
    ens = nim_amd.Ensemble(together)
    it_ens = [[] for _ in together]
    for step in range(2):
        for c in together:
            c.build_momentum_coeffs()
        rc, its, err = ens.inner_solve(TOL, inner=inner, min_iter=2 if (step == 0 and not inner) else 0)
        assert np.all(err <= TOL)
        for q, c in enumerate(together):
            it_ens[q].append(int(its[q]))
            c.end_step()
            c.advance_time()
    assert it_ens == it_alone
    for a, b in zip(alone, together):
        for n in NAMES:
            assert np.array_equal(a.get(n), b.get(n)), n
    ens.close()
    for c in alone + together:
        c.close()


def test_c3_cases_whole_steps_match_solo_steps():
    """Six cases drawn as the specification's 48-case ensemble draws them
    (host.ensemble_cases: Ra log-uniform in [1e3, 1e6], Pr in {0.71, 7},
    hexagon circumradius in [0.2, 0.4], n_r in {16, 24, 32, 48}, n_theta =
    2 n_r, n_z in {12, 16, 24}, seeded), two whole time steps through
    nim_ens_time_step (the per-step kernels batched over the cases and
    graph-captured around the solve) in the FP64 loop and in the mixed loop
    (8 FP32 passes per residual): every case's pass counts, end-of-step
    scalars, Nusselt numbers and fields are bitwise those of the same steps
    on its own context."""
    drawn = host.ensemble_cases()[:6]
    dims = {case.dims for case, _ in drawn}
    assert len(dims) >= 4
    setups = [(case, host.build_geometry(case, device=True), host.initial_state(case, amp=1e-3, seed=sd))
              for case, sd in drawn]
    tol = np.array([1e-8, 1e-8, 1e-8, 5e-5, 1e-8])
    for inner in (0, 8):
        print("C3 draw", [(c.dims, round(c.Ra), c.Pr) for c, _ in drawn], "inner", inner, flush=True)
        alone = contexts(setups)
        ref = []
        for c in alone:
            rows = []
            for step in range(2):
                c.build_momentum_coeffs()
                if inner:
                    rc, n, _ = c.inner_solve_mixed(tol, inner=inner)
                else:
                    rc, n, _ = c.inner_solve(tol, min_iter=2 if step == 0 else 0)
                scal = c.end_step()
                c.advance_time()
                rows.append((n, scal, c.nusselt()))
            ref.append(rows)
        together = contexts(setups)
        ens = nim_amd.Ensemble(together)
        got = [[] for _ in together]
        for step in range(2):
            rc, its, err, res = ens.time_step_batched(tol, inner=inner, min_iter=2 if (step == 0 and not inner) else 0)
            for q in range(len(together)):
                got[q].append((int(its[q]), res[q][0], res[q][1]))
        for q, (a, b) in enumerate(zip(alone, together)):
            for (n0, s0, nu0), (n1, s1, nu1) in zip(ref[q], got[q]):
                assert n0 == n1, (q, n0, n1)
                assert s0 == s1, (q, s0, s1)
                assert np.array_equal(nu0, nu1), (q, nu0, nu1)
            for n in NAMES + ["uxyz_old", "Txyz_old", "uxyzt", "Txyzt", "meshes_p", "meshes_dila"]:
                assert np.array_equal(a.get(n), b.get(n)), (q, n)
        ens.close()
        for c in alone + together:
            c.close()
