"""The accelerated inner loop (nim_inner_solve_accel: Anderson extrapolation
every few passes or mixed cycles) reaches the plain iteration's fixed point.

* On the 12 x 4 x 4 mesh of tests/test_gpu_fixed_point.py the oracle's
  colour order comes to rest (increments below 1e-15) after 6,000 passes.
  The accelerated FP64 loop, the same with an FP32 history, and the
  accelerated mixed loop stop on a pass (cycle) meeting tol 1e-13 and must
  land on that rest state to 1e-10 in every field, p (modulo its constant)
  to 1e-9 (its slowest mode leaves it undetermined by about the tolerance
  over that mode's rate), in a fraction of the passes; the mixed loop, whose
  p measure is held to 1e-11 by the rounding of its FP32 updates, to 1e-8.
* At the 48 x 24 x 16 case, one step from the seeded state: the plain loop
  at 1e-10 and the accelerated loop at 1e-12 agree (the same fixed point),
  as do the accelerated loops of both scopes at 1e-11 at the 1M-cell case,
  and at 1e-8 on every field (p as half the range of its increments) the
  plain FP64 loop, the accelerated FP64 loop and the accelerated mixed loop
  end within the distance that the stopping test leaves from that fixed
  point, the accelerated loops in a small fraction of the plain passes.
* The cap and an already converged state behave as in the plain loop.

Each holds for both scopes of the extrapolation: the pressure faces (the
default) and the whole iterate."""
import numpy as np
import pytest

import gpucase as G
import nim_amd
import oracle as O
from nim_amd import host

pytestmark = pytest.mark.gpu

FIELDS = ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]


def gaps(get_a, get_b):
    out = {}
    for f in FIELDS:
        a, b = get_a(f), get_b(f)
        if f == "faces_p":
            a, b = a - a.mean(), b - b.mean()
        out[f] = G.rel_l2(a, b)
    return out


@pytest.fixture(scope="module")
def rest():
    """The oracle's colour-order rest state on 12 x 4 x 4 (Taylor branch of
    the fits everywhere, so that the increments vanish to rounding)."""
    case = host.Case((12, 4, 4), 0.3, 1.0, 1.0, Ra=1e4, Pr=0.71, DT=1e-3, v_limit=0.3)
    geo = host.build_geometry(case, device=False)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    o = G.oracle_for_case(case, geo, st)
    o.set_p_measure(1)
    o.velocity_constants()
    rc, it, err = o.inner_solve(np.zeros(5), O.COLORED, max_iter=6000)
    assert it == 6000 and max(err[:3]) < 1e-15 and err[3] < 1e-13, err
    out = {f: o.get(f) for f in FIELDS}
    o.close()
    return case, geo, st, out


@pytest.mark.parametrize("form", ["fp64", "fp64_fp32_history", "mixed", "fp64_iterate", "mixed_iterate"])
def test_accelerated_solve_reaches_the_rest_state(rest, form):
    case, geo, st, ref = rest
    ctx = G.gpu_for_case(case, geo, st)
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    ctx.accel_options(depth=10, period=1 if form.startswith("mixed") else 4,
                      basis_bytes=4 if "fp32" in form else 8)
    ctx.accel_scope(nim_amd.ACCEL_ITERATE if form.endswith("iterate") else nim_amd.ACCEL_PRESSURE)
    mixed = form.startswith("mixed")
    ctx.build_momentum_coeffs()
    tol = np.full(5, 1e-13)
    if mixed:
        # the mixed loop measures p by its updates, FP32 corrections of which
        # the pressure level's drift (4 x 3e-6 per cycle here) is most: their
        # rounding alone is about 2e-13 per pass
        tol[3] = 1e-11
    rc, n, err = ctx.inner_solve_accel(tol, inner=4 if mixed else 0, max_iter=6000)
    assert rc == 0 and np.all(err <= tol), err
    passes = n * 4 if mixed else n
    g = gaps(ctx.get, ref.__getitem__)
    print(form, passes, "passes", g)
    assert passes < 2000, passes
    for f, v in g.items():
        assert v < (1e-8 if mixed else 1e-9 if f == "faces_p" else 1e-10), (f, v)
    ctx.close()


@pytest.fixture(scope="module")
def c1_steps():
    """One C1 step from the seeded state by the plain FP64 loop and the
    accelerated loops, to 1e-8 on every field (p as half the range of its
    increments, in the mixed loop of its updates per FP32 pass), and the
    references: the accelerated FP64 loop to 1e-12 and the plain loop to
    1e-10."""
    case = host.baseline_case("C1")
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    out = {}
    for form, tol in (("ref", 1e-12), ("plain_tight", 1e-10), ("plain", 1e-8), ("fp64", 1e-8), ("mixed", 1e-8),
                      ("mixed_budget", 0.), ("fp64_iterate", 1e-8)):
        ctx = G.gpu_for_case(case, geo, st)
        ctx.set_p_measure(nim_amd.P_HALFRANGE)
        ctx.build_momentum_coeffs()
        t = np.full(5, tol)
        if form.startswith("plain"):
            rc, n, err = ctx.inner_solve(t, max_iter=200000, min_iter=2)
        elif form == "mixed":
            ctx.accel_options(depth=10, period=1, basis_bytes=4)
            rc, n, err = ctx.inner_solve_accel(t, inner=16, max_iter=20000)
            n *= 16
        elif form == "mixed_budget":
            # no stopping test: a fixed budget of cycles
            ctx.accel_options(depth=10, period=1, basis_bytes=4)
            rc, n, err = ctx.inner_solve_accel(t, inner=16, max_iter=300, raise_on_cap=False)
            assert rc == nim_amd.NIM_E_MAXITER
            rc, n = 0, n * 16
        elif form == "fp64_iterate":
            ctx.accel_options(depth=16, period=6, basis_bytes=4)
            ctx.accel_scope(nim_amd.ACCEL_ITERATE)
            rc, n, err = ctx.inner_solve_accel(t, max_iter=200000, min_iter=2)
        else:
            ctx.accel_options(depth=12, period=1, basis_bytes=4)
            rc, n, err = ctx.inner_solve_accel(t, max_iter=200000, min_iter=2)
        assert rc == 0 and (form == "mixed_budget" or np.all(err <= t)), (form, err)
        if form == "ref":
            # the FP64 residual's half range for p stays at the pressure
            # drift's floor even here
            print("FP64 residual measure at the reference state", ctx.residual())
        out[form] = (n, {f: ctx.get(f) for f in FIELDS}, ctx.nusselt()[:3])
        ctx.close()
    return out


def test_c1_accelerated_loops_reach_the_plain_fixed_point(c1_steps):
    """The plain loop at 1e-10 and the accelerated loop at 1e-12 land on the
    same state, to what the plain loop's slow pressure mode leaves open at
    1e-10 (its increments fall by about 1e-4 per pass)."""
    n_ref, ref, nu_ref = c1_steps["ref"]
    n_pt, pt, nu_pt = c1_steps["plain_tight"]
    g = gaps(pt.__getitem__, ref.__getitem__)
    print("plain 1e-10 (%d passes) against accelerated 1e-12 (%d passes):" % (n_pt, n_ref), g)
    assert n_ref < n_pt / 10
    assert g["faces_p"] < 1e-4, g
    for f in ("faces_u", "faces_v", "faces_w", "uxyz", "vxyz", "wxyz"):
        assert g[f] < 5e-6, (f, g[f])
    for f in ("faces_T", "Txyz"):
        assert g[f] < 1e-9, (f, g[f])
    np.testing.assert_allclose(nu_pt, nu_ref, rtol=1e-9)


def test_c1_mixed_loop_reaches_the_fixed_point(c1_steps):
    """300 accelerated mixed cycles of 16 FP32 passes, no stopping test: the
    FP32 corrections limit only the rate, the fixed point is the FP64 one."""
    n_ref, ref, nu_ref = c1_steps["ref"]
    n, fields, nu = c1_steps["mixed_budget"]
    g = gaps(fields.__getitem__, ref.__getitem__)
    print("mixed, %d FP32 passes, against the fixed point:" % n, g)
    assert g["faces_p"] < 1e-4, g
    for f in ("faces_u", "faces_v", "faces_w", "uxyz", "vxyz", "wxyz"):
        assert g[f] < 5e-6, (f, g[f])
    for f in ("faces_T", "Txyz"):
        assert g[f] < 1e-9, (f, g[f])
    np.testing.assert_allclose(nu, nu_ref, rtol=1e-9)


def test_c1_strict_step_accelerated_loops(c1_steps):
    """At 1e-8 every loop stops within the distance that its stopping test
    leaves from the fixed point; the accelerated loops take a small fraction
    of the plain passes. Measured on the device: the plain loop's p is 3.5e-3
    from the fixed point (its slowest mode leaves it undetermined by up to
    the tolerance over that mode's rate), the accelerated FP64 loop's 7e-4."""
    n_ref, ref, nu_ref = c1_steps["ref"]
    n_plain = c1_steps["plain"][0]
    for form in ("plain", "fp64", "mixed", "fp64_iterate"):
        n, fields, nu = c1_steps[form]
        g = gaps(fields.__getitem__, ref.__getitem__)
        print(form, n, "passes, against the fixed point:", g)
        if form != "plain":
            assert n < n_plain / 8, (form, n, n_plain)
        assert g["faces_p"] < 1e-2, (form, g)
        for f in ("faces_u", "faces_v", "faces_w", "uxyz", "vxyz", "wxyz"):
            assert g[f] < 5e-4, (form, f, g[f])
        for f in ("faces_T", "Txyz"):
            assert g[f] < 1e-8, (form, f, g[f])
        np.testing.assert_allclose(nu, nu_ref, rtol=1e-6)


def test_cap_and_converged_entry(rest):
    case, geo, st, _ = rest
    ctx = G.gpu_for_case(case, geo, st)
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    ctx.accel_options(depth=4, period=4)
    ctx.build_momentum_coeffs()
    tol = np.full(5, 1e-13)
    rc, n, _ = ctx.inner_solve_accel(tol, max_iter=10, raise_on_cap=False)
    assert rc == nim_amd.NIM_E_MAXITER and n == 10
    rc, n, err = ctx.inner_solve_accel(tol, max_iter=6000)
    assert rc == 0
    # from a converged state the first pass meets the test
    rc, n, err = ctx.inner_solve_accel(tol, max_iter=6000)
    assert rc == 0 and n == 1, (n, err)
    with pytest.raises(nim_amd.NimError):
        ctx.accel_options(depth=17)
    with pytest.raises(nim_amd.NimError):
        ctx.accel_scope(2)
    ctx.close()


def test_c2_accelerated_step_both_loops():
    """At the benchmark's 1M-cell case, one step from the seeded state: the
    accelerated FP64 loop at 1e-11 over the pressure faces and over the whole
    iterate reach the same fixed point (to what their stopping test leaves:
    2.8e-5 in p, 1.2e-6 in v measured); at 1e-8 on every field the FP64 and
    mixed loops of both scopes converge in a few thousand passes (the plain
    loop needs 1.4e5 for p) and end within the distance that the stopping
    test leaves from that point (p is determined to about 1e-8 over the slow
    mode's rate, 1.5e-5 per pass: 1.0e-2 - 1.6e-2 here, v 4.7e-4 - 6.8e-4,
    T 3.4e-10 - 4.6e-10 measured); T and the Nusselt numbers, which the
    pressure mode hardly reaches, agree tightly."""
    case = host.baseline_case("C2")
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    ctx = G.gpu_for_case(case, geo, st)
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    P, I = nim_amd.ACCEL_PRESSURE, nim_amd.ACCEL_ITERATE
    out = {}
    for form, tol, inner, depth, period, scope in (("ref", 1e-11, 0, 12, 1, P), ("ref_iterate", 1e-11, 0, 16, 6, I),
                                                    ("fp64", 1e-8, 0, 12, 1, P), ("mixed", 1e-8, 4, 10, 1, P),
                                                    ("fp64_iterate", 1e-8, 0, 16, 6, I),
                                                    ("mixed_iterate", 1e-8, 16, 10, 1, I)):
        ctx.upload_state(st)
        ctx.accel_options(depth=depth, period=period)
        ctx.accel_scope(scope)
        ctx.build_momentum_coeffs()
        t = np.full(5, tol)
        rc, n, err = ctx.inner_solve_accel(t, inner=inner, max_iter=40000 // max(inner, 1),
                                           min_iter=2 if not inner else 0)
        passes = (n - 1) * inner if inner else n
        assert rc == 0 and np.all(err <= t), (form, err)
        assert passes < 10000, (form, passes)
        out[form] = ({f: ctx.get(f) for f in FIELDS}, ctx.nusselt()[:3], passes)
    ctx.close()
    ref = out["ref"][0]
    g = gaps(out["ref_iterate"][0].__getitem__, ref.__getitem__)
    print("references (%d, %d passes) against each other:" % (out["ref"][2], out["ref_iterate"][2]), g)
    # each stops within about 1e-3 of the 1e-8 slack below (p 1.5e-5, v 7e-7),
    # possibly on opposite sides
    assert g["faces_p"] < 1e-4 and max(g[f] for f in FIELDS if f != "faces_p") < 5e-6, g
    for form in ("fp64", "mixed", "fp64_iterate", "mixed_iterate"):
        g = gaps(out[form][0].__getitem__, ref.__getitem__)
        print(form, out[form][2], "passes, against the fixed point:", g)
        assert g["faces_p"] < 3e-2, (form, g)
        for f in ("faces_u", "faces_v", "faces_w", "uxyz", "vxyz", "wxyz"):
            assert g[f] < 1.5e-3, (form, f, g[f])
        for f in ("faces_T", "Txyz"):
            assert g[f] < 1e-8, (form, f, g[f])
        np.testing.assert_allclose(out[form][1], out["ref"][1], rtol=1e-6)


def test_ragged_iterate_length():
    """10 x 4 x 5 cells: the iterate's length 4 N + 5 NF is 2 mod 4 (a
    partial group of four at its end), the pressure faces start inside a
    group of the whole iterate and their count NF is 2 mod 4; the
    accelerated loop over either reaches the plain loop's rest state there."""
    case = host.Case((10, 4, 5), 0.3, 1.0, 1.0, Ra=1e4, Pr=0.71, DT=1e-3, v_limit=0.3)
    N, NF, _ = nim_amd.counts(*case.dims)
    assert (4 * N + 5 * NF) % 4 == 2 and (4 * N + 3 * NF) % 4 == 2 and NF % 4 == 2
    geo = host.build_geometry(case, device=False)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    out = {}
    for form in ("plain", "accel", "accel_iterate"):
        ctx = G.gpu_for_case(case, geo, st)
        ctx.set_p_measure(nim_amd.P_HALFRANGE)
        ctx.build_momentum_coeffs()
        if form == "plain":
            rc, n, err = ctx.inner_solve(np.zeros(5), max_iter=6000, raise_on_cap=False)
            assert n == 6000 and max(err[:3]) < 1e-15 and err[3] < 1e-13, err
        else:
            ctx.accel_options(depth=10, period=4)
            ctx.accel_scope(nim_amd.ACCEL_ITERATE if form == "accel_iterate" else nim_amd.ACCEL_PRESSURE)
            rc, n, err = ctx.inner_solve_accel(np.full(5, 1e-13), max_iter=6000)
            assert rc == 0 and n < 2000, (n, err)
        out[form] = {f: ctx.get(f) for f in FIELDS}
        ctx.close()
    for form in ("accel", "accel_iterate"):
        g = gaps(out[form].__getitem__, out["plain"].__getitem__)
        print(form, "against 6,000 plain passes:", g)
        for f, v in g.items():
            assert v < (1e-9 if f == "faces_p" else 1e-10), (form, f, v)


def test_high_rayleigh_ensemble_case():
    """The highest-Ra case of the 48-case ensemble draw (32 x 16 x 16, Ra
    7.5e5): the accelerated and plain loops at 1e-10 reach the same state
    from the case's seeded start, the accelerated loop in a fraction of the
    passes, with the mixed accelerated loop at 1e-8 within its stopping
    slack."""
    case, seed = max(host.ensemble_cases(), key=lambda t: t[0].Ra)
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, seed=seed)
    out = {}
    for form, tol in (("plain", 1e-10), ("fp64", 1e-10), ("mixed", 1e-8)):
        ctx = G.gpu_for_case(case, geo, st)
        ctx.set_p_measure(nim_amd.P_HALFRANGE)
        ctx.build_momentum_coeffs()
        t = np.full(5, tol)
        if form == "plain":
            rc, n, err = ctx.inner_solve(t, max_iter=400000, min_iter=2)
        elif form == "fp64":
            rc, n, err = ctx.inner_solve_accel(t, max_iter=100000, min_iter=2, raise_on_cap=False)
        else:
            ctx.accel_options(depth=10, period=1)
            rc, n, err = ctx.inner_solve_accel(t, inner=16, max_iter=10000, raise_on_cap=False)
            n = (n - 1) * 16
        print(form, rc, n, err)
        assert rc == 0 and np.all(err <= t), (form, err)
        out[form] = (n, {f: ctx.get(f) for f in FIELDS})
        ctx.close()
    g = gaps(out["fp64"][1].__getitem__, out["plain"][1].__getitem__)
    gm = gaps(out["mixed"][1].__getitem__, out["fp64"][1].__getitem__)
    print("passes", {k: v[0] for k, v in out.items()}, "fp64 against plain", g, "mixed against fp64", gm)
    assert out["fp64"][0] < out["plain"][0] / 4
    assert g["faces_p"] < 1e-4 and max(g[f] for f in FIELDS if f != "faces_p") < 1e-5, g
    assert gm["faces_p"] < 1e-1 and max(gm[f] for f in FIELDS if f != "faces_p") < 1e-3, gm
