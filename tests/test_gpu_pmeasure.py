"""The pressure convergence measure NIM_P_HALFRANGE against the oracle's,
in the same colour order: half the range of the pressure increments, which
a change of p's constant level (p is determined only up to a constant) does
not move. The same stopping pass and fields to rounding, with the FP64 loop.
In the mixed loop the measure is the half range of each cycle's pressure
update per FP32 pass: the FP64 residual of the drifting rest state keeps a
range (the drift times the sweep's splitting applied to a constant), which
nim_residual still reports."""
import numpy as np
import pytest

import gpucase as G
import nim_amd
import oracle as O
import refdata as R

pytestmark = pytest.mark.gpu

TOL = np.full(5, 1e-8)


@pytest.fixture(scope="module")
def geo():
    return R.geometry()


@pytest.fixture(scope="module")
def steps():
    return R.steps()


def test_halfrange_solve_matches_oracle(geo, steps):
    ctx = G.gpu_context(geo)
    o = G.oracle_context(geo)
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    o.set_p_measure(1)
    ctx.upload_state(G.state_from_steps(steps, "s3"))
    R.load_state(o, steps, "s3")
    ctx.build_momentum_coeffs()
    o.velocity_constants()
    rc, it_g, err_g = ctx.inner_solve(TOL, max_iter=200000)
    rc_o, it_o, err_o = o.inner_solve(TOL, O.COLORED, max_iter=200000)
    print("passes gpu %d oracle %d; err gpu %s oracle %s" % (it_g, it_o, err_g, err_o))
    assert rc == 0 and rc_o == 0
    assert abs(it_g - it_o) <= 2, (it_g, it_o)
    assert err_g[3] <= 1e-8 and err_o[3] <= 1e-8
    for f in "uvwpT":
        assert G.rel_l2(ctx.get("faces_" + f), o.get("faces_" + f)) < 1e-9, f
    ctx.close()
    o.close()


def test_halfrange_measure_of_mixed_loop(geo, steps):
    """The mixed loop converges in the half-range measure (its pressure
    updates) to the FP64 loop's state, within the distance that the two
    stopping tests leave; nim_residual's half range of w r / a at that state
    stays at least the drift's floor, and the other fields' residual
    measures do not depend on the choice."""
    ctx = G.gpu_context(geo)
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    s3 = G.state_from_steps(steps, "s3")
    ctx.upload_state(s3)
    ctx.build_momentum_coeffs()
    rc, it, err = ctx.inner_solve(TOL, max_iter=200000)
    assert rc == 0
    ref = {f: ctx.get("faces_" + f) for f in "uvwpT"}
    ctx.upload_state(s3)
    ctx.build_momentum_coeffs()
    rc, cycles, err = ctx.inner_solve_mixed(TOL, inner=8, max_cycles=50000)
    assert rc == 0 and err[3] <= 1e-8
    r_half = ctx.residual()
    ctx.set_p_measure(nim_amd.P_MAXABS)
    r_abs = ctx.residual()
    gaps = {}
    for f in "uvwpT":
        a, b = ctx.get("faces_" + f), ref[f]
        if f == "p":
            a, b = a - a.mean(), b - b.mean()
        gaps[f] = G.rel_l2(a, b)
    print("FP64 %d passes, mixed %d cycles; residual half range %s, max abs %s; against FP64 %s"
          % (it, cycles, r_half, r_abs, gaps))
    assert r_abs[3] >= r_half[3]
    np.testing.assert_array_equal(np.delete(r_half, 3), np.delete(r_abs, 3))
    assert gaps["p"] < 1e-3 and max(gaps[f] for f in "uvw") < 1e-4 and gaps["T"] < 1e-7, gaps
    ctx.close()
