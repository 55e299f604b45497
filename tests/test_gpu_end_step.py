"""The end-of-step reductions against the oracle: check_steady_state
(nse.f90:1316-1332: max |phi - phi_old| / (2 DT) of u, v, w and of T) and
evaluate_D (nse.f90:2504-2537: the dilatation of every cell and its RMS),
which integrate_nse runs after its loop. Both sides run the same passes in
the colour order from the same state, so the fields agree to rounding; the
reductions are then compared on the reference case (its own geometry from
the golden vectors, state after step 3) and at C1 (48 x 24 x 16, the
seeded initial state), through nim_end_step."""
import numpy as np
import pytest

import gpucase as G
import oracle as O
import refdata as R
from nim_amd import host

pytestmark = pytest.mark.gpu


def _compare(ctx, o, passes):
    ctx.build_momentum_coeffs()
    o.velocity_constants()
    zero = np.zeros(5)
    rc, it, _ = ctx.inner_solve(zero, max_iter=passes, raise_on_cap=False)
    rc_o, it_o, _ = o.inner_solve(zero, O.COLORED, max_iter=passes)
    assert it == it_o == passes
    got = ctx.end_step()
    o.check_steady_state()
    o.evaluate_D()
    want = o.scalars()
    print("end step gpu", got, "oracle", want)
    for k in ("dudt_max", "dTdt_max", "Dilatation"):
        assert want[k] > 0
        assert abs(got[k] - want[k]) <= 1e-9 * abs(want[k]), (k, got[k], want[k])
    assert G.rel_l2(ctx.get("meshes_dila"), o.get("dila")) < 1e-9


def test_end_step_reference_case():
    geo = R.geometry()
    ctx = G.gpu_context(geo)
    o = G.oracle_context(geo)
    steps = R.steps()
    ctx.upload_state(G.state_from_steps(steps, "s3"))
    R.load_state(o, steps, "s3")
    _compare(ctx, o, 300)
    ctx.close()
    o.close()


def test_end_step_c1():
    case = host.baseline_case("C1")
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    ctx = G.gpu_for_case(case, geo, st)
    o = G.oracle_for_case(case, geo, st)
    _compare(ctx, o, 200)
    ctx.close()
    o.close()
