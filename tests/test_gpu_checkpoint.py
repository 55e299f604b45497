"""Reversing a run of time steps with device checkpoints and the binomial
plan: every state before steps nsteps-1 .. 0 is visited, recomputed from at
most `nslots` snapshots, and equals bit for bit the state the forward run
had there (the time step is deterministic on the device)."""
import numpy as np
import pytest

import gpucase as G
import nim_amd
import refdata as R

pytestmark = pytest.mark.gpu

TOL = np.array([1e-8, 1e-8, 1e-8, 1e-5, 1e-8])


@pytest.mark.parametrize("inner", [0, 4])
def test_reverse_run_visits_forward_states(inner):
    geo, st = R.geometry(), R.steps()
    nsteps, nslots = 10, 3

    def make():
        ctx = G.gpu_context(geo)
        ctx.upload_state(G.state_from_steps(st, "s0"))
        return ctx

    def step(ctx):
        ctx.build_momentum_coeffs()
        if inner:
            ctx.inner_solve_mixed(TOL, inner=inner)
        else:
            ctx.inner_solve(TOL)
        ctx.end_step()
        ctx.advance_time()

    ref = make()
    states = []
    for _ in range(nsteps):
        states.append(ref.download_state())
        step(ref)
    ref.close()

    ctx = make()
    visited = []

    def reverse(k):
        s = ctx.download_state()
        for n, v in states[k].items():
            assert np.array_equal(s[n], v), (k, n)
        visited.append(k)

    fwd = nim_amd.reverse_run(ctx, nsteps, nslots, lambda: step(ctx), reverse)
    ctx.close()
    assert visited == list(range(nsteps - 1, -1, -1))
    plan, fwd_plan = nim_amd.revolve_plan(nsteps, nslots)
    assert fwd == fwd_plan < nsteps * (nsteps - 1) // 2
