"""The 8.4M-cell case (256 x 512 x 64 cells, n_r x n_theta x n_z; Ra 1e6,
DT 2.5e-4), which only these tests run at full size: every index of the
device code is largest here.

  * 10 FP64 passes from the seeded state against the oracle in the same
    colour order: every field and the Nusselt numbers to rounding;
  * the mixed loop (4 FP32 passes per FP64 residual) against the FP64 loop
    over the same 256-pass budget: the FP32 corrections follow the FP64
    passes (measured 1e-9 - 2e-8) while the fields move by O(1);
  * the reversed stages of the discrete adjoint (momentum sources,
    momentum sweep, pressure sources, energy sweep) are the transposed
    Jacobians of the forward kernels: <lam, (S(x + e v) - S(x - e v)) / 2e>
    against <S'^T lam, v> with random lam on the stage's outputs and random
    v on its state inputs. The arrays are moved one at a time (a full set
    of a stage's arrays is up to 2 GB here)."""
import numpy as np
import pytest

import gpucase as G
import nim_amd
import oracle as O
from nim_amd import host

pytestmark = pytest.mark.gpu

FIELDS = ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]


@pytest.fixture(scope="module")
def c5():
    case = host.baseline_case("C5")
    assert case.ncells == 8388608
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    return case, geo, st


def test_c5_passes_match_oracle(c5):
    case, geo, st = c5
    ctx = G.gpu_for_case(case, geo, st)
    o = G.oracle_for_case(case, geo, st)
    try:
        ctx.build_momentum_coeffs()
        o.velocity_constants()
        n = 10
        rc, it, _ = ctx.inner_solve(np.zeros(5), max_iter=n, raise_on_cap=False)
        rc_o, it_o, _ = o.inner_solve(np.zeros(5), O.COLORED, max_iter=n)
        assert it == it_o == n
        errs = {f: G.rel_l2(ctx.get(f), o.get(f)) for f in FIELDS}
        print("C5, 10 passes:", " ".join("%s %.1e" % kv for kv in errs.items()))
        assert max(errs.values()) < 1e-9, errs
        np.testing.assert_allclose(ctx.nusselt()[:3], o.nusselt()[:3], rtol=1e-9)
    finally:
        ctx.close()
        o.close()


def test_c5_mixed_follows_fp64(c5):
    case, geo, st = c5
    ctx = G.gpu_for_case(case, geo, st)
    try:
        ctx.build_momentum_coeffs()
        ctx.inner_solve(np.zeros(5), max_iter=256, raise_on_cap=False)
        a = {n: ctx.get(n) for n in FIELDS}
        ctx.upload_state(st)
        ctx.build_momentum_coeffs()
        rc, cycles, _ = ctx.inner_solve_mixed(np.zeros(5), inner=4, max_cycles=65, raise_on_cap=False)
        assert cycles == 65  # 64 cycles of 4 FP32 passes, then the 65th residual
        for n in FIELDS:
            d = G.rel_l2(ctx.get(n), a[n])
            moved = G.rel_l2(a[n], st[n]) if np.linalg.norm(st[n]) > 0 else 1.0
            print("C5 mixed against FP64 after 256 passes: %s %.1e (moved %.1e)" % (n, d, moved))
            assert d < 1e-7 and moved > 1e-2, (n, d, moved)
    finally:
        ctx.close()


# stage: (state inputs perturbed, outputs weighted)
STAGES = {
    0: (["faces_u", "faces_v", "faces_w", "faces_p", "faces_T", "Txyz", "Txyz_old"], ["R_B_u", "R_B_v", "R_B_w"]),
    1: (["faces_u", "faces_v", "faces_w", "uxyz", "vxyz", "wxyz", "uxyz_old", "vxyz_old", "wxyz_old", "R_B_u",
         "R_B_v", "R_B_w"], ["faces_u", "faces_v", "faces_w", "uxyz", "vxyz", "wxyz"]),
    2: (["faces_u", "faces_v", "faces_w", "faces_p", "faces_T"], ["R_b_p"]),
    6: (["faces_T", "Txyz", "Txyz_old", "R_D_T"], ["faces_T", "Txyz"]),
}


def rand(name, size, seed):
    return np.random.default_rng([seed, sum(map(ord, name))]).standard_normal(size)


@pytest.mark.parametrize("stage", sorted(STAGES))
def test_c5_reversed_stage_is_the_transposed_jacobian(c5, stage):
    case, geo, st = c5
    ins, outs = STAGES[stage]
    ctx = G.gpu_for_case(case, geo, st)
    try:
        ctx.build_momentum_coeffs()
        # a state that is not at rest: a few passes from the seeded one, with
        # the energy coefficients stored and the sources at that state
        ctx.inner_solve(np.zeros(5), max_iter=3, raise_on_cap=False)
        ctx.run_stage(4)
        ctx.run_stage(0)
        ctx.run_stage(2)
        ctx.run_stage(5)
        x0 = {n: ctx.get(n) for n in ins}
        scale = {n: max(np.sqrt(np.mean(x0[n] ** 2)), 1e-3) for n in ins}
        eps = 1e-6
        acc = {}
        for sgn in (1.0, -1.0):
            for n in ins:
                ctx.set(n, x0[n] + sgn * eps * scale[n] * rand(n, x0[n].size, 1))
            ctx.run_stage(stage)
            acc[sgn] = sum(float(np.dot(rand(n, ctx._len(n), 2), ctx.get(n))) for n in outs)
        lhs = (acc[1.0] - acc[-1.0]) / (2 * eps)
        for n in ins:
            ctx.set(n, x0[n])
        adj = nim_amd.Adjoint(ctx)
        try:
            adj.clear()
            for n in outs:
                adj.set(n, rand(n, ctx._len(n), 2))
            adj.reverse_stage(stage)
            rhs = sum(float(np.dot(adj.get(n), scale[n] * rand(n, x0[n].size, 1))) for n in ins)
        finally:
            adj.close()
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        print("C5 stage %d: <lam, dS v> %.12e, <S'^T lam, v> %.12e, rel %.1e" % (stage, lhs, rhs, rel))
        assert rel <= 1e-7, (stage, lhs, rhs)
    finally:
        ctx.close()
