"""HIP path against the CPU oracle on the reference's 36 x 6 x 12 case.

Tolerances: the GPU contracts multiply-adds into FMA and evaluates a few
sums in another order, so per-stage agreement is at rounding level
(relative max error 1e-12 for single stages). Converged fields are held to
the 1e-8 relative L2 bar of the specification and Nusselt numbers to 1e-6.
"""
import numpy as np
import pytest

import gpucase as G
import oracle as O
import refdata as R

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-12
# The exponential-fit constants (nim.f90:184-287) cancel catastrophically just
# above |Re| = v_limit: -k + u*coth(Re/2) ~ k*Re^2/12 and 1 - exp(-Re) ~ Re,
# so one ulp of exp() (the device's exp is not glibc's) or one fused
# multiply-add moves them by up to ~1e7 ulp. Measured on the CPU alone, FMA
# contraction changes M_F1 by 6e-11 relative. Coefficients are therefore held
# to 1e-8 and the oracle's coefficients are copied into the device before the
# kernels that consume them, so every later stage is checked at rounding level.
COEF_TOL = 1e-8


@pytest.fixture(scope="module")
def geo():
    return R.geometry()


@pytest.fixture(scope="module")
def steps():
    return R.steps()


def test_pressure_coefficients_match_reference(geo):
    ctx = G.gpu_context(geo)
    for n in ("AA_px", "AA_py", "AA_pz", "bc_ap"):
        got = ctx.get(n)
        ref = geo[n]
        # slots the reference never assigns (walls without a neighbour) stay 0
        mask = ref != 0
        assert G.rel_max(got[mask], ref[mask]) < 1e-14, n
    ctx.close()


def test_stages_match_oracle_colored(geo, steps):
    ctx = G.gpu_context(geo)
    o = G.oracle_context(geo)
    st = G.state_from_steps(steps, "s3")
    ctx.upload_state(st)
    R.load_state(o, steps, "s3")

    ctx.build_momentum_coeffs()
    o.velocity_constants()
    for n in ("M_F1", "M_F2", "M_F3", "M_F4"):
        ref = o.get(n)
        got = ctx.get(n)
        mask = ref != 0
        assert G.rel_max(got[mask], ref[mask]) < COEF_TOL, n
        ctx.set(n, ref)

    ctx.run_stage("velocity_sources")
    o.velocity_sources()
    for f in "uvw":
        ref = o.get("R_C_" + f) + o.get("R_D_" + f) + o.get("R_B_" + f)
        assert G.rel_max(ctx.get("R_B_" + f), ref) < STAGE_TOL, f

    ctx.run_stage("velocity_sweep")
    o.velocity_sweep(O.COLORED)
    for f in "uvw":
        assert G.rel_max(ctx.get("faces_" + f), o.get("faces_" + f)) < STAGE_TOL, f
        assert G.rel_max(ctx.get(f + "xyz"), o.get(f + "xyz")) < STAGE_TOL, f

    ctx.run_stage("pressure_sources")
    o.pressure_sources()
    assert G.rel_max(ctx.get("R_b_p"), o.get("R_D_p") + o.get("R_B_p")) < STAGE_TOL
    assert G.rel_max(ctx.get("meshes_dila"), o.get("dila")) < STAGE_TOL

    ctx.run_stage("pressure_stage")
    o.pressure_stage(O.COLORED)
    assert G.rel_max(ctx.get("faces_p"), o.get("faces_p")) < STAGE_TOL

    ctx.run_stage("temperature_coeffs")
    o.temperature_constants()
    o.zero_flux()
    for n in ("MT_F1", "MT_F2", "MT_F3", "MT_F4"):
        ref = o.get(n)
        mask = ref != 0
        assert G.rel_max(ctx.get(n)[mask], ref[mask]) < COEF_TOL, n
        ctx.set(n, ref)
    assert G.rel_max(ctx.get("faces_T"), o.get("faces_T")) < STAGE_TOL

    ctx.run_stage("temperature_sources")
    o.temperature_sources()
    assert G.rel_max(ctx.get("R_D_T"), o.get("R_D_T")) < STAGE_TOL

    ctx.run_stage("temperature_sweep")
    o.temperature_sweep(O.COLORED)
    assert G.rel_max(ctx.get("faces_T"), o.get("faces_T")) < STAGE_TOL
    assert G.rel_max(ctx.get("Txyz"), o.get("Txyz")) < STAGE_TOL
    ctx.close()


# Convergence tolerances (max increment per pass for u, v, w, p, T). The
# pressure equation has only Neumann-type walls, so p is determined up to a
# constant whose mode relaxes at ~1 - 1.5e-5 per pass: its max increment
# stalls near 1e-8 even in the reference's own order (measured with the
# oracle), which is why the reference itself tests p 5000x more loosely than
# u (parameters.f90:146-150). Pressure is therefore compared modulo its mean.
TOL_COLORED = np.array([1e-9, 1e-9, 1e-9, 1e-6, 1e-9])
TOL_CONVERGED = np.array([1e-10, 1e-10, 1e-10, 1e-6, 1e-10])


@pytest.mark.parametrize("graph", ["1", "0"])
def test_inner_solve_matches_oracle_colored(geo, steps, graph, monkeypatch):
    """Same sweep order on both sides: the whole solve (~1800 passes) agrees
    to rounding."""
    monkeypatch.setenv("NIM_GRAPH", graph)
    ctx = G.gpu_context(geo)
    o = G.oracle_context(geo)
    ctx.upload_state(G.state_from_steps(steps, "s3"))
    R.load_state(o, steps, "s3")
    ctx.build_momentum_coeffs()
    o.velocity_constants()
    rc, it_g, err_g = ctx.inner_solve(TOL_COLORED)
    rc_o, it_o, err_o = o.inner_solve(TOL_COLORED, O.COLORED)
    assert rc == 0 and rc_o == 0
    assert abs(it_g - it_o) <= 2, (it_g, it_o)
    for f in "uvwpT":
        assert G.rel_l2(ctx.get("faces_" + f), o.get("faces_" + f)) < 1e-9, f
    for f in "uvwT":
        assert G.rel_l2(ctx.get(f + "xyz"), o.get(f + "xyz")) < 1e-9, f
    ctx.close()


def test_fixed_passes_match_oracle_colored(geo, steps):
    """A whole time step's worth of passes (4000) in the same order on both
    sides, without a stopping test: every field agrees to rounding. (Where
    a tolerance stops the loop, the pass it stops at is itself sensitive to
    rounding near u's floor, see test_converged_step_vs_reference_order.)"""
    ctx = G.gpu_context(geo)
    o = G.oracle_context(geo)
    ctx.upload_state(G.state_from_steps(steps, "s3"))
    R.load_state(o, steps, "s3")
    ctx.build_momentum_coeffs()
    o.velocity_constants()
    zero = np.zeros(5)
    rc, it_g, _ = ctx.inner_solve(zero, max_iter=4000, raise_on_cap=False)
    rc_o, it_o, _ = o.inner_solve(zero, O.COLORED, max_iter=4000)
    assert it_g == it_o == 4000
    for f in "uvwpT":
        assert G.rel_l2(ctx.get("faces_" + f), o.get("faces_" + f)) < 1e-9, f
    for f in "uvwT":
        assert G.rel_l2(ctx.get(f + "xyz"), o.get(f + "xyz")) < 1e-9, f
    ctx.close()


def test_fused_energy_sweep_is_bitwise_the_two_kernels(geo, steps, monkeypatch):
    """NIM_FUSE_T builds the energy coefficients inside the energy sweep
    (same formulas, same rounding) instead of storing them from a kernel of
    their own: 500 passes give the same fields bit for bit."""
    out = []
    for fuse in ("0", "1"):
        monkeypatch.setenv("NIM_FUSE_T", fuse)
        ctx = G.gpu_context(geo)
        ctx.upload_state(G.state_from_steps(steps, "s3"))
        ctx.build_momentum_coeffs()
        rc, it, _ = ctx.inner_solve(np.zeros(5), max_iter=500, raise_on_cap=False)
        assert it == 500
        out.append({n: ctx.get(n) for n in ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]})
        ctx.close()
    for n in out[0]:
        assert np.array_equal(out[0][n], out[1][n]), n


def test_converged_step_vs_reference_order(geo, steps):
    """GPU (4-colour order) against the oracle in the reference's own order,
    one time step from the reference's state after step 3, both converged
    to TOL_CONVERGED. Temperature and Nusselt numbers meet the parity bar
    (1e-8 relative L2, 1e-6).

    Velocity and pressure are only determined by the reference's iteration
    to the level its stopping pass is: p's constant mode never converges
    (its increment stays ~1.5e-8 per pass) and reaches u through the
    discrete gradient of a constant, which the face-centre normals make
    nonzero, so u's increment decreases very slowly near 1e-10 and the pass
    where it crosses the tolerance moves by tens of passes under rounding.
    The oracle alone shows the spread: a 1e-15 relative perturbation of the
    initial p moves its converged u by 7e-7, v by 1.2e-6, p by 1.2e-5
    (test_oracle_sensitivity.py), and its two orders differ by u 7e-7,
    p 6e-6. The bounds below allow a few times that spread."""
    ctx = G.gpu_context(geo)
    o = G.oracle_context(geo)
    ctx.upload_state(G.state_from_steps(steps, "s3"))
    R.load_state(o, steps, "s3")
    it_g, err_g, scal_g, nu_g = ctx.time_step(TOL_CONVERGED)
    rc, scal_o, nu_o = o.time_step(TOL_CONVERGED, O.NATURAL)
    assert rc == 0
    np.testing.assert_allclose(nu_g[:3], nu_o[:3], rtol=1e-6)
    np.testing.assert_allclose(nu_g[3:], nu_o[3:], rtol=1e-13)
    assert G.rel_l2(ctx.get("faces_T"), o.get("faces_T")) < 1e-8
    for suf in ("xyz", "xyz_old", "xyzt"):
        assert G.rel_l2(ctx.get("T" + suf), o.get("T" + suf)) < 1e-8, suf
    for f in "uvw":
        assert G.rel_l2(ctx.get("faces_" + f), o.get("faces_" + f)) < 2e-5, f
        assert G.rel_l2(ctx.get(f + "xyz"), o.get(f + "xyz")) < 2e-5, f
    pg, po = ctx.get("faces_p"), o.get("faces_p")
    assert G.rel_l2(pg - pg.mean(), po - po.mean()) < 2e-4
    ctx.close()


def test_iteration_cap_reports_error(geo, steps):
    ctx = G.gpu_context(geo)
    ctx.upload_state(G.state_from_steps(steps, "s0"))
    ctx.build_momentum_coeffs()
    import nim_amd
    rc, it, err = ctx.inner_solve(np.full(5, 1e-14), max_iter=3, raise_on_cap=False)
    assert rc == nim_amd.NIM_E_MAXITER and it == 3
    ctx.close()


@pytest.mark.parametrize("slab", ["2", "5"])
def test_slab_schedule_is_bitwise_the_colour_order(geo, steps, slab, monkeypatch):
    """The skewed slab schedule (colour c on slab s in launch s + 2c) gives
    every cell the same inputs as the four full-domain colour phases, so
    500 passes of either loop end on bit-identical states."""
    names = ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]

    def run(slab_env, mixed):
        if slab_env is None:
            monkeypatch.delenv("NIM_SLAB", raising=False)
        else:
            monkeypatch.setenv("NIM_SLAB", slab_env)
        ctx = G.gpu_context(geo)
        ctx.upload_state(G.state_from_steps(steps, "s3"))
        ctx.build_momentum_coeffs()
        if mixed:
            ctx.inner_solve_mixed(np.zeros(5), inner=4, max_cycles=126, raise_on_cap=False)
        else:
            ctx.inner_solve(np.zeros(5), max_iter=500, raise_on_cap=False)
        out = {n: ctx.get(n) for n in names}
        ctx.close()
        return out

    for mixed in (False, True):
        a, b = run(None, mixed), run(slab, mixed)
        for n in names:
            assert np.array_equal(a[n], b[n]), (n, mixed)


def test_plane_fastest_cell_order_is_bitwise(geo, steps, monkeypatch):
    """The per-cell kernels compute every cell from the same inputs whatever
    order their blocks run in: NIM_KFAST=1 (plane-fastest blocks) ends 300
    passes of either loop on the same bits."""
    names = ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]

    def run(flag, mixed):
        monkeypatch.setenv("NIM_KFAST", flag)
        ctx = G.gpu_context(geo)
        ctx.upload_state(G.state_from_steps(steps, "s3"))
        ctx.build_momentum_coeffs()
        if mixed:
            ctx.inner_solve_mixed(np.zeros(5), inner=4, max_cycles=76, raise_on_cap=False)
        else:
            ctx.inner_solve(np.zeros(5), max_iter=300, raise_on_cap=False)
        out = {n: ctx.get(n) for n in names}
        ctx.close()
        return out

    for mixed in (False, True):
        a, b = run("0", mixed), run("1", mixed)
        for n in names:
            assert np.array_equal(a[n], b[n]), (n, mixed)


def _bitwise_under(monkeypatch, var, dims, passes=200):
    """Run `passes` passes of the FP64 loop and of the mixed loop with the
    environment switch `var` at 0 and at 1 from the same seeded state of a
    case of dimensions `dims` (n_theta, n_r, n_z); every field must end on
    the same bits."""
    from nim_amd import host
    import nim_amd
    case = host.Case(dims, 0.25, 1.0, 0.8, Ra=1e4, Pr=0.71, DT=1e-3)
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-2, seed=7)
    names = ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]

    def run(flag, mixed):
        monkeypatch.setenv(var, flag)
        ctx = nim_amd.Context(*case.dims, case)
        ctx.upload_geometry(geo)
        ctx.build_pressure_coeffs()
        ctx.upload_state(st)
        ctx.build_momentum_coeffs()
        if mixed:
            ctx.inner_solve_mixed(np.zeros(5), inner=4, max_cycles=passes // 4, raise_on_cap=False)
        else:
            ctx.inner_solve(np.zeros(5), max_iter=passes, raise_on_cap=False)
        out = {n: ctx.get(n) for n in names}
        ctx.close()
        return out

    for mixed in (False, True):
        a, b = run("0", mixed), run("1", mixed)
        for n in names:
            assert np.array_equal(a[n], b[n]), (var, dims, n, mixed)


def test_radial_slab_order_is_bitwise(monkeypatch):
    """One radial slab per XCD (the default wherever n_r splits into 8 slabs
    of whole blocks) changes which thread takes which cell, not what a cell
    computes. 128 x 64 x 4 qualifies for the per-cell kernels (128 * 64 / 8 =
    1024 cells per slab and plane) and for the sweeps (64 * 32 / 8 = 256
    cells of a colour per slab and plane)."""
    _bitwise_under(monkeypatch, "NIM_RSLAB", (128, 64, 4))


@pytest.mark.parametrize("dims", [(128, 64, 4), (96, 13, 5)])
def test_tiled_sources_are_bitwise(monkeypatch, dims):
    """The energy sources from LDS-staged tiles (NIM_TILE=1, where n_theta is
    a multiple of 32) compute each term as the per-cell kernel does: 64-wide
    tiles at n_theta = 128, 32-wide tiles with a partial last row of tiles at
    n_theta = 96, n_r = 13."""
    _bitwise_under(monkeypatch, "NIM_TILE", dims)


@pytest.mark.parametrize("dims", [(128, 64, 4), (20, 6, 5), (36, 6, 12)])
def test_paired_sources_are_bitwise(monkeypatch, dims):
    """The FP64 source kernels with two cells of a half-row per thread
    (NIM_PAIR, the default where n_theta % 4 == 0) compute each cell as the
    per-cell kernels do, in the FP64 loop and in the mixed loop's FP64
    residual; on radial slabs at 128 x 64 x 4 (1,024 cells per slab and plane, two
    512-cell blocks), in storage order with a partial last block at
    20 x 6 x 5 and at the reference's 36 x 6 x 12."""
    _bitwise_under(monkeypatch, "NIM_PAIR", dims)


@pytest.mark.gpu
def test_non_temporal_sweep_loads_are_bitwise(monkeypatch):
    """The sweeps with non-temporal loads of what a pass reads once
    (NIM_NT=1, the default from 2^19 cells on) load the same values as with
    plain loads, in the FP64 loop and in the FP32 passes of the mixed loop."""
    _bitwise_under(monkeypatch, "NIM_NT", (128, 64, 4))


def test_alternating_plane_direction_is_bitwise(monkeypatch):
    """With radial slabs, the FP64 loop's kernels walk the planes alternately
    from the bottom and from the top (NIM_KREV, default on), so that each
    starts where its predecessor ended; only the order of the cells within
    a launch changes, so every field keeps its bits."""
    _bitwise_under(monkeypatch, "NIM_KREV", (128, 64, 4))
