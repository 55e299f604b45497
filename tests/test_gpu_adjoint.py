"""Discrete adjoint of the time step.

The reference has no adjoint, so parity is with the mathematics:
  - each reversed stage is the transpose of the Jacobian of the forward
    stage (nim_run_stage) at a converged state: <lam, (S(x + e v) - S(x - e v))
    / 2e> = <S'(x)^T lam, v> for random lam and v over every array the stages
    read and write (dot-product test against central differences of the
    forward kernels themselves);
  - one whole reversed pass, likewise, with the pass's intermediates and
    parameters in place;
  - dNu/dT_hot and dNu/dRa over a run of time steps reversed with fewer
    snapshots than steps (binomial checkpointing) against central finite
    differences of the forward run, to 1e-5;
  - the reversed fixed point solved by GMRES gives the plain iteration's
    gradient with far fewer reversed passes;
  - the torch.autograd harness returns the same gradient.
The specification's case (48 x 96 x 24, per-face wall temperatures, the
time-averaged Nusselt number) is in tests/test_gpu_adjoint_c4.py.
"""
import numpy as np
import pytest

import nim_amd
import nim_amd.autograd
from nim_amd import host

pytestmark = pytest.mark.gpu

CELLS = ["uxyz", "vxyz", "wxyz", "Txyz"]
FACES = ["faces_u", "faces_v", "faces_w", "faces_p", "faces_T"]
OLD = ["uxyz_old", "vxyz_old", "wxyz_old", "Txyz_old"]
TC = ["uxyzt", "vxyzt", "wxyzt"]
SRC = ["R_B_u", "R_B_v", "R_B_w", "R_b_p", "R_D_T"]
COEF = ["M_F1", "M_F2", "M_F3", "M_F4", "MT_F1", "MT_F2", "MT_F3", "MT_F4"]
ALL = CELLS + FACES + OLD + TC + SRC + COEF


def small_case(**kw):
    args = dict(Ra=1e4, Pr=0.71, DT=2e-3)
    args.update(kw)
    return host.Case((24, 8, 8), 0.3, 1.0, 1.0, **args)


def converged_context(case, amp=0.2):
    """A context at the converged state of one step from a strongly
    perturbed initial state (velocities large enough that the exponential
    branch of the fits is taken in many cells)."""
    ctx = nim_amd.autograd.make_context(case, amp=amp, seed=7)
    ctx.build_momentum_coeffs()
    # p is determined up to a constant, and that level drifts by a constant
    # amount per pass; the other fields converge
    tol = np.array([1e-13, 1e-13, 1e-13, 1e0, 1e-13])
    rc, it, err = ctx.inner_solve(tol, max_iter=20000, raise_on_cap=False)
    print("converged state: passes", it, "increments", err.tolist())
    # the FP64 loop builds the energy coefficients inside its energy sweep;
    # stage 4 stores them (and moves the adiabatic walls by the increment)
    ctx.run_stage(4)
    # (T creeps by ~5e-11 per pass on a state of size 0.2 here; far below
    # what the dot-product tests resolve)
    assert np.all(err[[0, 1, 2, 4]] < 1e-9), (it, err.tolist())
    return ctx


@pytest.fixture(scope="module")
def conv():
    case = small_case()
    ctx = converged_context(case)
    x0 = {n: ctx.get(n) for n in ALL}
    yield case, ctx, x0
    ctx.close()


def scale(a):
    s = np.sqrt(np.mean(a * a))
    return s if s > 0 else 1.0


def run_forward(ctx, stage):
    if stage == 7:
        ctx.build_momentum_coeffs()
    elif stage == "pass":
        for s in range(7):
            ctx.run_stage(s)
    else:
        ctx.run_stage(stage)


def restore(ctx, x0):
    for n, v in x0.items():
        ctx.set(n, v)


def dot_test(ctx, x0, stage, lam, v, eps, order=2):
    """<lam, central difference of the stage along v> and the adjoint side
    <S'^T lam, v> (with the adjoint taken at x0). order 4 uses the points
    +-eps and +-2 eps."""
    out = {}
    steps = (1.0, -1.0) if order == 2 else (1.0, -1.0, 2.0, -2.0)
    for sgn in steps:
        restore(ctx, {n: x0[n] + sgn * eps * v[n] for n in ALL})
        run_forward(ctx, stage)
        out[sgn] = {n: ctx.get(n) for n in ALL}

    def deriv(n):
        d1 = (out[1.0][n] - out[-1.0][n]) / (2 * eps)
        if order == 2:
            return d1
        d2 = (out[2.0][n] - out[-2.0][n]) / (4 * eps)
        return (4 * d1 - d2) / 3

    lhs = sum(np.dot(lam[n], deriv(n)) for n in ALL if n in lam)
    restore(ctx, x0)
    adj = nim_amd.Adjoint(ctx)
    try:
        adj.clear()
        for n, l in lam.items():
            adj.set(n, l)
        if stage == "pass":
            adj.reverse_pass()
        else:
            adj.reverse_stage(stage)
        rhs = sum(np.dot(adj.get(n), v[n]) for n in ALL)
    finally:
        adj.close()
    restore(ctx, x0)
    return lhs, rhs


STAGE_NAMES = {0: "momentum sources", 1: "momentum sweep", 2: "pressure sources", 3: "pressure sweep",
               4: "energy coefficients and adiabatic walls", 5: "energy sources", 6: "energy sweep",
               7: "momentum coefficients"}


# The exponential fits (stages 4 and 7) lose accuracy to cancellation in
# their closed forms just above the Taylor switch (relative error about
# 1e-16 / Re^2), which swamps a central difference at small steps; those
# stages use a fourth-order difference at a step of 1e-3 instead.
FD_STEP = {4: (1e-3, 4), 7: (1e-3, 4)}


@pytest.mark.parametrize("stage", list(range(8)), ids=lambda s: STAGE_NAMES[s].replace(" ", "_"))
def test_reversed_stage_is_the_transposed_jacobian(conv, stage):
    case, ctx, x0 = conv
    rng = np.random.default_rng(100 + stage)
    v = {n: rng.standard_normal(x0[n].size) * scale(x0[n]) for n in ALL}
    lam = {n: rng.standard_normal(x0[n].size) for n in ALL}
    eps, order = FD_STEP.get(stage, (1e-6, 2))
    lhs, rhs = dot_test(ctx, x0, stage, lam, v, eps, order)
    print(STAGE_NAMES[stage], lhs, rhs, (lhs - rhs) / max(abs(lhs), abs(rhs)))
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs)), (STAGE_NAMES[stage], lhs, rhs)


def dirichlet_mask(case, name):
    """Face slots of the reference numbering that no pass writes: u, v, w on
    every wall, T on the hexagon and cylinder (BC groups 2-5 and 4-5 of
    create_geometry.f90:626-664, the walls at the end of the numbering)."""
    nt, nr, nz = case.dims
    N = nt * nr * nz
    NF = 3 * N + nt * nz + nt * nr
    o = NF - 2 * nt * nr - 2 * nt * nz
    m = np.zeros(NF, dtype=bool)
    if name in ("faces_u", "faces_v", "faces_w"):
        m[o:] = True
    elif name == "faces_T":
        m[o + 2 * nt * nr:] = True
    return m


def test_reversed_pass_is_the_transposed_jacobian(conv):
    """One outer pass as a map of the iterate (cells and faces) with the
    previous level, momentum coefficients and wall values as parameters: the
    reversed pass gives the transposed Jacobian for both."""
    case, ctx, x0 = conv
    rng = np.random.default_rng(7)
    v = {n: rng.standard_normal(x0[n].size) * scale(x0[n]) for n in ALL}
    lam = {}
    for n in CELLS + FACES:
        lam[n] = rng.standard_normal(x0[n].size)
        if n.startswith("faces_"):
            lam[n][dirichlet_mask(case, n)] = 0.0
    lhs, rhs = dot_test(ctx, x0, "pass", lam, v, 1e-6)
    print("pass", lhs, rhs, (lhs - rhs) / max(abs(lhs), abs(rhs)))
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs)), (lhs, rhs)


NSTEPS, NSLOTS, PASSES = 4, 2, 1500


def nusselt_after(case):
    ctx = nim_amd.autograd.make_context(case)
    try:
        return nim_amd.run_steps(ctx, NSTEPS, passes=PASSES)[1]
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def adjoint_run():
    case = small_case(Ra=5e4)
    ctx = nim_amd.autograd.make_context(case)
    try:
        J, grad, info = nim_amd.nusselt_gradient(ctx, case, NSTEPS, NSLOTS, 1e-13, passes=PASSES, solver="plain")
    finally:
        ctx.close()
    return case, J, grad, info


def test_gradient_matches_finite_differences(adjoint_run):
    """Hexagon Nusselt number after 4 steps, reversed with 2 snapshots:
    the adjoint dNu/dT_hot and dNu/dRa against central differences."""
    case, J, grad, info = adjoint_run
    print("J", J, "adjoint", grad, info)
    assert info["forward_steps"] > NSTEPS - 1  # the plan re-advanced from a snapshot
    assert abs(J - nusselt_after(case)) <= 1e-12 * abs(J)
    hT, hR = 1e-4, 1e-4 * case.Ra
    fd_T = (nusselt_after(case.replace(T_hot=case.T_hot + hT)) - nusselt_after(case.replace(T_hot=case.T_hot - hT))) / (2 * hT)
    fd_R = (nusselt_after(case.replace(Ra=case.Ra + hR)) - nusselt_after(case.replace(Ra=case.Ra - hR))) / (2 * hR)
    print("finite differences: T_hot", fd_T, "Ra", fd_R)
    assert abs(grad["T_hot"] - fd_T) <= 1e-5 * abs(fd_T), (grad, fd_T, info)
    assert abs(grad["Ra"] - fd_R) <= 1e-5 * abs(fd_R), (grad, fd_R, info)


def test_gmres_matches_plain_iteration(adjoint_run):
    """Restarted GMRES on (I - G_x^T) lam = seed against the plain reversed
    iteration (about 7,000 passes per step here): the same gradient, the
    per-face wall gradients summing to the scalar one, at a small fraction
    of the reversed passes."""
    case, J, grad, info = adjoint_run
    ctx = nim_amd.autograd.make_context(case)
    try:
        Jg, gg, ig = nim_amd.nusselt_gradient(ctx, case, NSTEPS, NSLOTS, 1e-12, passes=PASSES, solver="gmres")
    finally:
        ctx.close()
    plain = sum(n for n, _ in info["adjoint_passes"])
    krylov = sum(n for n, _ in ig["adjoint_passes"])
    print("reversed passes: plain", plain, "GMRES", krylov, ig["adjoint_passes"])
    assert Jg == J
    assert abs(gg["T_hot"] - grad["T_hot"]) <= 1e-9 * abs(grad["T_hot"])
    assert abs(gg["Ra"] - grad["Ra"]) <= 1e-9 * abs(grad["Ra"])
    assert abs(gg["T_faces"].sum() - gg["T_hot"]) <= 1e-12 * abs(gg["T_hot"])
    np.testing.assert_allclose(gg["T_faces"], grad["T_faces"], rtol=1e-8, atol=1e-12 * abs(grad["T_hot"]))
    assert krylov * 10 < plain


@pytest.mark.parametrize("basis,augment,degree", [("fp64", 0, 1), ("fp64", 3, 1), ("fp32", 3, 1), ("fp64", 3, 4)])
def test_gmres_options_match_plain_iteration(adjoint_run, basis, augment, degree):
    """Short cycles (GMRES(20)) with and without LGMRES augmentation, the
    Krylov basis stored in FP64 or FP32 (all arithmetic FP64 either way,
    each cycle started from the FP64 true residual), one or four reversed
    passes per operator application (the Neumann-preconditioned system):
    the plain iteration's gradient to the solve's tolerance."""
    case, J, grad, info = adjoint_run
    ctx = nim_amd.autograd.make_context(case)
    try:
        Jg, gg, ig = nim_amd.nusselt_gradient(ctx, case, NSTEPS, NSLOTS, 1e-12, passes=PASSES, solver="gmres",
                                              restart=20, basis=basis, augment=augment, degree=degree)
    finally:
        ctx.close()
    print(basis, "basis,", augment, "augmentation vectors, degree", degree, ": reversed passes", ig["adjoint_passes"])
    assert Jg == J
    assert abs(gg["T_hot"] - grad["T_hot"]) <= 1e-9 * abs(grad["T_hot"])
    assert abs(gg["Ra"] - grad["Ra"]) <= 1e-9 * abs(grad["Ra"])
    np.testing.assert_allclose(gg["T_faces"], grad["T_faces"], rtol=1e-8, atol=1e-12 * abs(grad["T_hot"]))
    assert all(r <= 1e-12 for _, r in ig["adjoint_passes"])


def test_autograd_harness(adjoint_run):
    import torch
    case, J, grad, info = adjoint_run
    T = torch.tensor(case.T_hot, dtype=torch.float64, requires_grad=True)
    Ra = torch.tensor(case.Ra, dtype=torch.float64, requires_grad=True)
    run = dict(case=case, nsteps=NSTEPS, nslots=NSLOTS, adj_tol=1e-13, passes=PASSES, solver="plain")
    out = nim_amd.autograd.nusselt(T, Ra, run)
    gT, gR = torch.autograd.grad(3.0 * out, (T, Ra))
    assert abs(out.item() - J) <= 1e-12 * abs(J)
    assert abs(gT.item() - 3.0 * grad["T_hot"]) <= 1e-9 * abs(3.0 * grad["T_hot"])
    assert abs(gR.item() - 3.0 * grad["Ra"]) <= 1e-9 * abs(3.0 * grad["Ra"])


def test_reversed_stages_beyond_one_grid():
    """The reversed sources and coefficient builds loop over faces or cells
    in grid-stride fashion on a grid of at most 2048 blocks of 256 threads,
    so that each block adds its parameter sum with one atomic. Here (128 x
    48 x 32 cells, 600,064 face slots > 2048 * 256) every thread takes more
    than one item. The state is the perturbed initial one, not converged:
    each stage's transpose holds at any state. Random velocities spread the
    cell Peclet numbers Re = 2 |u| h / nu over every value up to their
    maximum, and near the switch between the fits' Taylor and closed forms
    (|Re| = 1e-2) the closed forms' cancellation (about 5e-14 / Re^2)
    swamps a finite difference: 3e-7 and 1e-7 of the coefficient stages'
    dot products at perturbations of 0.2 and 1e-2. At 1e-4 (Re below 2e-3)
    every fit takes its Taylor branch; the closed forms are covered by the
    converged case above."""
    case = host.Case((128, 48, 32), 0.3, 1.0, 1.0, Ra=1e4, Pr=0.71, DT=2e-3)
    ctx = nim_amd.autograd.make_context(case, amp=1e-4, seed=11)
    try:
        ctx.build_momentum_coeffs()
        ctx.run_stage(4)
        x0 = {n: ctx.get(n) for n in ALL}
        assert x0["faces_u"].size > 2048 * 256
        for stage in (0, 2, 4, 5, 7):
            rng = np.random.default_rng(300 + stage)
            v = {n: rng.standard_normal(x0[n].size) * scale(x0[n]) for n in ALL}
            lam = {n: rng.standard_normal(x0[n].size) for n in ALL}
            eps, order = FD_STEP.get(stage, (1e-6, 2))
            lhs, rhs = dot_test(ctx, x0, stage, lam, v, eps, order)
            print(STAGE_NAMES[stage], lhs, rhs, (lhs - rhs) / max(abs(lhs), abs(rhs)))
            assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs)), (STAGE_NAMES[stage], lhs, rhs)
    finally:
        ctx.close()
