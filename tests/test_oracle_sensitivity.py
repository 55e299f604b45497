"""How far the reference's own iteration determines a converged step.

The oracle runs the reference's algorithm (nse.f90:117-176) from the
reference's state after step 3 twice: once as is, once with the initial
pressure faces perturbed by 1e-15 relative. Temperature and the Nusselt
numbers stay within the parity bar. Velocity and pressure move by orders of
magnitude more than the perturbation, because the pass at which the
stopping test fires moves (p's constant mode never converges and reaches u
through the nonzero discrete gradient of a constant). The spread measured
here bounds what any sweep order or rounding can be held to for u, v, w
and p at these tolerances; the GPU comparison against the reference order
uses a few times this spread.
"""
import numpy as np
import pytest

import gpucase as G
import oracle as O
import refdata as R

TOL_CONVERGED = np.array([1e-10, 1e-10, 1e-10, 1e-6, 1e-10])


@pytest.fixture(scope="module")
def runs():
    geo, st = R.geometry(), R.steps()
    out = {}
    for name, eps in (("base", 0.0), ("perturbed", 1e-15)):
        o = G.oracle_context(geo)
        R.load_state(o, st, "s3")
        if eps:
            p = o.get("faces_p").copy()
            rng = np.random.default_rng(1)
            o.set("faces_p", p * (1 + eps * rng.standard_normal(p.shape)))
        rc, scal, nu = o.time_step(TOL_CONVERGED, O.COLORED)
        assert rc == 0
        out[name] = dict(iters=scal[0], nu=np.array(nu), **{f: o.get("faces_" + f).copy() for f in "uvwpT"})
        o.close()
    return out


def test_temperature_and_nusselt_are_determined(runs):
    a, b = runs["base"], runs["perturbed"]
    assert G.rel_l2(a["T"], b["T"]) < 1e-8
    np.testing.assert_allclose(a["nu"][:3], b["nu"][:3], rtol=1e-6)


def test_velocity_and_pressure_spread(runs):
    a, b = runs["base"], runs["perturbed"]
    du = max(G.rel_l2(a[f], b[f]) for f in "uvw")
    dp = G.rel_l2(a["p"] - a["p"].mean(), b["p"] - b["p"].mean())
    # the stopping pass moves, and with it u and p, far beyond 1e-15 ...
    assert a["iters"] != b["iters"]
    assert du > 1e-8 and dp > 1e-7
    # ... but within the bounds the GPU comparison uses
    assert du < 2e-5 and dp < 2e-4
