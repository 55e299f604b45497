"""The oracle against what the mathematics fixes, beside the bitwise
comparison with the compiled reference (test_oracle_reference.py).

- The exponential-fit constants (nim.f90:184-287): the closed forms
  evaluated at 50 digits (mpmath) against the oracle's double evaluation,
  in both branches. Below the switch |Re| = v_limit the oracle uses the
  reference's Taylor series: agreement to rounding there pins the series
  coefficients of the orders that are visible in double precision (t, t^3,
  t^5 with t = Re/2; a wrong sign in the t^5 term alone moves A1 by 1e-14;
  constants are numbered 1..6 as M_A(., 1..6)).
  Terms of order t^7 and above are below 1e-19 relative at |Re| <= 0.01,
  so only the bitwise comparison with the compiled reference pins them. (The
  reference's t^7 term of A1 and A2 has the opposite sign to the expansion
  of the closed form; it changes no double.)
  Above the switch the closed forms cancel: A5 keeps only about
  5e-14 / Re^2 relative accuracy, 3.5e-10 just above the switch; the bound
  below is that law.
- The limits Re -> 0: A5 -> a / (3 k), A6 -> 1/2, A1 -> -gp / k,
  A3 -> -gp / 2 (a = 1).
- The colour order (the GPU's) and the reference's natural order converge
  to the same temperature field and Nusselt numbers on the reference case.
"""
import mpmath as mp
import numpy as np
import pytest

import gpucase as G
import oracle as O
import refdata as R


def closed_forms(u, k, gp, gm):
    """nim.f90:184-287 (a = 1), the exponential branch, at 50 digits."""
    mp.mp.dps = 50
    u, k, gp, gm = mp.mpf(u), mp.mpf(k), mp.mpf(gp), mp.mpf(gm)
    Re = 2 * u / k
    e = mp.e ** Re
    return [(2 / (k * (1 / e - 1)) + 1 / u) * gp,
            (-(2 / (k * (1 - e))) - 1 / u) * gm,
            (u / (k * (1 / e - 1))) * gp,
            (u / (k * (1 - e))) * gm,
            (-k + u * (1 / e + 1) / (-1 / e + 1)) / (u * u),
            1 / (-1 / e + 1) - 1 / Re]


RES = [1e-6, 1e-3, 0.0099, 0.0101, 0.05, 0.5, 5.0, 50.0]


@pytest.mark.parametrize("k", [0.01, 1.0, 30.0])
def test_exponential_fits_against_high_precision(k):
    gp, gm = 0.7, 1.3
    for Re in RES:
        for sgn in (1.0, -1.0):
            u = sgn * Re * k / 2
            a = O.nim_direction(u, k, gp, gm)
            b = closed_forms(u, k, gp, gm)
            for m in range(6):
                rel = float(abs(a[m] - b[m]) / abs(b[m]))
                if Re <= 0.01:
                    bound = 1e-15
                elif m == 4:
                    bound = max(5e-14 / Re ** 2, 1e-15)
                else:
                    bound = 1e-12
                assert rel <= bound, (k, sgn * Re, m + 1, rel, bound)


def test_limits_at_zero_velocity():
    k, gp, gm = 2.5, 0.7, 1.3
    a = O.nim_direction(0.0, k, gp, gm)
    np.testing.assert_allclose(a, [-gp / k, -gm / k, -gp / 2, -gm / 2, 1 / (3 * k), 0.5], rtol=1e-15)


def test_colour_and_natural_orders_share_the_fixed_point():
    """From the reference's state after step 3, one step converged to 1e-10
    (p 1e-6) in each order: T and the Nusselt numbers agree within the
    parity bar. u, v, w and p differ by where each loop stops, which the
    pressure level's drift moves (test_oracle_sensitivity.py)."""
    geo, st = R.geometry(), R.steps()
    tol = np.array([1e-10, 1e-10, 1e-10, 1e-6, 1e-10])
    out = {}
    for order in (O.NATURAL, O.COLORED):
        o = G.oracle_context(geo)
        R.load_state(o, st, "s3")
        rc, scal, nu = o.time_step(tol, order)
        assert rc == 0
        out[order] = (o.get("faces_T").copy(), o.get("Txyz").copy(), np.array(nu[:3]))
        o.close()
    (Tn, Tcn, nun), (Tc, Tcc, nuc) = out[O.NATURAL], out[O.COLORED]
    assert G.rel_l2(Tc, Tn) < 1e-8
    assert G.rel_l2(Tcc, Tcn) < 1e-8
    np.testing.assert_allclose(nuc, nun, rtol=1e-6)
