"""The device's rest state on the 12 x 4 x 4 mesh of
tests/test_oracle_fixed_point.py is the oracle's in the same colour order:
6,000 passes from the seeded state, every field to rounding: 1e-12 with
the Taylor branch of the exponential fits everywhere (v_limit 0.3), 1e-9
with the reference's switch, where the closed forms amplify the last bit
of exp. With the oracle test, which shows
that the colour and natural orders share this fixed point once the pressure
system is compatible, this ties the device to the reference's fixed point."""
import numpy as np
import pytest

import gpucase as G
import nim_amd
import oracle as O
from nim_amd import host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("v_limit", [1e-2, 0.3])
def test_device_rest_state_is_the_oracle_colour_order(v_limit):
    case = host.Case((12, 4, 4), 0.3, 1.0, 1.0, Ra=1e4, Pr=0.71, DT=1e-3, v_limit=v_limit)
    geo = host.build_geometry(case, device=False)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    ctx = G.gpu_for_case(case, geo, st)
    o = G.oracle_for_case(case, geo, st)
    ctx.set_p_measure(nim_amd.P_HALFRANGE)
    o.set_p_measure(1)
    ctx.build_momentum_coeffs()
    o.velocity_constants()
    n = 6000
    rc, it, err = ctx.inner_solve(np.zeros(5), max_iter=n, raise_on_cap=False)
    rc_o, it_o, err_o = o.inner_solve(np.zeros(5), O.COLORED, max_iter=n)
    assert it == it_o == n
    assert max(err[:3]) < 1e-15 and err[3] < 1e-13, err
    gaps = {}
    for f in ["faces_" + f for f in "uvwpT"] + [f + "xyz" for f in "uvwT"]:
        a, b = ctx.get(f), o.get(f)
        if f == "faces_p":
            a, b = a - a.mean(), b - b.mean()
        gaps[f] = G.rel_l2(a, b)
    print("v_limit", v_limit, gaps)
    for f, g in gaps.items():
        # just above the switch the fits amplify one ulp of exp (the
        # device's is not glibc's) by about 1e7
        assert g < (1e-9 if v_limit == 1e-2 else 1e-12), (f, g)
    ctx.close()
    o.close()
