"""The oracle's half-range pressure measure, on the CPU: a uniform shift of
every pressure face gives a zero half range and a max |dp| of the shift,
and a shift plus a ramp gives half the ramp's range, computed independently
here."""
import numpy as np

import oracle as O
import gpucase as G
import refdata as R


def test_halfrange_is_blind_to_a_constant():
    geo = R.geometry()
    o = G.oracle_context(geo)
    steps = R.steps()
    R.load_state(o, steps, "s3")
    o.find_error()  # old = current
    p = o.get("faces_p")
    rng = np.random.default_rng(3)
    ramp = rng.uniform(-1e-6, 2e-6, p.size)
    for mode in (0, 1):
        o.set_p_measure(mode)
        o.set("faces_p", p)
        o.find_error()
        o.set("faces_p", p + 0.125)
        e = o.find_error()
        dp = (p + 0.125) - p
        want = np.max(np.abs(dp)) if mode == 0 else (dp.max() - dp.min()) / 2
        assert e[3] == want
        assert (e[3] == 0.125) if mode == 0 else (e[3] < 1e-15)
        o.set("faces_p", p + 0.125 + ramp)
        e = o.find_error()
        dp = (p + 0.125 + ramp) - (p + 0.125)
        want = np.max(np.abs(dp)) if mode == 0 else (dp.max() - dp.min()) / 2
        assert e[3] == want, (mode, e[3], want)
        if mode == 1:
            assert abs(e[3] - (ramp.max() - ramp.min()) / 2) < 1e-15
    o.close()
