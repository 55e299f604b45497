"""The GPU against the oracle at the sizes the benchmark is quoted on, from
the seeded initial state: C1 (48 x 24 x 16 cells, theta x r x z) and C2
(256 x 128 x 32, 1,048,576 cells). Both sides run the same number of passes
in the colour order, with no stopping test; every face and cell field and
the Nusselt numbers must agree to rounding. At C2 the per-cell kernels and
the FP64 sweeps give each XCD one radial slab (n_r = 128 splits into 8), at
C1 they keep storage order (n_theta * n_r / 8 is not a multiple of 256), so
both block orders are checked at full size. The pass counts keep the oracle
to a few seconds on the host's threads."""
import numpy as np
import pytest

import gpucase as G
import oracle as O
from nim_amd import host

pytestmark = pytest.mark.gpu

PASSES = {"C1": 1000, "C2": 48}


@pytest.mark.parametrize("name", list(PASSES))
def test_passes_match_oracle_at_benchmark_size(name):
    case = host.baseline_case(name)
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    ctx = G.gpu_for_case(case, geo, st)
    o = G.oracle_for_case(case, geo, st)
    ctx.build_momentum_coeffs()
    o.velocity_constants()
    n = PASSES[name]
    zero = np.zeros(5)
    rc, it_g, _ = ctx.inner_solve(zero, max_iter=n, raise_on_cap=False)
    rc_o, it_o, _ = o.inner_solve(zero, O.COLORED, max_iter=n)
    assert it_g == it_o == n
    errs = {}
    for f in "uvwpT":
        errs["faces_" + f] = G.rel_l2(ctx.get("faces_" + f), o.get("faces_" + f))
    for f in "uvwT":
        errs[f + "xyz"] = G.rel_l2(ctx.get(f + "xyz"), o.get(f + "xyz"))
    print(name, n, "passes:", " ".join("%s %.1e" % kv for kv in errs.items()))
    assert max(errs.values()) < 1e-9, errs
    np.testing.assert_allclose(ctx.nusselt()[:3], o.nusselt()[:3], rtol=1e-9)
    ctx.close()
    o.close()
