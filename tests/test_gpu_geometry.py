"""Per-cell metrics on the device (nim_device_metrics) against the host's
Fortran restatement, which is itself bit-identical to the reference on the
reference's case (test_host_geometry.py): the same expression order without
fused multiply-adds, so every array agrees bit for bit."""
import time

import numpy as np
import pytest

from nim_amd import host

pytestmark = pytest.mark.gpu

CASES = {"reference": host.reference_case(), "C1": host.baseline_case("C1"),
         "uneven": host.Case((30, 7, 5), 0.25, 1.0, 0.8, Ra=1e4, Pr=0.71, DT=1e-3)}


@pytest.mark.parametrize("name", list(CASES))
def test_device_metrics_bitwise(name):
    case = CASES[name]
    t0 = time.perf_counter()
    a = host.build_geometry(case)
    t1 = time.perf_counter()
    b = host.build_geometry(case, device=True)
    t2 = time.perf_counter()
    print("%s: host %.3f s, device %.3f s" % (name, t1 - t0, t2 - t1))
    diff = {}
    for k in a:
        if not np.array_equal(a[k], b[k]):
            x, y = np.asarray(a[k], dtype=np.float64), np.asarray(b[k], dtype=np.float64)
            diff[k] = float(np.max(np.abs(x - y)) / max(np.max(np.abs(x)), 1e-300))
    assert not diff, diff
