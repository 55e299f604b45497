"""nim_end_step_noalloc (the allocation-free form of nim_end_step) against
nim_end_step through the C-ABI: the same kernels with their partial results
in the library's fixed device array, so the three outputs and the
dilatation field must be bit-identical. Two calls in a row, the second on a
smaller case, check that the shared array carries nothing from one call to
the next. A host output pointer is rejected with NIM_E_ARG before any launch.
(The size limit, 16,777,216 cells, is checked on the host before the
pointer; a context that large needs 32 GB and is not made here.)"""
import ctypes

import numpy as np
import pytest

import gpucase as G
import nim_amd
from nim_amd import host

pytestmark = pytest.mark.gpu


def _stepped_context(case, passes):
    geo = host.build_geometry(case, device=True)
    st = host.initial_state(case, amp=1e-3, seed=20221)
    ctx = G.gpu_for_case(case, geo, st)
    ctx.build_momentum_coeffs()
    ctx.inner_solve(np.zeros(5), max_iter=passes, raise_on_cap=False)
    return ctx


def test_noalloc_end_step_matches_end_step_twice():
    import torch
    torch.cuda.set_device(0)
    cases = [host.baseline_case("C1"), host.Case((24, 8, 6), 0.3, 1.0, 1.0, Ra=3e4, Pr=0.71, DT=1e-3)]
    for n, case in enumerate(cases):
        print("noalloc end step, case", n, case.dims, flush=True)
        ctx = _stepped_context(case, 40)
        ref = ctx.end_step()
        dila_ref = ctx.get("meshes_dila")
        ctx.set("meshes_dila", np.zeros(ctx.nmesh))
        out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        ctx.end_step_noalloc(out.data_ptr())
        ctx.synchronize()
        got = out.cpu().numpy()
        want = np.array([ref["dudt_max"], ref["dTdt_max"], ref["Dilatation"]])
        assert np.all(np.isfinite(want)) and want[2] > 0
        assert np.array_equal(got, want), (got, want)
        assert np.array_equal(ctx.get("meshes_dila"), dila_ref)
        ctx.close()


def test_noalloc_end_step_rejects_bad_output_pointer():
    case = host.Case((24, 8, 6), 0.3, 1.0, 1.0, Ra=3e4, Pr=0.71, DT=1e-3)
    ctx = _stepped_context(case, 2)
    L = nim_amd.lib()
    host_buf = np.zeros(3)
    assert L.nim_end_step_noalloc(ctx.h, ctypes.c_void_p(host_buf.ctypes.data)) == nim_amd.NIM_E_ARG
    assert L.nim_end_step_noalloc(ctx.h, None) == nim_amd.NIM_E_ARG
    ctx.close()
