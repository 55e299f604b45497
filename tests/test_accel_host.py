"""The host half of the accelerated loop: the least-squares solve of each
extrapolation step (libnim_hip.so's nim_internal_ls_solve, the routine the
solver calls; no device use). It takes the Gram matrix G = A^T A and
r = A^T f of the history A and must return a minimiser of ||f - A gamma||:
checked against numpy's least squares on full-rank histories, and, on
histories with dependent columns (a converging history's columns become
nearly parallel), for the residual it reaches."""
import ctypes

import numpy as np
import pytest

import nim_amd

P = ctypes.POINTER(ctypes.c_double)


def ls_solve(A, f):
    L = nim_amd.lib()
    L.nim_internal_ls_solve.restype = ctypes.c_int32
    L.nim_internal_ls_solve.argtypes = [P, P, ctypes.c_int32, P]
    G = np.ascontiguousarray(A.T @ A)
    r = np.ascontiguousarray(A.T @ f)
    g = np.zeros(A.shape[1])
    assert L.nim_internal_ls_solve(G.ctypes.data_as(P), r.ctypes.data_as(P), A.shape[1], g.ctypes.data_as(P)) == 0
    return g


@pytest.mark.parametrize("n", [1, 2, 5, 16])
def test_full_rank_history(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((400, n)) * np.logspace(0, -3, n)  # columns of different scales
    f = rng.standard_normal(400)
    g = ls_solve(A, f)
    ref = np.linalg.lstsq(A, f, rcond=None)[0]
    np.testing.assert_allclose(g, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())


def test_dependent_columns_reach_the_least_squares_residual():
    rng = np.random.default_rng(7)
    B = rng.standard_normal((500, 4))
    # 10 columns spanning 4 directions, two of them equal up to 1e-12
    A = np.column_stack([B @ rng.standard_normal(4) for _ in range(9)] + [B[:, 0] * (1 + 1e-12)])
    f = rng.standard_normal(500)
    g = ls_solve(A, f)
    assert np.all(np.isfinite(g))
    best = np.linalg.norm(f - A @ np.linalg.lstsq(A, f, rcond=None)[0])
    assert np.linalg.norm(f - A @ g) <= best * (1 + 1e-9)


def test_bad_arguments():
    L = nim_amd.lib()
    L.nim_internal_ls_solve.restype = ctypes.c_int32
    L.nim_internal_ls_solve.argtypes = [P, P, ctypes.c_int32, P]
    a = np.zeros(4)
    assert L.nim_internal_ls_solve(a.ctypes.data_as(P), a.ctypes.data_as(P), 0, a.ctypes.data_as(P)) == nim_amd.NIM_E_ARG
    assert L.nim_internal_ls_solve(a.ctypes.data_as(P), a.ctypes.data_as(P), 17, a.ctypes.data_as(P)) == nim_amd.NIM_E_ARG
