"""C-ABI library: loads, exports every declared entry point, and its host-side
layout mapping agrees with the reference's connectivity (no device use)."""
import ctypes
import os
import re

import numpy as np

import nim_amd
import refdata as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "nim_hip.h")).read()
    return sorted(set(re.findall(r"\b(nim_[a-z_]+)\s*\(", text)))


def test_library_exports_all_declared_symbols():
    L = ctypes.CDLL(nim_amd.LIB_PATH)
    names = declared_entry_points()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(nim_amd.exported_symbols()) == names


def _check_face_map(nt, nr, nz, meshes_faces):
    """nim_face_map is a permutation, and the device slot of every
    cell-local face holds the reference face meshes_faces (1-based) names."""
    nmesh, nface, _ = nim_amd.counts(nt, nr, nz)
    fm = nim_amd.face_map(nt, nr, nz)
    assert sorted(fm.tolist()) == list(range(nface))
    # device slot of each cell-local face must be the reference's meshes_faces
    mf = np.asarray(meshes_faces).reshape(nmesh, 6) - 1
    oT = nt * (nr + 1) * nz
    oZ = oT + nmesh

    def P(i):  # storage position of azimuthal index i: even cells, then odd
        return (i % 2) * (nt // 2) + i // 2

    for c in range(nmesh):
        i, j, k = c % nt, (c // nt) % nr, c // (nt * nr)
        slots = [P(i) + nt * (j + (nr + 1) * k), P(i) + nt * (j + 1 + (nr + 1) * k), oT + P(i) + nt * (j + nr * k),
                 oT + P((i + 1) % nt) + nt * (j + nr * k), oZ + P(i) + nt * (j + nr * k),
                 oZ + P(i) + nt * (j + nr * (k + 1))]
        assert list(fm[slots]) == list(mf[c]), c


def test_counts_and_face_map_match_reference_numbering():
    geo = R.geometry()
    nt, nr, nz = R.REF_DIMS
    nmesh, nface, nbc = nim_amd.counts(nt, nr, nz)
    assert (nmesh, nface, nbc) == (int(geo["dims"][3]), int(geo["dims"][4]), int(geo["dims"][5]))
    _check_face_map(nt, nr, nz, geo["meshes_faces"])


def test_face_map_at_the_48_x_24_x_16_case():
    """The same at the 48 x 24 x 16 case, against the host geometry's
    meshes_faces, which tests/golden/ref_c1.json pins bit for bit to the
    reference built at that size (test_oracle_reference_c1.py)."""
    from nim_amd import host
    case = host.baseline_case("C1")
    geo = host.build_geometry(case)
    _check_face_map(*case.dims, geo["meshes_faces"])


def test_unsupported_dimensions_rejected():
    import pytest
    with pytest.raises(nim_amd.NimError):
        nim_amd.counts(35, 6, 12)  # odd azimuthal count cannot be 4-coloured


def test_cases_beyond_32_bit_device_indexing_rejected():
    """The device code indexes slot groups as slot * N + c in 32-bit ints
    (up to 45 N for the pressure coefficients, 5 NF for the face fields):
    nim_counts and nim_ctx_create refuse larger cases with NIM_E_ARG before
    any device use."""
    import pytest
    # 47,185,920 cells: the largest 45 N within INT32_MAX here
    assert nim_amd.counts(1024, 512, 90)[0] == 47185920
    for dims in ((1024, 512, 134), (1024, 512, 92), (65536, 65536, 65536), (2**30, 2**30, 2**30)):
        with pytest.raises(nim_amd.NimError):
            nim_amd.counts(*dims)
        L = nim_amd.lib()
        h = ctypes.c_void_p()
        prm = nim_amd.Params()
        assert L.nim_ctx_create(ctypes.byref(nim_amd.Dims(*dims)), ctypes.byref(prm), ctypes.byref(h)) == \
            nim_amd.NIM_E_ARG
        assert b"47,721,858" in L.nim_last_error()


def test_argument_errors_before_any_device_use():
    """Entry points validate their arguments before they touch a device, so
    these run without one and return NIM_E_ARG with a message."""
    L = nim_amd.lib()
    out = ctypes.c_void_p()
    assert L.nim_adj_create(None, ctypes.byref(out)) == nim_amd.NIM_E_ARG
    assert L.nim_adj_reverse_step(None, 1e-10, 10, None, None) == nim_amd.NIM_E_ARG
    assert L.nim_adj_destroy(None) == nim_amd.NIM_OK
    assert b"reverse-step" in L.nim_last_error()
    a = np.zeros(8)
    P = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    bad = nim_amd.Dims(35, 6, 12)
    assert L.nim_device_metrics(ctypes.byref(bad), P, *[P] * 12) == nim_amd.NIM_E_ARG
    assert L.nim_device_metrics(ctypes.byref(nim_amd.Dims(36, 6, 12)), None, *[P] * 12) == nim_amd.NIM_E_ARG
    codes = np.zeros(4, dtype=np.int32)
    n, fwd = ctypes.c_int64(), ctypes.c_int64()
    PI = codes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    assert L.nim_revolve_plan(0, 2, PI, PI, 4, ctypes.byref(n), ctypes.byref(fwd)) == nim_amd.NIM_E_ARG


def test_accelerated_solve_arguments_checked_before_device_use():
    """nim_inner_solve_accel, nim_accel_options and nim_accel_scope refuse a
    null context (and the solve a null tolerance) with NIM_E_ARG, without a
    device."""
    L = nim_amd.lib()
    tol = np.full(5, 1e-8)
    err = np.zeros(5)
    it = ctypes.c_int32()
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.nim_accel_options(None, 16, 8, 4) == nim_amd.NIM_E_ARG
    assert L.nim_accel_scope(None, nim_amd.ACCEL_PRESSURE) == nim_amd.NIM_E_ARG
    assert L.nim_inner_solve_accel(None, P(tol), 100, 0, 0, ctypes.byref(it), P(err)) == nim_amd.NIM_E_ARG
    assert b"null" in L.nim_last_error()
