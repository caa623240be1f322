"""The C ABI's host logic on the CPU: csrc/capi.hip built host-only against a mock HIP runtime
(tests/host_mock: device memory is host memory, kernel launches do nothing and are counted).

  * ITOLBM_PAD: a stride pad that is negative or not a whole number is refused by lbm_create, and
    the state buffers' size does not depend on the pad;
  * ITOLBM_CPAD: likewise for the slot-pitch pad of the FP64 compact form, whose state buffers hold
    39 doubles per site of the padded pitch (either FP64 form fits them); unset, widths that make the
    pitch a multiple of 64 KiB are padded by 64 elements;
  * graph capture: while the context's stream is capturing (a caller's graph), every entry point
    that would allocate, free, copy from host memory or synchronise returns LBM_E_STATE and does none
    of it; stepping a prepared context queues kernel launches only, and needs an even step count.
"""
import ctypes as C
import os
import subprocess

import pytest

import _pkg

HERE = os.path.dirname(os.path.abspath(__file__))
MOCK_DIR = os.path.join(HERE, "host_mock")
MOCK_LIB = os.path.join(MOCK_DIR, "libitolbm_mock.so")

LBM_OK, LBM_E_ARG, LBM_E_STATE = 0, -1, -3
COUNT_NAMES = ("malloc", "free", "memcpy", "memset", "sync", "launches", "captures", "graph_launches")


@pytest.fixture(scope="module")
def mock():
    subprocess.run(["make", "-s", "-C", MOCK_DIR], check=True)
    L = C.CDLL(MOCK_LIB)
    P = C.c_void_p
    m = _pkg.load()
    L.lbm_create.argtypes = [C.POINTER(m.CaseDesc), C.c_int, C.c_int, C.POINTER(P)]
    L.lbm_destroy.argtypes = [P]
    L.lbm_last_error.argtypes = [P]
    L.lbm_last_error.restype = C.c_char_p
    L.lbm_device_bytes.argtypes = [P]
    L.lbm_device_bytes.restype = C.c_longlong
    L.lbm_step.argtypes = [P, C.c_int, C.c_int]
    L.lbm_step_case.argtypes = [P, C.c_int, C.c_int, C.c_int]
    L.lbm_set_active.argtypes = [P, C.c_int, C.c_int]
    L.lbm_init_drop.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double]
    L.lbm_init_band.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double]
    L.lbm_upload_state.argtypes = [P, C.c_int, P, P]
    L.lbm_download_state.argtypes = [P, C.c_int, C.c_int, P, P]
    L.lbm_download_aux.argtypes = [P, C.c_int, P, P]
    L.lbm_set_reference.argtypes = [P, C.c_int, C.c_double, C.c_double]
    L.lbm_sync.argtypes = [P]
    L.lbm_set_stream.argtypes = [P, P]
    L.lbm_diag_drop.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(m.DropDiag)]
    L.lbm_diag_incl.argtypes = [P, C.c_int, C.c_int, C.c_int, C.POINTER(m.InclDiag), P]
    L.mock_set_capturing.argtypes = [C.c_int]
    L.mock_get_counts.argtypes = [C.POINTER(C.c_longlong)]
    return L, m


def counts(L):
    a = (C.c_longlong * 8)()
    L.mock_get_counts(a)
    return dict(zip(COUNT_NAMES, a))


def create(L, m, cases, prec=0):
    arr = (m.CaseDesc * len(cases))(*cases)
    h = C.c_void_p()
    rc = L.lbm_create(arr, len(cases), prec, C.byref(h))
    return rc, h


@pytest.mark.parametrize("pad", ["-128", "-1", "abc", "12x", "", "4096", str(1 << 21)])
def test_stride_pad_checked(mock, pad, monkeypatch):
    L, m = mock
    monkeypatch.setenv("ITOLBM_PAD", pad)
    rc, h = create(L, m, [m.case_desc(40, 30, "P3 0.5")], prec=1)
    ok = pad in ("", "4096")
    assert rc == (LBM_OK if ok else LBM_E_ARG), (pad, rc)
    if ok:
        # two FP32 state buffers of 2 x 19 floats per site, whatever the pad
        assert L.lbm_device_bytes(h) == 2 * (4 * 2 * 19 * 40 * 30)
        assert L.lbm_destroy(h) == LBM_OK


@pytest.mark.parametrize("pad", ["-8", "abc", "3.5", "", "0", "64", "4097"])
def test_compact_pad_checked(mock, pad, monkeypatch):
    L, m = mock
    monkeypatch.setenv("ITOLBM_CPAD", pad)
    rc, h = create(L, m, [m.case_desc(40, 30, "P3 0.5")], prec=0)
    ok = pad in ("", "0", "64")
    assert rc == (LBM_OK if ok else LBM_E_ARG), (pad, rc)
    if ok:
        px = 40 + (int(pad) if pad else 0)
        # two FP64 state buffers of 39 doubles per site of the padded pitch
        assert L.lbm_device_bytes(h) == 2 * (8 * 39 * px * 30)
        assert L.lbm_destroy(h) == LBM_OK


@pytest.mark.parametrize("nx,pad", [(8192, 64), (16384, 64), (4096, 0), (8200, 0)])
def test_compact_pad_default(mock, nx, pad, monkeypatch):
    """Unset ITOLBM_CPAD: box widths that make the slot pitch a multiple of 64 KiB are padded by 64."""
    L, m = mock
    monkeypatch.delenv("ITOLBM_CPAD", raising=False)
    rc, h = create(L, m, [m.case_desc(nx, 2, "P3 0.5")], prec=0)
    assert rc == LBM_OK
    assert L.lbm_device_bytes(h) == 2 * (8 * 39 * (nx + pad) * 2)
    assert L.lbm_destroy(h) == LBM_OK


def test_capture_guard(mock):
    L, m = mock
    cases = [m.case_desc(40, 30, "P3 0.5"), m.case_desc(36, 28, "F2", 1.0, 0.1), m.case_desc(30, 30, "F1", 1.0, 0.1)]
    rc, h = create(L, m, cases)
    assert rc == LBM_OK
    for i in range(3):
        assert L.lbm_init_drop(h, i, 7.0, 3.45, 15, 14, 0.0) == LBM_OK
    # eager steps prepare every family (tables, aux) and one case's own table
    assert L.lbm_step(h, 3, 1) == LBM_OK
    assert L.lbm_step_case(h, 0, 2, 1) == LBM_OK

    L.mock_set_capturing(1)
    L.mock_reset_counts()
    assert L.lbm_step(h, 4, 0) == LBM_OK
    assert L.lbm_step(h, 2, 1) == LBM_OK                   # records into the aux fields allocated before
    assert L.lbm_step_case(h, 0, 2, 0) == LBM_OK
    c = counts(L)
    assert c["launches"] == 3 * 6 + 2, c                 # kernel launches only
    for k in ("malloc", "free", "memcpy", "memset", "sync", "captures", "graph_launches"):
        assert c[k] == 0, (k, c)

    # odd step counts, and everything that would allocate, free, copy or synchronise, are refused
    assert L.lbm_step(h, 3, 0) == LBM_E_ARG
    assert L.lbm_step_case(h, 0, 1, 0) == LBM_E_ARG
    assert L.lbm_step_case(h, 1, 2, 0) == LBM_E_STATE     # case 1's own table was never written
    buf = (C.c_double * (40 * 30 * 19))()
    buf2 = (C.c_double * (40 * 30 * 19))()
    refused = {
        "init_drop": L.lbm_init_drop(h, 0, 7.0, 3.45, 20, 15, 0.0),
        "init_band": L.lbm_init_band(h, 0, 2, 1, 3.45),
        "upload": L.lbm_upload_state(h, 0, buf, buf2),
        "download": L.lbm_download_state(h, 0, 1, buf, buf2),
        "download_aux": L.lbm_download_aux(h, 0, buf, None),
        "diag_drop": L.lbm_diag_drop(h, 0, 0.0, 1, 20.0, 15.0, C.byref(m.DropDiag())),
        "diag_incl": L.lbm_diag_incl(h, 0, 2, 1, C.byref(m.InclDiag()), None),
        "set_reference": L.lbm_set_reference(h, 0, 1.1, 0.9),
        "sync": L.lbm_sync(h),
        "set_stream": L.lbm_set_stream(h, None),
        "destroy": L.lbm_destroy(h),
    }
    assert all(v == LBM_E_STATE for v in refused.values()), refused
    assert b"capturing" in L.lbm_last_error(h)
    # leaving a family out needs nothing new; bringing a case back needs its family's table rewritten:
    # refused, and nothing allocated or copied
    assert L.lbm_set_active(h, 2, 0) == LBM_OK
    assert L.lbm_step(h, 2, 0) == LBM_OK
    assert L.lbm_set_active(h, 2, 1) == LBM_OK
    assert L.lbm_step(h, 2, 0) == LBM_E_STATE
    c = counts(L)
    for k in ("malloc", "free", "memcpy", "memset", "sync"):
        assert c[k] == 0, (k, c)

    # the same calls work once the capture has ended; whole graphs of the library's own again
    L.mock_set_capturing(0)
    assert L.lbm_step(h, 2, 0) == LBM_OK
    L.mock_reset_counts()
    assert L.lbm_step(h, 64, 0) == LBM_OK
    c = counts(L)
    assert c["captures"] == 1 and c["graph_launches"] == 2, c
    assert L.lbm_download_state(h, 0, 1, buf, buf2) == LBM_OK
    assert L.lbm_destroy(h) == LBM_OK


def test_first_recording_step_under_capture_refused(mock):
    L, m = mock
    rc, h = create(L, m, [m.case_desc(40, 30, "P1")])
    assert rc == LBM_OK
    assert L.lbm_step(h, 2, 0) == LBM_OK                  # table written, no aux fields yet
    L.mock_set_capturing(1)
    L.mock_reset_counts()
    assert L.lbm_step(h, 2, 1) == LBM_E_STATE             # would allocate the aux fields
    assert counts(L)["malloc"] == 0
    L.mock_set_capturing(0)
    assert L.lbm_destroy(h) == LBM_OK


def test_noalloc_step_touches_no_memory_api(mock):
    """lbm_step_case_noalloc: one table-fill launch plus one launch per step; no allocation, free,
    copy, memset or synchronisation, also under a caller's capture (even step counts there)."""
    L, m = mock
    L.lbm_step_case_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    rc, h = create(L, m, [m.case_desc(100, 80, "F1", 1.0, 0.1), m.case_desc(40, 30, "P3 0.5")], prec=3)
    assert rc == LBM_OK
    L.mock_reset_counts()
    assert L.lbm_step_case_noalloc(h, 0, 5, 0) == LBM_OK
    assert L.lbm_step_case_noalloc(h, 1, 3, 1) == LBM_E_STATE      # no aux fields yet
    c = counts(L)
    assert c["launches"] == 1 + 5, c
    for k in ("malloc", "free", "memcpy", "memset", "sync", "captures", "graph_launches"):
        assert c[k] == 0, (k, c)
    assert L.lbm_step_case(h, 1, 1, 1) == LBM_OK                     # allocates the aux fields
    L.mock_set_capturing(1)
    L.mock_reset_counts()
    assert L.lbm_step_case_noalloc(h, 1, 4, 1) == LBM_OK
    assert L.lbm_step_case_noalloc(h, 0, 3, 0) == LBM_E_ARG          # odd count under capture
    c = counts(L)
    assert c["launches"] == 1 + 4, c
    for k in ("malloc", "free", "memcpy", "memset", "sync", "captures", "graph_launches"):
        assert c[k] == 0, (k, c)
    L.mock_set_capturing(0)
    assert L.lbm_destroy(h) == LBM_OK
