"""lbm_step_case_noalloc, the allocation-free variant of lbm_step_case, through the C ABI.

Its results must equal lbm_step_case's bit for bit (same kernels, same tile shape; the case table
and tile map sit in the library's static device scratch instead of allocated memory), on two calls
in a row, the second with a smaller case, in FP64 and in FP32 z-symmetric storage.  The context's
device-byte count must not change across each call.
"""
import sys

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu

LBM_E_STATE = -3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(m, prec):
    cases = [m.case_desc(200, 150, "F2", 1.0, 0.1), m.case_desc(70, 45, "P3 0.5", 0.8, 0.225)]
    ctx = m.Context(cases, precision=prec)
    ctx.init_drop(37.0, width=3.45, v0=0.03, case=0)
    ctx.init_drop(11.0, width=-1.0, case=1)
    return ctx


def same_state(a, b, case, what):
    m = _pkg.load()
    ra, ba = a.download(m.STATE_POST_COLLISION, case=case)
    rb, bb = b.download(m.STATE_POST_COLLISION, case=case)
    assert np.array_equal(ra, rb), f"{what}: red differs, max |d| {np.abs(ra - rb).max()}"
    assert np.array_equal(ba, bb), f"{what}: blue differs, max |d| {np.abs(ba - bb).max()}"


@pytest.mark.parametrize("prec,compact", [(0, None), (0, "1"), (3, None)], ids=["f64", "f64-compact", "f32z"])
def test_noalloc_step_equals_step_case(prec, compact, monkeypatch):
    """f64-compact: the FP64 P1/P3 case held in the compact form (ITOLBM_F64_COMPACT=1), so the
    allocation-free step runs the compact kernel too (the F2 case keeps the full form)."""
    if compact:
        monkeypatch.setenv("ITOLBM_F64_COMPACT", compact)
    m = _pkg.load()
    say(f"[gpu] lbm_step_case_noalloc vs lbm_step_case, precision {prec}, compact {compact}: 200x150 F2 then 70x45 P3")
    ref, na = make(m, prec), make(m, prec)
    if compact:
        assert na.stored_bytes(case=1) == 384 and na.stored_bytes(case=0) == 608
    def noalloc(n, case, record=False):
        before = na.device_bytes
        na.step_noalloc(n, record=record, case=case)
        assert na.device_bytes == before, "the allocation-free step allocated device memory"

    # first call: the larger case; second call: the smaller one
    ref.step(7, case=0)
    noalloc(7, 0)
    same_state(ref, na, 0, "200x150 after 7 steps")
    ref.step(5, case=1)
    noalloc(5, 1)
    same_state(ref, na, 1, "70x45 after 5 steps")
    # a recording step needs the aux fields: refused until a recording lbm_step_case has made them
    with pytest.raises(m.LBMError) as e:
        na.step_noalloc(2, record=True, case=0)
    assert f"{LBM_E_STATE}" in str(e.value)
    ref.step(1, record=True, case=0)
    na.step(1, record=True, case=0)
    ref.step(4, record=True, case=0)
    noalloc(4, 0, record=True)
    same_state(ref, na, 0, "200x150 after 12 steps, recorded")
    g0, F0 = ref.aux(case=0)
    g1, F1 = na.aux(case=0)
    assert np.array_equal(g0, g1) and np.array_equal(F0, F1), "aux fields differ"
    # lbm_step_case and lbm_step continue from the state the allocation-free steps left (odd count)
    ref.step(3, case=1)
    noalloc(3, 1)
    ref.step(2, case=1)
    na.step(2, case=1)
    same_state(ref, na, 1, "70x45 after 10 steps, mixed entry points")
    ref.step(2)
    na.step(2)
    same_state(ref, na, 0, "200x150 after lbm_step")
    same_state(ref, na, 1, "70x45 after lbm_step")
    ref.close()
    na.close()
