"""lbm_step's launch forms for FP64 contexts: one launch per family (0), one launch of every family's
tiles (1) and that launch persistent (2) give the same bits, with and without graph replay, for a
batch of uneven cases of all seven operators; and the tuner settles on one of them."""
import os
import sys

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu

CASES = [(70, 45, "P1", 0.225), (33, 21, "P3 -0.4", 0.225), (131, 33, "P3 0.5", 0.225), (64, 64, "P3 2", 0.225),
         (40, 30, "F1", 0.1), (100, 20, "F2", 0.1), (20, 100, "F3 1.569988", 0.1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(m, mode, steps, graph):
    old = os.environ.get("ITOLBM_FUSED")
    os.environ["ITOLBM_FUSED"] = str(mode)
    try:
        ctx = m.Context([m.case_desc(nx, ny, op, 0.8, A, 0.8) for nx, ny, op, A in CASES])
        ctx.set_graph(graph)
        for i, (nx, ny, _, _) in enumerate(CASES):
            ctx.init_drop(min(nx, ny) / 4.0, width=3.45, case=i)
        ctx.step(steps, record=True)
        got = mode, ctx.step_mode()
        out = []
        for i in range(len(CASES)):
            out.append(ctx.download(m.STATE_POST_COLLISION, case=i) + ctx.aux(case=i))
        ctx.close()
    finally:
        if old is None:
            os.environ.pop("ITOLBM_FUSED", None)
        else:
            os.environ["ITOLBM_FUSED"] = old
    assert got[1] == got[0], got
    return out


@pytest.mark.parametrize("graph", [0, 8])
def test_launch_forms_give_the_same_bits(graph):
    m = _pkg.load()
    ref = None
    for mode in (0, 1, 2):
        say(f"[gpu] 7 uneven FP64 cases, launch form {mode}, graphs of {graph} steps, 27 steps")
        out = run(m, mode, 27, graph)
        if ref is None:
            ref = out
            continue
        for i, (a, b) in enumerate(zip(out, ref)):
            for j, (x, y) in enumerate(zip(a, b)):
                assert np.array_equal(x, y), f"launch form {mode}, case {i}, array {j}"


def test_tuner_chooses_a_launch_form():
    m = _pkg.load()
    say("[gpu] launch form chosen by timing (600 steps, 7 uneven FP64 cases)")
    os.environ.pop("ITOLBM_FUSED", None)
    ctx = m.Context([m.case_desc(nx, ny, op, 0.8, A, 0.8) for nx, ny, op, A in CASES])
    assert ctx.step_mode() == -1
    for i, (nx, ny, _, _) in enumerate(CASES):
        ctx.init_drop(min(nx, ny) / 4.0, width=3.45, case=i)
    ctx.step(600)
    assert ctx.step_mode() in (0, 1, 2)
    ctx.set_active(False, case=3)   # a change of the active set chooses again
    assert ctx.step_mode() == -1
    ctx.close()
    lp = m.Context(m.case_desc(64, 64, "P3 0.5", 1.0, 0.225, 0.8), precision=m.PREC_F32)
    lp.init_drop(16.0)
    lp.step(200)
    assert lp.step_mode() == 0   # FP32 / FP16: one launch per family
    lp.close()
