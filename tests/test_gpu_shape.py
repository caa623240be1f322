"""GPU: a step's result does not depend on the tile shape or the pipeline depth.

The autotuner picks, per kernel family, the rows per wavefront segment and (FP32 / FP16 storage) the
number of rows whose loads are in flight, by timing real steps of the caller's; that is sound only if
every shape computes every site the same way.  Contexts with fixed shapes (ITOLBM_LY rows per segment, ITOLBM_DEPTH depth) are
compared bit for bit with one whose shapes the tuner chose.  The cases cover the four kernel families,
widths that are not a multiple of the 64-column strip (lanes outside the case) and heights that are
not a multiple of the segment (short last segments, look-ahead rows past a segment's end), and the
aux fields of a record step.
"""
import os
import sys

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu

CASES = [(96, 37, "P3 0.5", 1.0, 0.225), (70, 45, "F1", 0.8, 0.1), (64, 29, "F2", 1.0, 0.1),
         (33, 64, "F3 1.569988", 0.6, 0.1)]
SHAPES = [(1, 1), (2, 2), (3, 3), (5, 2), (7, 1), (64, 3), (13, 3)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(m, prec, shape, steps=4):
    saved = {k: os.environ.get(k) for k in ("ITOLBM_LY", "ITOLBM_DEPTH")}
    try:
        if shape:
            os.environ["ITOLBM_LY"], os.environ["ITOLBM_DEPTH"] = str(shape[0]), str(shape[1])
        else:
            os.environ.pop("ITOLBM_LY", None)
            os.environ.pop("ITOLBM_DEPTH", None)
        descs = [m.case_desc(nx, ny, op, tau, A) for nx, ny, op, tau, A in CASES]
        ctx = m.Context(descs, precision=prec)
        try:
            for i, d in enumerate(descs):
                ctx.init_drop(min(d.nx, d.ny) / 4, width=3.45, case=i)
            ctx.step(steps)
            ctx.step(1, record=True)
            out = []
            for i in range(len(CASES)):
                out += list(ctx.download(m.STATE_POST_COLLISION, case=i))
                out += list(ctx.aux(case=i))
            return out
        finally:
            ctx.close()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("prec", [0, 1, 2, 3, 4])
def test_result_independent_of_tile_shape(prec):
    m = _pkg.load()
    say(f"tile shapes: precision {prec}, tuned")
    ref = run(m, prec, None)
    for shape in SHAPES:
        say(f"tile shapes: precision {prec}, rows per segment {shape[0]}, depth {shape[1]}")
        got = run(m, prec, shape)
        for j, (x, y) in enumerate(zip(ref, got)):
            assert np.array_equal(x, y), f"shape {shape}, array {j}"


@pytest.mark.parametrize("prec", [0, 1, 3])
def test_online_tuning_keeps_results(prec):
    """A call long enough for the library to tune its tile shapes on the call's own steps (every
    family's candidate shapes timed in turn) ends in the same state as a fixed shape."""
    m = _pkg.load()
    say(f"online tuning: precision {prec}, 480 steps")
    ref = run(m, prec, (8, 2), steps=480)
    got = run(m, prec, None, steps=480)
    for j, (x, y) in enumerate(zip(ref, got)):
        assert np.array_equal(x, y), f"array {j}"
