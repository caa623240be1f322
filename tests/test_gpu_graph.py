"""GPU: steps replayed from a captured HIP graph equal the same steps launched one by one.

lbm_step queues whole runs of steps_per_graph steps as one graph replay.  Two contexts holding the
same cases, one with graphs (4 steps per graph) and one without, are advanced by the same sequence
of step counts; after every call the states are compared bit for bit.  The sequence mixes graph
replays with eager steps and odd counts (the family parity changes, so the graph is captured again),
records aux on some calls, and changes the active set and (FP32 / FP16) the storage reference in the
middle, which also recaptures.  Three kernel families run on forked streams inside the graph.
"""
import sys

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu

CASES = [(48, 40, "P3 0.5", 1.0, 0.225), (70, 45, "F1", 0.8, 0.1), (33, 64, "F3 1.569988", 1.0, 0.1),
         (64, 64, "P1", 0.6, 0.375)]
SEQ = [10, 7, 9, 4, 1, 12, 8]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(m, prec, graph):
    descs = [m.case_desc(nx, ny, op, tau, A) for nx, ny, op, tau, A in CASES]
    ctx = m.Context(descs, precision=prec)
    ctx.set_graph(graph)
    for i, d in enumerate(descs):
        ctx.init_drop(min(d.nx, d.ny) / 4, width=3.45, case=i)
    return ctx


def states(ctx):
    m = _pkg.load()
    return [ctx.download(m.STATE_POST_COLLISION, case=i) for i in range(len(CASES))]


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_graph_replay_equals_eager(prec):
    m = _pkg.load()
    say(f"graph vs eager: precision {prec}")
    a, b = make(m, prec, 0), make(m, prec, 4)
    try:
        for call, n in enumerate(SEQ):
            rec = call % 2 == 1
            if call == 3:   # a member leaves the batch: its family's table is rewritten
                a.set_active(False, case=3)
                b.set_active(False, case=3)
            if call == 5 and prec:
                a.set_reference(1.01, 0.99, case=0)
                b.set_reference(1.01, 0.99, case=0)
            a.step(n, record=rec)
            b.step(n, record=rec)
            for i, (sa, sb) in enumerate(zip(states(a), states(b))):
                for x, y in zip(sa, sb):
                    assert np.array_equal(x, y), f"call {call} ({n} steps), case {i}"
            if rec:
                da = a.diag_drop(0.0, 0, case=0)
                db = b.diag_drop(0.0, 0, case=0)
                assert da.delta_p == db.delta_p and da.raio == db.raio
    finally:
        a.close()
        b.close()


def test_graph_steps_must_be_even():
    m = _pkg.load()
    ctx = m.Context([m.case_desc(32, 32, "P1", 1.0, 0.225)], precision=0)
    try:
        with pytest.raises(m.LBMError):
            ctx.set_graph(3)
        ctx.set_graph(0)
        ctx.set_graph(2)
    finally:
        ctx.close()
