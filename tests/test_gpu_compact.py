"""GPU: the compact FP64 form of the P1/P3 and F3 cases against the oracle and against the full form.

In the compact form a site stores its 19 post-collision totals and the five scalars its recolouring
is formed from (cR, rho, the mediator gradient); the site that pulls a population forms its two colour
parts with the sender's own expressions.  A launch covering many sites runs its cases in the compact
form, a small one in the full form; ITOLBM_F64_COMPACT=1 / 0 fixes either.  These tests check, through
the C ABI:

- the bytes per lattice update each case reports (lbm_stored_bytes), and the form each launch size
  chooses;
- drops of every P/F3 operator in the compact form equal the oracle bit for bit, and so does the full
  form;
- cases moved between the forms by launches of different sizes step exactly as in one form;
- raw sites (near-empty sites the collision leaves untouched, Operadores_tensao.cpp:347) placed on
  strip edges, on the box's periodic seam and across segment boundaries, where the kernel falls back
  to per-population loads, give the oracle's bits;
- a batch of all seven operators steps to the same bits in both forms.
"""
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

OPS7 = [("P1", 0.225), ("P3 -0.4", 0.225), ("P3 0.5", 0.225), ("P3 2", 0.225), ("F1", 0.1), ("F2", 0.1),
        ("F3 1.569988", 0.1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_same(a, b, what):
    bad = np.flatnonzero(~(a == b))
    assert bad.size == 0, f"{what}: {bad.size} values differ, first {bad[:5]}, max |d| {np.abs(a - b).max()}"


def oracle_steps(m, sim, ctx, checkpoints):
    done = 0
    for target in checkpoints:
        n = target - done
        for _ in range(n):
            sim.step()
        ctx.step(n, record=True)
        done = target
        gR, gB = ctx.download(m.STATE_POST_COLLISION)
        assert_same(gR, sim.post_R, f"red after {target} steps")
        assert_same(gB, sim.post_B, f"blue after {target} steps")
        gm, F = ctx.aux()
        assert_same(gm, sim.gm, f"mediator momentum after {target} steps")
        if sim.F is not None:
            assert_same(F, sim.F, f"force after {target} steps")
        nR, nB = ctx.download(m.STATE_POST_STREAM)
        assert_same(nR, sim.fR, f"post-stream red after {target} steps")
        assert_same(nB, sim.fB, f"post-stream blue after {target} steps")


def test_stored_bytes(monkeypatch):
    m = _pkg.load()
    descs = [m.case_desc(40, 30, op, 1.0, A, 0.8) for op, A in OPS7]
    expect = {"P1": 384, "P3": 384, "F3": 384, "F1": 608, "F2": 608}
    ctx = m.Context(descs)   # small launches: the full form
    assert [ctx.stored_bytes(case=i) for i in range(len(descs))] == [608] * len(descs)
    ctx.close()
    monkeypatch.setenv("ITOLBM_F64_COMPACT", "1")
    ctx = m.Context(descs)
    got = [ctx.stored_bytes(case=i) for i in range(len(descs))]
    ctx.close()
    assert got == [expect[op.split()[0]] for op, _ in OPS7], got
    monkeypatch.setenv("ITOLBM_F64_COMPACT", "0")
    ctx = m.Context(descs)
    assert [ctx.stored_bytes(case=i) for i in range(len(descs))] == [608] * len(descs)
    ctx.close()
    monkeypatch.delenv("ITOLBM_F64_COMPACT")
    for prec, b in ((1, 304), (2, 152), (3, 224), (4, 112)):
        ctx = m.Context(descs[:1], precision=prec)
        assert ctx.stored_bytes() == b, (prec, ctx.stored_bytes())
        ctx.close()


@pytest.mark.parametrize("op,A,depth", [OPS7[2] + (None,), OPS7[6] + (None,), OPS7[4] + ("1",), OPS7[4] + ("2",),
                                        OPS7[5] + ("1",), OPS7[5] + ("2",)])
def test_full_form_bitwise(op, A, depth, monkeypatch):
    """The full FP64 form (38 populations stored per site), chosen with ITOLBM_F64_COMPACT=0; F1 / F2
    with the kernel that reads a row's populations again (depth 1) and the one that keeps them (2)."""
    say(f"[gpu] full FP64 form parity drop 70x45 {op} depth {depth}")
    monkeypatch.setenv("ITOLBM_F64_COMPACT", "0")
    if depth:
        monkeypatch.setenv("ITOLBM_LY", "8")
        monkeypatch.setenv("ITOLBM_DEPTH", depth)
    m = _pkg.load()
    sim = O.Sim(O.make_case(70, 45, op, 0.8, A, 0.8))
    ctx = m.Context(m.case_desc(70, 45, op, 0.8, A, 0.8))
    assert ctx.stored_bytes() == 608
    sim.init_drop(11.0, width=3.45)
    ctx.upload(sim.fR, sim.fB)
    oracle_steps(m, sim, ctx, [1, 2, 7, 20])
    ctx.close()


@pytest.mark.parametrize("op,A,depth", [OPS7[0] + (None,), OPS7[2] + (None,), OPS7[6] + (None,),
                                        OPS7[4] + ("2",), OPS7[5] + ("2",), OPS7[5] + ("1",)])
def test_compact_raw_sites_at_strip_edges(op, A, depth, monkeypatch):
    """3 x 3 blocks of sites scaled to a density of ~1e-33 (left uncollided, so stored raw) on the
    edge between two 64-column strips, on the periodic seam x = 0 / nx - 1, on a partial last strip
    and across 4-row segment boundaries; rows pulling from them take the per-population path.  F1 / F2
    (full form) with the kernel that keeps a row's totals for its collision (depth 2), which reads an
    uncollided site's populations again, and with the one that reads every row again (depth 1)."""
    say(f"[gpu] compact raw sites 130x40 {op} depth {depth}")
    monkeypatch.setenv("ITOLBM_LY", "4")
    monkeypatch.setenv("ITOLBM_F64_COMPACT", "1")
    if depth:
        monkeypatch.setenv("ITOLBM_DEPTH", depth)
    m = _pkg.load()
    nx, ny = 130, 40
    sim = O.Sim(O.make_case(nx, ny, op, 1.0, A, 0.8))
    ctx = m.Context(m.case_desc(nx, ny, op, 1.0, A, 0.8))
    assert ctx.stored_bytes() == (608 if op in ("F1", "F2") else 384)
    sim.init_drop(14.0, width=3.45)
    fR = sim.fR.reshape(ny, nx, 19)
    fB = sim.fB.reshape(ny, nx, 19)
    for y, x in [(20, 64), (8, 0), (33, 128), (16, 100), (4, 63)]:
        ys = [(y + d) % ny for d in (-1, 0, 1)]
        xs = [(x + d) % nx for d in (-1, 0, 1)]
        for yy in ys:
            for xx in xs:
                fR[yy, xx, :] *= 1e-33
                fB[yy, xx, :] *= 1e-33
    ctx.upload(sim.fR, sim.fB)
    oracle_steps(m, sim, ctx, [1, 2, 3, 6])
    ctx.close()


@pytest.mark.parametrize("op,A", [OPS7[0], OPS7[1], OPS7[2], OPS7[3], OPS7[6]])
def test_compact_drop_bitwise(op, A, monkeypatch):
    """A diffuse drop and a sharp translating one in the compact form, against the oracle."""
    say(f"[gpu] compact form parity {op}")
    monkeypatch.setenv("ITOLBM_F64_COMPACT", "1")
    m = _pkg.load()
    for nx, ny, R, w, v0, steps in ((70, 45, 11.0, 3.45, 0.0, [1, 2, 7, 20]), (131, 33, 9.0, -1.0, 0.05, [1, 10])):
        sim = O.Sim(O.make_case(nx, ny, op, 0.8, A, 0.8))
        ctx = m.Context(m.case_desc(nx, ny, op, 0.8, A, 0.8))
        assert ctx.stored_bytes() == 384
        sim.init_drop(R, width=w, v0=v0)
        ctx.upload(sim.fR, sim.fB)
        oracle_steps(m, sim, ctx, steps)
        ctx.close()


def test_form_follows_launch_size(monkeypatch):
    """Eight 192^2 cases (295 k sites per lbm_step launch: the compact form) stepped by lbm_step, then
    one of them by lbm_step_case (a 37 k-site launch: the full form, converted on the device), then
    all by lbm_step again: the same bits as the same steps in the full form throughout."""
    say("[gpu] storage form by launch size, 8 x 192^2")
    m = _pkg.load()
    descs = [m.case_desc(192, 192, "P3 0.5", tau, 0.225, 0.8) for tau in (0.6, 1.0) * 4]
    out = {}
    for form in (None, "0"):
        if form is None:
            monkeypatch.delenv("ITOLBM_F64_COMPACT", raising=False)
        else:
            monkeypatch.setenv("ITOLBM_F64_COMPACT", form)
        ctx = m.Context(descs)
        for i in range(len(descs)):
            ctx.init_drop(40.0, width=-1.0, v0=0.02, case=i)
        ctx.step(12)
        b1 = [ctx.stored_bytes(case=i) for i in range(len(descs))]
        ctx.step(9, case=3)
        b2 = ctx.stored_bytes(case=3)
        ctx.step(6, record=True)
        b3 = ctx.stored_bytes(case=3)
        if form is None:
            assert b1 == [384] * 8 and b2 == 608 and b3 == 384, (b1, b2, b3)
        else:
            assert b1 == [608] * 8 and b2 == 608 and b3 == 608
        out[form] = [ctx.download(m.STATE_POST_COLLISION, case=i) for i in range(len(descs))]
        ctx.close()
    for i in range(len(descs)):
        assert_same(out[None][i][0], out["0"][i][0], f"case {i} red")
        assert_same(out[None][i][1], out["0"][i][1], f"case {i} blue")


def test_compact_equals_full_batch(monkeypatch):
    """All seven operators at two tau on uneven boxes, batched: 60 steps in the compact form and in
    the full form give the same populations and the same drop diagnostics."""
    say("[gpu] compact vs full FP64 form, 14-case batch")
    m = _pkg.load()
    descs, sizes = [], [(96, 70), (130, 33)]
    for k, (op, A) in enumerate(OPS7 * 2):
        nx, ny = sizes[k % 2]
        descs.append(m.case_desc(nx, ny, op, 0.6 if k < 7 else 1.0, A, 0.8))
    out = {}
    for form in ("1", "0"):
        monkeypatch.setenv("ITOLBM_F64_COMPACT", form)
        ctx = m.Context(descs)
        for i in range(len(descs)):
            ctx.init_drop(12.0, width=-1.0, v0=0.03 * (i % 3), case=i)
        ctx.step(60, record=True)
        states = [ctx.download(m.STATE_POST_COLLISION, case=i) for i in range(len(descs))]
        diags = [ctx.diag_drop(0.0, 60, case=i).delta_p for i in range(len(descs))]
        out[form] = (states, diags)
        ctx.close()
    for i in range(len(descs)):
        assert_same(out["1"][0][i][0], out["0"][0][i][0], f"case {i} red")
        assert_same(out["1"][0][i][1], out["0"][0][i][1], f"case {i} blue")
    assert out["1"][1] == out["0"][1]


def test_compact_graph_replay_and_form_switch(monkeypatch):
    """Eight 192^2 cases with the library's step graphs (4 steps per graph): runs of lbm_step (the six
    P1/P3 cases' launch, 221 k sites, in the compact form, the two F3 cases' launch, 74 k sites, in the
    full form, replayed from graphs), lbm_step_case on one case (the full form: converted on the
    device, the graph dropped), lbm_step again (converted back, the graph captured again), recording
    aux on some calls; the same bits as eager steps in the full form after every call."""
    say("[gpu] compact form under graph replay with form switches, 8 x 192^2")
    m = _pkg.load()
    descs = [m.case_desc(192, 192, op, tau, A, 0.8) for (op, A) in (OPS7[0], OPS7[2], OPS7[6], OPS7[3])
             for tau in (0.6, 1.0)]

    def make(form, graph):
        if form is None:
            monkeypatch.delenv("ITOLBM_F64_COMPACT", raising=False)
        else:
            monkeypatch.setenv("ITOLBM_F64_COMPACT", form)
        ctx = m.Context(descs)
        ctx.set_graph(graph)
        for i in range(len(descs)):
            ctx.init_drop(40.0, width=-1.0, v0=0.02 * (i % 3), case=i)
        return ctx

    ops = [op for (op, A) in (OPS7[0], OPS7[2], OPS7[6], OPS7[3]) for tau in (0.6, 1.0)]
    family_bytes = [608 if op.startswith("F3") else 384 for op in ops]
    a = make(None, 4)    # form chosen by launch size, graphs
    b = make("0", 0)     # full form, eager
    try:
        plan = [(9, None, False), (12, None, True), (6, 2, False), (8, None, False), (4, 5, True), (10, None, True)]
        for call, (n, case, rec) in enumerate(plan):
            a.step(n, record=rec, case=case)
            b.step(n, record=rec, case=case)
            if case is None:
                assert [a.stored_bytes(case=i) for i in range(len(descs))] == family_bytes, call
            else:
                assert a.stored_bytes(case=case) == 608, call
            for i in range(len(descs)):
                ra, ba = a.download(m.STATE_POST_COLLISION, case=i)
                rb, bb = b.download(m.STATE_POST_COLLISION, case=i)
                assert_same(ra, rb, f"call {call} case {i} red")
                assert_same(ba, bb, f"call {call} case {i} blue")
            if rec:
                k = case if case is not None else 0
                ga, Fa = a.aux(case=k)
                gb, Fb = b.aux(case=k)
                assert_same(ga, gb, f"call {call} mediator momentum")
                assert_same(Fa, Fb, f"call {call} force")
    finally:
        a.close()
        b.close()
