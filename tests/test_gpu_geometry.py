"""GPU parity on boxes with solid sites (lbm_set_geometry): the FP64 site-parallel kernels against the
oracle's orc_step_geo, which is itself pinned bit for bit to the reference's functions on the same
geometry (tests/test_oracle.py::test_solid_sites_bitwise_vs_reference_build).

Populations (fluid sites; solid sites hold none and read 0), mediator momentum and force must be
IEEE-equal to the oracle's after every checked step, and the fluid sites' populations must hash to the
reference's own (golden "ref_geo": the reference stores only fluid sites, in site order).
"""
import hashlib
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

Q = 19


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def meio_of(nx, ny, solids):
    m = np.ones(nx * ny, np.int32)
    for x, y in solids:
        m[x + nx * y] = 0
    m[m != 0] = np.arange(1, int((m != 0).sum()) + 1)
    return m


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def assert_same(a, b, what):
    bad = np.flatnonzero(~(a == b))
    assert bad.size == 0, f"{what}: {bad.size} values differ, first {bad[:5]}, max |d| {np.abs(a - b).max()}"


def check(m, sim, ctx, fluid, case, what):
    gR, gB = ctx.download(m.STATE_POST_COLLISION, case=case)
    assert_same(gR.reshape(-1, Q)[fluid], sim.post_R.reshape(-1, Q)[fluid], f"{what}: red")
    assert_same(gB.reshape(-1, Q)[fluid], sim.post_B.reshape(-1, Q)[fluid], f"{what}: blue")
    assert not gR.reshape(-1, Q)[~fluid].any() and not gB.reshape(-1, Q)[~fluid].any(), f"{what}: solid sites"
    gm, F = ctx.aux(case=case)
    assert_same(gm, sim.gm, f"{what}: mediator momentum")
    if sim.F is not None:
        assert_same(F, sim.F, f"{what}: force")
    nR, nB = ctx.download(m.STATE_POST_STREAM, case=case)
    assert_same(nR, sim.fR, f"{what}: post-stream red")
    assert_same(nB, sim.fB, f"{what}: post-stream blue")
    return gR, gB


@pytest.mark.parametrize("idx", range(7))
def test_solid_sites_bitwise_vs_oracle_and_reference(golden, idx):
    ent = golden["ref_geo"][idx]
    a = ent["args"]
    nx, ny = int(a["nx"]), int(a["ny"])
    say(f"[gpu] solid sites {ent['label']}: {nx}x{ny}, wett {a['wett']}")
    m = _pkg.load()
    meio = meio_of(nx, ny, a["solid"])
    fluid = meio != 0
    cs = O.make_case(nx, ny, a["op"], float(a["tau"]), float(a["A"]), float(a["beta"]))
    sim = O.Sim(cs, solid=(meio == 0).astype(np.uint8), wett=float(a["wett"]))
    sim.init_drop(float(a["R"]), width=float(a["width"]))
    ctx = m.Context(m.case_desc(nx, ny, a["op"], float(a["tau"]), float(a["A"]), float(a["beta"])))
    ctx.set_geometry(meio, wett=float(a["wett"]))
    ctx.init_drop(float(a["R"]), width=float(a["width"]))
    dumps = [int(x) for x in str(a["dump"]).split(",")]
    for st in range(int(a["steps"])):
        sim.step()
        ctx.step(1, record=True)
        gR, gB = check(m, sim, ctx, fluid, 0, f"step {st}")
        if st in dumps:
            ref = ent["hash_nofma"][str(st)]
            assert h(gR.reshape(-1, Q)[fluid]) == ref["R"] and h(gB.reshape(-1, Q)[fluid]) == ref["B"], \
                f"{ent['label']}: populations at step {st} differ from the reference build's"
            if "F" in ref:
                _, F = ctx.aux()
                assert h(F.reshape(-1, 3)[fluid]) == ref["F"]
    ctx.close()


def test_solid_sites_batched_with_fluid_cases_and_graphs(golden):
    """A case with solids steps in the same lbm_step as all-fluid cases of the family kernels, with
    graph replay on, and a geometry change keeps the post-stream state."""
    m = _pkg.load()
    ent = golden["ref_geo"][5]   # F2
    a = ent["args"]
    nx, ny = int(a["nx"]), int(a["ny"])
    meio = meio_of(nx, ny, a["solid"])
    fluid = meio != 0
    descs = [("P3 0.5", 0.225, 40, 30), (a["op"], float(a["A"]), nx, ny), ("F1", 0.1, 33, 21)]
    say("[gpu] solid-site case batched with two all-fluid cases, graphs of 4 steps")
    ctx = m.Context([m.case_desc(x, y, op, 0.8, A, 0.8) for op, A, x, y in descs])
    ctx.set_graph(4)
    sims = []
    for i, (op, A, x, y) in enumerate(descs):
        cs = O.make_case(x, y, op, 0.8, A, 0.8)
        s = O.Sim(cs, solid=(meio == 0).astype(np.uint8) if i == 1 else None, wett=float(a["wett"]))
        s.init_drop(min(x, y) / 4.0, width=3.45)
        ctx.upload(s.fR, s.fB, case=i)
        sims.append(s)
    ctx.set_geometry(meio, wett=float(a["wett"]), case=1)   # after the upload: the state is kept
    for n in (9, 12):
        for s in sims:
            for _ in range(n):
                s.step()
        ctx.step(n, record=True)
        for i, s in enumerate(sims):
            fl = fluid if i == 1 else np.ones(s.cs.nx * s.cs.ny, bool)
            check(m, s, ctx, fl, i, f"case {i} after {n} more steps")
    # back to an all-fluid box: the post-stream state is kept (the former solid sites hold 0), and the
    # family kernel steps the case again
    before = ctx.download(m.STATE_POST_STREAM, case=1)
    ctx.set_geometry(None, case=1)
    after = ctx.download(m.STATE_POST_STREAM, case=1)
    assert_same(after[0], before[0], "post-stream red kept")
    assert_same(after[1], before[1], "post-stream blue kept")
    s2 = O.Sim(sims[1].cs)
    s2.init_drop(6.0, width=3.45)
    ctx.upload(s2.fR, s2.fB, case=1)
    for _ in range(3):
        s2.step()
    ctx.step(3, record=True)
    check(m, s2, ctx, np.ones(nx * ny, bool), 1, "all fluid again")
    ctx.close()


def test_channel_wall_wetting_long_run():
    """A drop against the solid walls of a 96 x 64 channel (two solid rows), P3 and F3, 300 steps,
    bit for bit the oracle; the wall wettability changes the drop's footprint on the wall."""
    m = _pkg.load()
    nx, ny = 96, 64
    solids = [[x, 0] for x in range(nx)] + [[x, 1] for x in range(nx)]
    meio = meio_of(nx, ny, solids)
    fluid = meio != 0
    out = {}
    for op, A in (("P3 0.5", 0.225), ("F3 1.569988", 0.1)):
        for wett in (0.6, -0.6):
            say(f"[gpu] channel 96x64 {op} wett {wett}, 300 steps")
            cs = O.make_case(nx, ny, op, 1.0, A, 0.8)
            sim = O.Sim(cs, solid=(meio == 0).astype(np.uint8), wett=wett)
            ctx = m.Context(m.case_desc(nx, ny, op, 1.0, A, 0.8))
            ctx.set_geometry(meio, wett=wett)
            ctx.init_drop(14.0, width=3.45, x0=48, y0=12)
            sim.init_drop(14.0, width=3.45, x0=48, y0=12)
            for _ in range(300):
                sim.step()
            ctx.step(300, record=True)
            gR, _ = check(m, sim, ctx, fluid, 0, f"{op} wett {wett} after 300 steps")
            out[(op, wett)] = gR.reshape(ny, nx, Q).sum(axis=2)[2]   # red density on the first fluid row
            ctx.close()
        # a wetting wall (wett > 0 attracts red) holds more red on its first row than a non-wetting one
        assert out[(op, 0.6)].sum() > out[(op, -0.6)].sum()


def test_solid_sites_refusals():
    m = _pkg.load()
    meio = meio_of(20, 16, [[3, 3]])
    ctx = m.Context(m.case_desc(20, 16, "P3 0.5", 1.0, 0.225, 0.8), precision=m.PREC_F32)
    with pytest.raises(m.LBMError):
        ctx.set_geometry(meio)
    ctx.close()
    ctx = m.Context(m.case_desc(20, 16, "P3 0.5", 1.0, 0.225, 0.8))
    ctx.set_geometry(meio, wett=0.2)
    ctx.init_drop(4.0)
    ctx.step(2, record=True)
    with pytest.raises(m.LBMError):
        ctx.diag_drop()
    with pytest.raises(m.LBMError):
        ctx.step_noalloc(2)
    with pytest.raises(m.LBMError):
        ctx.set_geometry(np.zeros(20 * 16, np.int32))   # no fluid site
    ctx.close()
