"""GPU: batched launches and the reduced-precision forms, through the C ABI.

Batched (many cases per launch) FP64 must equal the oracle bit for bit, like the one-case path.

The FP32 and FP16 forms compute in FP32 (colour sums in FP64) and store, per site and direction, the
pair (h_i, e_i) with e_i = (f_R,i - f_B,i) / (f_R,i + f_B,i) and h_i the deviation
s_i = f_R,i + f_B,i - w_i scaled by a power of two near 1 / w_i, less the case's reference density,
in FP32 or FP16.  They are compared with the FP64 path (itself equal to the oracle, see
test_gpu_parity.py) at the tolerances written in each test:

* populations after 20 steps (70 x 45 drop, width 3.45): max |df| <= 5e-7 (FP32), 1e-4 (FP16);
* Laplace delta_p after 1000 steps (128^2 sharp drop R = 25.6): relative <= 2e-4 (FP32), 3e-2 (FP16);
  radius <= 1e-5 (FP32), 1e-3 (FP16);
* colour masses drift by <= 1e-5 (FP32) / 5e-4 (FP16) relative over 1000 steps (exact in FP64);
* a pure phase stays pure: e_i = +-1 is stored exactly and the colour-weighted sums use the
  same order as the plain sums, so f_B = 0 inside a sharp drop is kept exactly;
* random non-equilibrium states (no symmetry in z), 3 steps: max |df| <= 5e-7 (FP32), 2e-4 (FP16).
"""
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

OPS7 = [("P1", 0.225), ("P3 -0.4", 0.225), ("P3 0.5", 0.225), ("P3 2", 0.225), ("F1", 0.1), ("F2", 0.1),
        ("F3 1.569988", 0.1)]
POP_TOL = {1: 5e-7, 2: 1e-4}
DP_RTOL = {1: 2e-4, 2: 3e-2}
R_RTOL = {1: 1e-5, 2: 1e-3}
MASS_RTOL = {1: 1e-5, 2: 5e-4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_batched_fp64_uneven_cases_match_oracle():
    """Six cases of different sizes and operators (two per family for P) advanced by one lbm_step."""
    say("[gpu] batched FP64: 6 uneven cases, all families")
    m = _pkg.load()
    spec = [(70, 45, "P1", 0.225, 0.8), (50, 36, "P3 2", 0.3, 1.0), (64, 40, "F2", 0.1, 0.6),
            (33, 50, "F1", 0.05, 1.0), (40, 40, "F3 1.569988", 0.1, 0.8), (129, 17, "P3 -0.4", 0.225, 1.5)]
    sims, descs = [], []
    for k, (nx, ny, op, A, tau) in enumerate(spec):
        s = O.Sim(O.make_case(nx, ny, op, tau, A, 0.8))
        s.init_drop(min(nx, ny) * 0.3, width=3.45, v0=0.02 * (k % 2))
        sims.append(s)
        descs.append(m.case_desc(nx, ny, op, tau, A, 0.8))
    ctx = m.Context(descs)
    for k, s in enumerate(sims):
        ctx.upload(s.fR, s.fB, case=k)
    done = 0
    for target in (1, 6):
        for s in sims:
            for _ in range(target - done):
                s.step()
        ctx.step(target - done, record=True)
        done = target
        for k, s in enumerate(sims):
            gR, gB = ctx.download(m.STATE_POST_COLLISION, case=k)
            assert np.array_equal(gR, s.post_R) and np.array_equal(gB, s.post_B), (spec[k], target)
            gm, F = ctx.aux(case=k)
            assert np.array_equal(gm, s.gm), (spec[k], target)
    ctx.close()


def test_mixed_single_and_batched_steps():
    """lbm_step_case on one member then lbm_step on all: the group tables follow each case's buffer."""
    say("[gpu] mixed single/batched FP64 steps")
    m = _pkg.load()
    spec = [(48, 40, "P3 0.5", 0.225), (44, 52, "P1", 0.225)]
    sims = [O.Sim(O.make_case(nx, ny, op, 1.0, A, 0.8)) for nx, ny, op, A in spec]
    ctx = m.Context([m.case_desc(nx, ny, op, 1.0, A, 0.8) for nx, ny, op, A in spec])
    for k, s in enumerate(sims):
        s.init_drop(12.0)
        ctx.upload(s.fR, s.fB, case=k)
    ctx.step(3, case=0)
    for _ in range(3):
        sims[0].step()
    ctx.step(2)
    for s in sims:
        for _ in range(2):
            s.step()
    for k, s in enumerate(sims):
        gR, gB = ctx.download(m.STATE_POST_COLLISION, case=k)
        assert np.array_equal(gR, s.post_R) and np.array_equal(gB, s.post_B), spec[k]
    ctx.close()


@pytest.mark.parametrize("prec", [1, 2], ids=["f32", "f16"])
@pytest.mark.parametrize("op,A", OPS7)
def test_lp_step_close_to_oracle(prec, op, A):
    say(f"[gpu] precision {prec} step 70x45 {op}")
    m = _pkg.load()
    sim = O.Sim(O.make_case(70, 45, op, 0.8, A, 0.8))
    sim.init_drop(11.0, width=3.45, v0=0.03)
    ctx = m.Context(m.case_desc(70, 45, op, 0.8, A, 0.8), precision=prec)
    ctx.upload(sim.fR, sim.fB)
    for _ in range(20):
        sim.step()
    ctx.step(20, record=True)
    gR, gB = ctx.download(m.STATE_POST_COLLISION)
    err = max(np.abs(gR - sim.post_R).max(), np.abs(gB - sim.post_B).max())
    say(f"    max |df| = {err:.3e}")
    assert err <= POP_TOL[prec]
    gm, F = ctx.aux()
    assert np.abs(gm - sim.gm).max() <= 50 * POP_TOL[prec]
    if sim.F is not None:
        assert np.abs(F - sim.F).max() <= 50 * POP_TOL[prec] * max(1.0, np.abs(sim.F).max())
    nR, nB = ctx.download(m.STATE_POST_STREAM)
    assert np.abs(nR - sim.fR).max() <= POP_TOL[prec]
    ctx.close()


@pytest.mark.parametrize("prec", [1, 2], ids=["f32", "f16"])
@pytest.mark.parametrize("op,A", [OPS7[2], OPS7[4]])
def test_lp_wide_case_close_to_oracle(prec, op, A, monkeypatch):
    """A case 4096 wide (64 strips, a multiple of 32: its tiles are dealt in bands of 32 strips) and
    24 rows in 8-row segments, a drop on the band edge x = 2048: 5 steps against the oracle."""
    say(f"[gpu] precision {prec} step 4096x24 {op}")
    monkeypatch.setenv("ITOLBM_LY", "8")
    m = _pkg.load()
    nx, ny = 4096, 24
    sim = O.Sim(O.make_case(nx, ny, op, 0.8, A, 0.8))
    sim.init_drop(9.0, width=3.45, x0=2048, y0=12, v0=0.03)
    ctx = m.Context(m.case_desc(nx, ny, op, 0.8, A, 0.8), precision=prec)
    ctx.upload(sim.fR, sim.fB)
    for _ in range(5):
        sim.step()
    ctx.step(5)
    gR, gB = ctx.download(m.STATE_POST_COLLISION)
    err = max(np.abs(gR - sim.post_R).max(), np.abs(gB - sim.post_B).max())
    say(f"    max |df| = {err:.3e}")
    assert err <= POP_TOL[prec]
    ctx.close()


RANDOM_BOXES = [(5, 3, "F2", 0.1), (7, 1, "F1", 0.1), (6, 2, "F2", 0.1), (9, 2, "P3 0.5", 0.225),
                (4, 1, "F3 1.569988", 0.1), (3, 4, "F1", 0.1), (64, 40, "F2", 0.1), (70, 45, "P3 0.5", 0.225)]
RANDOM_TOL = {1: 5e-7, 2: 2e-4}


@pytest.mark.parametrize("prec", [1, 2], ids=["f32", "f16"])
@pytest.mark.parametrize("nx,ny,op,A", RANDOM_BOXES)
def test_lp_random_states(prec, nx, ny, op, A):
    """Populations within 10 % of the weights, independently random per direction, with a random colour:
    no symmetry in z (momentum in z), strong gradients everywhere.  Boxes narrower than a strip and
    shorter than the stencils (rows wrap more than once) and ordinary ones.  3 steps against the oracle:
    max |df| <= 5e-7 (FP32), 2e-4 (FP16)."""
    say(f"[gpu] precision {prec} random state {nx}x{ny} {op}")
    m = _pkg.load()
    sim = O.Sim(O.make_case(nx, ny, op, 0.8, A, 0.8))
    rng = np.random.default_rng(7)
    w = np.tile(np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12), nx * ny)
    f = w * rng.uniform(0.9, 1.1, w.size)
    c = rng.uniform(0.0, 1.0, w.size)
    sim.fR[:] = f * c
    sim.fB[:] = f * (1.0 - c)
    ctx = m.Context(m.case_desc(nx, ny, op, 0.8, A, 0.8), precision=prec)
    ctx.upload(sim.fR, sim.fB)
    for _ in range(3):
        sim.step()
    ctx.step(3)
    gR, gB = ctx.download(m.STATE_POST_COLLISION)
    err = max(np.abs(gR - sim.post_R).max(), np.abs(gB - sim.post_B).max())
    say(f"    max |df| = {err:.3e}")
    assert err <= RANDOM_TOL[prec]
    ctx.close()


@pytest.mark.parametrize("prec", [1, 2], ids=["f32", "f16"])
def test_lp_pure_phase_stays_pure(prec):
    say(f"[gpu] precision {prec} pure phase")
    m = _pkg.load()
    ctx = m.Context(m.case_desc(96, 96, "P3 0.5", 1.0, 0.225, 0.8), precision=prec)
    ctx.init_drop(24.0, width=-1.0)
    ctx.step(10)
    fR, fB = ctx.download(m.STATE_POST_COLLISION)
    fR, fB = fR.reshape(96, 96, 19), fB.reshape(96, 96, 19)
    y, x = np.mgrid[0:96, 0:96]
    r = np.hypot(x - 48, y - 48)
    inside, outside = r < 12, r > 40
    assert np.all(fB[inside] == 0.0) and np.all(fR[outside] == 0.0)
    ctx.close()


def test_lp_upload_roundtrip_f32():
    """Upload stores (s/w, e) in FP32; reading back reproduces the populations to FP32 resolution."""
    say("[gpu] precision 1 upload round trip")
    m = _pkg.load()
    rng = np.random.default_rng(5)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    n = 37 * 29
    rho = rng.uniform(0.9, 1.1, (n, 1))
    c = rng.uniform(0, 1, (n, 1))
    f = w * rho * (1 + rng.uniform(-0.05, 0.05, (n, 19)))
    fR, fB = (c * f).reshape(-1), ((1 - c) * f).reshape(-1)
    ctx = m.Context(m.case_desc(37, 29, "P3 0.5", 1.0, 0.225), precision=1)
    ctx.upload(fR, fB)
    gR, gB = ctx.download(m.STATE_POST_STREAM)
    assert np.abs(gR - fR).max() <= 1e-7 and np.abs(gB - fB).max() <= 1e-7
    ctx.close()


@pytest.mark.parametrize("prec", [1, 2], ids=["f32", "f16"])
def test_lp_laplace_vs_fp64(prec):
    """128^2 sharp drop R = 25.6, tau = 1, sigma = 0.1, all seven operators in one batched context per
    precision, 1000 steps: delta_p, radius and colour mass against the FP64 path."""
    say(f"[gpu] precision {prec} Laplace 128^2 x 7 operators, 1000 steps")
    m = _pkg.load()
    descs = [m.case_desc(128, 128, op, 1.0, A, 0.8) for op, A in OPS7]
    res = {}
    for p in (0, prec):
        ctx = m.Context(descs, precision=p)
        for k in range(len(descs)):
            ctx.init_drop(25.6, width=-1.0, case=k)
        ctx.step(1, record=True)
        d0 = [ctx.diag_drop(0.0, 0, case=k) for k in range(len(descs))]
        ctx.step(499, record=True)
        if p:
            # re-centre the deviation storage on the measured phase densities, as a driver does
            for k in range(len(descs)):
                d = ctx.diag_drop(0.0, 499, case=k)
                ctx.set_reference(3 * d.p_in, 3 * d.p_out, case=k)
        ctx.step(500, record=True)
        res[p] = (d0, [ctx.diag_drop(0.0, 999, case=k) for k in range(len(descs))])
        ctx.close()
    for k, (op, _) in enumerate(OPS7):
        a, b = res[prec][1][k], res[0][1][k]
        edp = abs(a.delta_p - b.delta_p) / abs(b.delta_p)
        er = abs(a.raio - b.raio) / b.raio
        em = abs(a.massa_R - res[prec][0][k].massa_R) / res[prec][0][k].massa_R
        say(f"    {op:12s} dp {edp:.2e}  R {er:.2e}  massR drift {em:.2e}")
        assert edp <= DP_RTOL[prec], (op, edp)
        assert er <= R_RTOL[prec], (op, er)
        assert em <= MASS_RTOL[prec], (op, em)
        assert a.ok == b.ok and a.n_ruim == 0


@pytest.mark.parametrize("prec", [1, 2], ids=["f32", "f16"])
def test_lp_batched_equals_single(prec):
    say(f"[gpu] precision {prec} batched vs single launch")
    m = _pkg.load()
    descs = [m.case_desc(80, 64, "F2", 1.0, 0.1), m.case_desc(70, 90, "P3 0.5", 1.0, 0.225),
             m.case_desc(65, 65, "F3 1.569988", 1.0, 0.1), m.case_desc(100, 30, "F1", 1.0, 0.1)]
    a = m.Context(descs, precision=prec)
    b = m.Context(descs, precision=prec)
    for k in range(len(descs)):
        a.init_drop(14.0, case=k)
        b.init_drop(14.0, case=k)
    a.step(7)
    for k in range(len(descs)):
        b.step(7, case=k)
    for k in range(len(descs)):
        ra, _ = a.download(m.STATE_POST_COLLISION, case=k)
        rb, _ = b.download(m.STATE_POST_COLLISION, case=k)
        assert np.array_equal(ra, rb), k
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["f64", "f32", "f16"])
def test_set_reference_keeps_state(prec):
    """Re-expressing the stored deviations against new reference densities changes the populations by
    rounding only (FP64: not at all)."""
    say(f"[gpu] precision {prec} set_reference")
    m = _pkg.load()
    ctx = m.Context(m.case_desc(64, 48, "F2", 1.0, 0.1), precision=prec)
    ctx.init_drop(13.0, width=3.45, v0=0.05)
    ctx.step(5)
    a = ctx.download(m.STATE_POST_COLLISION)
    ctx.set_reference(1.012, 0.994)
    b = ctx.download(m.STATE_POST_COLLISION)
    tol = {0: 0.0, 1: 1e-7, 2: 1e-5}[prec]
    assert np.abs(a[0] - b[0]).max() <= tol and np.abs(a[1] - b[1]).max() <= tol
    ctx.step(3)   # steps after a change of reference run from the re-expressed state
    c = ctx.download(m.STATE_POST_COLLISION)
    assert np.all(np.isfinite(c[0]))
    ctx.close()
