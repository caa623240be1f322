"""GPU: the z-symmetric storage forms (LBM_PREC_F32_ZSYM, LBM_PREC_F16_ZSYM) against the full ones.

With one node in z, the z mirror of a velocity pulls from the same site, and every state the programs
create is mirror-symmetric in z; the FP32/FP16 collision keeps that symmetry exactly (its z moment is a
sum of mirror-pair differences, 0 in a symmetric state).  The z-symmetric forms store the 14 distinct
populations and alias the mirrors on load, so on a symmetric state they must give the same bits as the
full forms: populations, mediator gradient and force after every checked step, and the diagnostics.
A state that is not symmetric cannot be held and is rejected on upload.
"""
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

CASES = [(96, 80, "P1", 0.6, 0.375), (96, 80, "P3 -0.4", 1.0, 0.225), (70, 45, "P3 2", 0.8, 0.28),
         (96, 80, "F1", 1.0, 0.1), (70, 64, "F2", 0.8, 0.1), (64, 70, "F3 1.569988", 1.0, 0.1)]
PAIRS = [(1, 3), (2, 4)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(m, prec, v0):
    descs = [m.case_desc(nx, ny, op, tau, A) for nx, ny, op, tau, A in CASES]
    ctx = m.Context(descs, precision=prec)
    for i, d in enumerate(descs):
        ctx.init_drop(min(d.nx, d.ny) / 4, width=3.45, v0=v0, case=i)
    return ctx


@pytest.mark.parametrize("full,zsym", PAIRS, ids=["f32", "f16"])
@pytest.mark.parametrize("v0", [0.0, 0.05])
def test_zsym_equals_full(full, zsym, v0):
    m = _pkg.load()
    say(f"z-symmetric storage {zsym} vs {full}, v0 {v0}")
    a, b = make(m, full, v0), make(m, zsym, v0)
    try:
        for call, n in enumerate([1, 6, 13]):
            if call == 2:
                for i in range(len(CASES)):
                    a.set_reference(1.02, 0.99, case=i)
                    b.set_reference(1.02, 0.99, case=i)
            a.step(n, record=True)
            b.step(n, record=True)
            for i in range(len(CASES)):
                for x, y in zip(a.download(m.STATE_POST_COLLISION, case=i), b.download(m.STATE_POST_COLLISION, case=i)):
                    assert np.array_equal(x, y), f"populations, call {call}, case {i}"
                for x, y in zip(a.aux(case=i), b.aux(case=i)):
                    assert np.array_equal(x, y), f"aux, call {call}, case {i}"
                da, db = a.diag_drop(v0, 0, case=i), b.diag_drop(v0, 0, case=i)
                assert da.delta_p == db.delta_p and da.raio == db.raio, f"diagnostics, call {call}, case {i}"
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("zsym", [3, 4], ids=["f32", "f16"])
def test_zsym_upload(zsym):
    """A symmetric state from the oracle's initial condition is accepted and steps as the full form
    does; a state with random (non-symmetric) populations is rejected."""
    m = _pkg.load()
    say(f"z-symmetric storage {zsym}: upload")
    nx, ny = 70, 45
    sim = O.Sim(O.make_case(nx, ny, "F2", 0.8, 0.1, 0.8))
    sim.init_drop(11.0, width=3.45, v0=0.03)
    full = zsym - 2
    a = m.Context(m.case_desc(nx, ny, "F2", 0.8, 0.1, 0.8), precision=full)
    b = m.Context(m.case_desc(nx, ny, "F2", 0.8, 0.1, 0.8), precision=zsym)
    try:
        a.upload(sim.fR, sim.fB)
        b.upload(sim.fR, sim.fB)
        for x, y in zip(a.download(m.STATE_POST_STREAM), b.download(m.STATE_POST_STREAM)):
            assert np.array_equal(x, y)
        a.step(5)
        b.step(5)
        for x, y in zip(a.download(m.STATE_POST_COLLISION), b.download(m.STATE_POST_COLLISION)):
            assert np.array_equal(x, y)
        rng = np.random.default_rng(5)
        fR = sim.fR * rng.uniform(0.9, 1.1, sim.fR.size)
        with pytest.raises(m.LBMError):
            b.upload(fR, sim.fB)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("full,zsym", PAIRS, ids=["f32", "f16"])
def test_zsym_band_equals_full(full, zsym):
    """The inclined-band program's state (lbm_init_band, Imm_INCL.cpp:309-334) in both storages: the
    same populations after each call and the same inclined-interface diagnostics."""
    m = _pkg.load()
    cases = [(96, 96, "P3 0.5", 1.0, 0.225, 2, 1), (80, 80, "F2", 0.8, 0.1, 1, 1), (128, 64, "F1", 1.0, 0.1, 1, 0)]
    say(f"z-symmetric storage {zsym} vs {full}: inclined band")
    descs = [m.case_desc(nx, ny, op, tau, A) for nx, ny, op, tau, A, _, _ in cases]
    a, b = m.Context(descs, precision=full), m.Context(descs, precision=zsym)
    try:
        for i, c in enumerate(cases):
            a.init_band(c[5], c[6], case=i)
            b.init_band(c[5], c[6], case=i)
        for call, n in enumerate([3, 20]):
            a.step(n, record=True)
            b.step(n, record=True)
            for i, c in enumerate(cases):
                for x, y in zip(a.download(m.STATE_POST_COLLISION, case=i), b.download(m.STATE_POST_COLLISION, case=i)):
                    assert np.array_equal(x, y), f"populations, call {call}, case {i}"
                da, db = a.diag_incl(c[5], c[6], case=i).as_dict(), b.diag_incl(c[5], c[6], case=i).as_dict()
                for k, va in da.items():   # equal, or both NaN (a fit that did not converge)
                    assert va == db[k] or (va != va and db[k] != db[k]), f"diagnostic {k}, call {call}, case {i}"
    finally:
        a.close()
        b.close()
