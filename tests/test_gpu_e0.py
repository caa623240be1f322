"""E0 operator checks (verifica_operadores.cpp) on the device path: synthetic fields are uploaded, one
FP64 step is taken through the C ABI, and E0's own quantities and tolerances are evaluated on what the
HIP kernels return (populations and the recorded mediator momentum / force field).

  1) class P (verifica_operadores.cpp:57-98): on a state whose colour field is linear around the test
     sites, the perturbation the step adds is the difference of the total post-collision populations
     of a run with amplitude A and one with A = 0 (BGK and recolouring are the same in both, and
     recolouring keeps f_R + f_B).  Its moments: M0 = 0, M1 = 0, M2 = (2/9) A |g| (nn - lambda I) with g
     the recorded mediator momentum, 1e-14 as E0.  The slab has in-plane gradients only, so the normals
     are drawn in the plane (E0 draws them in 3D).
  1b) the in-plane fourth moment M4_nnnn / (A |g|) at 0, 15, 22.5, 30, 45 degrees equals the table the
     reference printed (E0_resultado.txt, 6 digits; verifica_operadores.cpp:100-124).
  2) the z-fold of P3(chi) equals the D2Q9 family member chi + 2 (verifica_operadores.cpp:128-166), at
     the angle 0.3 rad, chi in {-0.4, 0.5, 2, 7}; E0's 1e-15.
  3) Guo forcing (verifica_operadores.cpp:172-186): through the step, the total post-collision momentum
     of every site exceeds the pulled momentum by exactly the recorded force F, and the mass is kept
     (the source's moments sum S = 0, sum S c = (1 - 1/2tau) F together with the shifted equilibrium).
  4) class F on tanh profiles, 256^2, sigma = 0.1, R = 60, xi in {1.569988, 3, 6}
     (verifica_operadores.cpp:188-268): the recorded force equals the oracle's bit for bit; the plane's
     max |F| / (sigma / xi) and the disc's radial integral / (sigma / R) - 1 are within E0's tolerances.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
Q = 19
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
P_CASES = [("P1", 0.0), ("P3 -0.4", 1.0), ("P3 0.5", 1.0), ("P3 2", 1.0)]   # (operator, lambda)
A_P, MG = 0.37, 0.81     # E0's amplitude and |g|


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def linear_patches(n, angles, mg):
    """Colour field phi on an n x n box: around each test site (4i + 1, 4j + 1) the 3 x 3 neighbourhood
    holds phi = -mg n . (x - x_c) (so the mediator momentum, -grad phi, is mg n), elsewhere 0; the state
    is the rest equilibrium rho = 1, f_R = w c_R, f_B = w (1 - c_R), c_R = (1 + phi) / 2."""
    phi = np.zeros((n, n))
    sites = []
    k = 0
    for j in range(n // 4):
        for i in range(n // 4):
            if k >= len(angles):
                break
            xc, yc = 4 * i + 1, 4 * j + 1
            nx_, ny_ = math.cos(angles[k]), math.sin(angles[k])
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    phi[yc + dy, xc + dx] = -mg * (nx_ * dx + ny_ * dy)
            sites.append(xc + n * yc)
            k += 1
    cR = 0.5 * (1.0 + phi.reshape(-1))
    fR = (cR[:, None] * W[None, :]).reshape(-1)
    fB = ((1.0 - cR)[:, None] * W[None, :]).reshape(-1)
    return fR, fB, np.array(sites)


def perturbation_by_difference(m, op, angles, n=64):
    """Per test site: the 19 values the perturbation added (A run minus A = 0 run) and the recorded g."""
    fR, fB, sites = linear_patches(n, angles, MG)
    out = {}
    for A in (A_P, 0.0):
        ctx = m.Context(m.case_desc(n, n, op, 1.0, A, 0.8))
        ctx.upload(fR, fB)
        ctx.step(1, record=True)
        R, B = ctx.download(m.STATE_POST_COLLISION)
        gm, _ = ctx.aux()
        ctx.close()
        out[A] = ((R + B).reshape(-1, Q), gm.reshape(-1, 3))
    pert = out[A_P][0][sites] - out[0.0][0][sites]
    g = out[A_P][1][sites]
    assert np.array_equal(g, out[0.0][1][sites])
    return pert, g


def test_e0_class_p_moments():
    m = _pkg.load()
    rng = np.random.default_rng(7)
    angles = rng.uniform(0.0, 2 * math.pi, 200)
    for op, lam in P_CASES:
        say(f"[gpu] E0 1) class P moments through the step, {op}, 200 in-plane normals")
        pert, g = perturbation_by_difference(m, op, angles)
        e0 = e1 = e2 = 0.0
        for f, gv in zip(pert, g):
            mg = math.sqrt(gv @ gv)
            assert abs(mg - MG) <= 1e-14
            nvec = gv / mg
            M0 = f.sum()
            M1 = np.array([(f * CX).sum(), (f * CY).sum(), (f * CZ).sum()])
            C = np.stack([CX, CY, CZ])
            M2 = (C[:, None, :] * C[None, :, :] * f).sum(-1)
            target = 2.0 / 9.0 * A_P * mg * (np.outer(nvec, nvec) - lam * np.eye(3))
            e0, e1, e2 = max(e0, abs(M0)), max(e1, np.abs(M1).max()), max(e2, np.abs(M2 - target).max())
        say(f"    M0 {e0:.2e}  M1 {e1:.2e}  M2 {e2:.2e}")
        assert e0 <= 1e-14 and e1 <= 1e-14 and e2 <= 1e-14, (op, e0, e1, e2)


def test_e0_fourth_moment_table():
    m = _pkg.load()
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        table = np.array(json.load(f)["e0"]["m4"])
    angles = [math.radians(t) for t in (0.0, 15.0, 22.5, 30.0, 45.0)]
    for row, (op, _) in zip(table, P_CASES):
        say(f"[gpu] E0 1b) M4_nnnn / (A |g|) in the plane, {op}")
        pert, g = perturbation_by_difference(m, op, angles)
        vals = []
        for f, gv in zip(pert, g):
            mg = math.sqrt(gv @ gv)
            nvec = gv / mg
            cn = CX * nvec[0] + CY * nvec[1]
            vals.append((f * cn ** 4).sum() / (A_P * mg))
        say("    " + " ".join(f"{v:9.6f}" for v in vals))
        assert np.allclose(vals, row, atol=5e-7), (op, vals, row)


def test_e0_d2q9_fold():
    m = _pkg.load()
    for chi in (-0.4, 0.5, 2.0, 7.0):
        say(f"[gpu] E0 2) z-fold of P3 chi={chi} vs D2Q9 chi + 2")
        pert, g = perturbation_by_difference(m, f"P3 {chi}", [0.3])
        f, gv = pert[0], g[0]
        mg = math.sqrt(gv @ gv)
        f9 = np.zeros((3, 3))
        for i in range(Q):
            f9[CX[i] + 1, CY[i] + 1] += f[i]
        c2 = chi + 2.0
        B9 = [-c2 / (3.0 * (c2 + 2.0)), c2 / (6.0 * (c2 + 2.0)), 1.0 / (6.0 * (c2 + 2.0))]
        err = 0.0
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                s = abs(a) + abs(b)
                w9 = 4.0 / 9.0 if s == 0 else (1.0 / 9.0 if s == 1 else 1.0 / 36.0)
                cg = a * gv[0] + b * gv[1]
                err = max(err, abs(A_P * mg * (w9 * cg * cg / (mg * mg) - B9[s]) - f9[a + 1, b + 1]))
        say(f"    error {err:.2e}")
        assert err <= 1e-15, (chi, err)


def tanh_state(n, xi, disc, R=60.0):
    """verifica_operadores.cpp:206-223: phi = tanh(s / xi), s the signed distance to the disc of radius R
    or to the two planes |x - n/2| = n/4; f_R = w (1 + phi) / 2, f_B = w (1 - phi) / 2."""
    y, x = np.mgrid[0:n, 0:n].astype(float)
    if disc:
        s = R - np.hypot(x - n / 2.0, y - n / 2.0)
    else:
        ax = np.abs(x - n / 2.0)
        s = np.where(ax < n / 4.0, n / 4.0 - ax, -(ax - n / 4.0))
    ph = np.tanh(s / xi).reshape(-1)
    fR = (W[None, :] * 0.5 * (1 + ph)[:, None]).reshape(-1)
    fB = (W[None, :] * 0.5 * (1 - ph)[:, None]).reshape(-1)
    return fR, fB


@pytest.mark.parametrize("xi", [1.569988, 3.0, 6.0])
def test_e0_class_f_tanh_profiles_and_guo(xi):
    m = _pkg.load()
    n, sigma, R = 256, 0.1, 60.0
    y, x = np.mgrid[0:n, 0:n].astype(float)
    dx, dy = (x - n / 2.0).reshape(-1), (y - n / 2.0).reshape(-1)
    r = np.hypot(dx, dy)
    inside = r >= 1e-9
    for op in ("F1", "F2", f"F3 {xi!r}"):
        say(f"[gpu] E0 4) {op} on tanh plane and disc, 256^2, xi={xi}")
        res = {}
        for disc in (False, True):
            fR, fB = tanh_state(n, xi, disc, R)
            ctx = m.Context(m.case_desc(n, n, op, 1.0, sigma, 0.8))
            ctx.upload(fR, fB)
            ctx.step(1, record=True)
            _, F = ctx.aux()
            pR, pB = ctx.download(m.STATE_POST_COLLISION)
            ctx.close()
            # the oracle's force field on the same state, bit for bit
            oF = np.zeros(3 * n * n)
            O.lib().orc_force_field(O.make_case(n, n, op, 1.0, sigma, 0.8), fR, fB, oF)
            assert np.array_equal(F, oF), (op, disc, np.abs(F - oF).max())
            # 3) Guo: the step adds exactly F to each site's momentum and keeps its mass
            pre = (fR + fB).reshape(-1, Q)
            post = (pR + pB).reshape(-1, Q)
            Fv = F.reshape(-1, 3)
            dmom = np.stack([(post - pre) @ CX, (post - pre) @ CY, (post - pre) @ CZ], axis=1)
            assert np.abs(dmom - Fv).max() <= 1e-14, (op, disc, np.abs(dmom - Fv).max())
            assert np.abs(post.sum(1) - pre.sum(1)).max() <= 1e-14
            res[disc] = Fv
        fmax = np.max(np.abs(res[False][:, 0]) + np.abs(res[False][:, 1]))
        Fd = res[True]
        soma = np.sum(-(Fd[inside, 0] * dx[inside] + Fd[inside, 1] * dy[inside]) / r[inside] / (2.0 * math.pi * r[inside]))
        f3 = op.startswith("F3")
        tol_p = 0.2 / (xi * xi) if f3 else 1e-3
        tol_d = 1.5 * (xi / R) ** 2 + 5e-4 + (0.25 / (xi * xi) if f3 else 0.0)
        ep, ed = fmax / (sigma / xi), soma / (sigma / R) - 1.0
        say(f"    plane max|F|/(sigma/xi) {ep:.3e} (tol {tol_p:.1e}); disc {ed:.3e} (tol {tol_d:.1e})")
        assert abs(ep) <= tol_p and abs(ed) <= tol_d, (op, ep, ed)
