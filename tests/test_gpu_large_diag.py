"""GPU drop diagnostics on a box of more than 2^24 sites, against a host FP64 recomputation.

A 4200^2 sharp drop (17.6 M sites), R = 300, P3 chi = 0.5, in FP32 z-symmetric storage, 1000 steps
(the Laplace pressure has built up: delta_p ~ sigma / R).  From the downloaded FP64 post-collision
populations, on the host:

  * the drop's centre, area and radius, and the in / out region counts n_in, n_out (the masks
    delta_p is taken over, Other_functions.cpp:471-528) from the oracle's diagnostics, which sum in site
    order as the reference does: the device's must be equal bit for bit (it sums them in site order at
    every box size);
  * p_in, p_out and delta_p recomputed with exactly rounded sums (math.fsum) over those masks:
    the device's within 1e-12 (delta_p: 1e-12 of p_in).  The reference itself sums these with an
    OpenMP reduction (Other_functions.cpp:494), so its order is not defined; a sequential sum over
    1.5e6 sites is off by ~5e-12 relative, which is why the exact sum is the yardstick here;
  * every other field within 1e-10 of the oracle's (sequential sums over 1.8e7 sites carry ~1e-11).
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

import _oracle as O
import _pkg

pytestmark = pytest.mark.gpu

PRESSURE_FIELDS = ("p_in", "p_out", "delta_p", "sigma", "sd_in", "sd_out", "p_centro", "sigma_linha")
FIT_FIELDS = ("K_perfil", "r_perfil", "esp_perfil", "res_perfil")   # shell histogram: FP64 atomics
COUNTS = ("n_in", "n_out", "n_perfil", "negative_total", "negative_color", "ok", "n_ruim")
FLOOR = {"qx": 1e-12, "qy": 1e-12, "qz": 1e-12, "mean_x": 1e-14, "mean_y": 1e-14, "mean_z": 1e-14,
         "drop_x": 1e-14, "drop_y": 1e-14, "center_dx": 1e-10, "center_dy": 1e-10, "a2": 1e-13, "a4": 1e-13,
         "a8": 1e-13, "raio_linha": 1e-10, "max_mean": 1e-14, "rms_mean": 1e-14}


def image(d, L):
    d = np.where(d > 0.5 * L, d - L, d)
    return np.where(d < -0.5 * L, d + L, d)


def test_diag_drop_above_2_24_sites_matches_host_recomputation():
    m = _pkg.load()
    n, R, steps = 4200, 300.0, 1000
    assert n * n > 1 << 24
    print(f"[gpu] 4200^2 drop R=300, f32z, {steps} steps, lbm_diag_drop vs host FP64 recomputation", flush=True)
    ctx = m.Context(m.case_desc(n, n, "P3 0.5", 1.0, 0.225, 0.8), precision=m.PREC_F32_ZSYM)
    ctx.init_drop(R, width=-1.0)
    ctx.step(steps, record=True)
    g = ctx.diag_drop(0.0, steps - 1)
    fR, fB = ctx.download(m.STATE_POST_COLLISION)
    ctx.close()
    o = O.DropDiag()
    O.lib().orc_drop_diag(C.byref(O.make_case(n, n, "P3 0.5", 1.0, 0.225, 0.8)), fR, fB, None, 0.0, steps - 1,
                          float(n // 2), float(n // 2), C.byref(o))
    rho = fR.reshape(-1, 19).sum(1) + fB.reshape(-1, 19).sum(1)
    del fR, fB
    print(f"    n_in {g.n_in} / {o.n_in}, n_out {g.n_out} / {o.n_out}; delta_p {g.delta_p!r}", flush=True)
    # centre, radius and masks: bitwise the site-ordered sums
    for name in ("x0", "y0", "area", "raio"):
        assert getattr(g, name) == getattr(o, name), (name, getattr(g, name), getattr(o, name))
    for name in COUNTS:
        assert getattr(g, name) == getattr(o, name), (name, getattr(g, name), getattr(o, name))
    # pressures over the masks with exactly rounded sums
    y, x = np.divmod(np.arange(n * n), n)
    dx, dy = image(x - g.x0, n), image(y - g.y0, n)
    r = np.sqrt(dx * dx + dy * dy)
    r_in_max = 0.5 * g.raio
    r_out_min = min(max(2.0 * g.raio, g.raio + 4.0 * 3.45), 0.45 * n)
    p = rho / 3.0
    inside, outside = r < r_in_max, r >= r_out_min
    assert inside.sum() == g.n_in and outside.sum() == g.n_out
    p_in = math.fsum(p[inside]) / g.n_in
    p_out = math.fsum(p[outside]) / g.n_out
    print(f"    p_in {g.p_in!r} / exact {p_in!r}; p_out {g.p_out!r} / exact {p_out!r}; oracle (sequential) "
          f"{o.p_in!r} / {o.p_out!r}", flush=True)
    sys.stdout.flush()
    assert abs(g.p_in - p_in) <= 1e-12 * p_in
    assert abs(g.p_out - p_out) <= 1e-12 * p_out
    assert abs(g.delta_p - (p_in - p_out)) <= 1e-12 * p_in
    assert g.delta_p > 1e-5   # the Laplace pressure of the drop (sigma / R ~ 3e-4)
    # the rest against the oracle
    bad = []
    for name, _ in g._fields_:
        if name in PRESSURE_FIELDS or name in COUNTS:
            continue
        gv, ov = getattr(g, name), getattr(o, name)
        rtol = 1e-8 if name in FIT_FIELDS else 1e-10
        if not (abs(gv - ov) <= rtol * abs(ov) + FLOOR.get(name, 1e-300) or (gv != gv and ov != ov)):
            bad.append((name, gv, ov))
    assert not bad, bad
