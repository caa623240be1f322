"""Pins the CPU oracle before it is trusted as the checker of the GPU library.

  * against the reference's own printed E0 output (E0_resultado.txt, verifica_operadores.cpp);
  * bit for bit against the reference's functions built from its sources with -ffp-contract=off
    (hashes of post-collision populations and force fields, full-precision diagnostics);
  * against the committed Lotes/ time series of the reference (galilean.dat 17 digits, laplace.dat,
    iso.dat, inclinada.dat 10 digits), produced by the reference's optimised build;
  * against what the mathematics fixes: equilibrium fixed point, moments of f^eq, colour-mass
    conservation, the isotropic-stencil gradient of a sinusoid, the push-stream permutation.
"""
import hashlib

import numpy as np
import pytest

import _oracle as O

Q = 19


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# --------------------------------------------------------------------------- E0

def test_e0_matches_reference_output(golden):
    fails, err, tol, m4 = O.e0()
    assert fails == 0
    g = golden["e0"]
    assert g["tudo_certo"]
    printed = ["%10.3e" % e for e in err]
    # rows 0-11 use random normals (the reference draws them from mt19937(7)): tolerance only;
    # every other row is deterministic and must print exactly as the reference printed it
    for i, (p, gold) in enumerate(zip(printed, g["erro"])):
        if i < 12:
            assert abs(err[i]) <= tol[i]
        else:
            assert p.strip() == gold, (i, p, gold)
    assert np.allclose(m4.reshape(4, 5), np.array(g["m4"]), atol=5e-7)


# --------------------------------------------------------------------------- bitwise vs reference

def _run_small(ent):
    a = ent["args"]
    cs = O.make_case(int(a["nx"]), int(a["ny"]), a["op"], float(a["tau"]), float(a["A"]), float(a["beta"]))
    s = O.Sim(cs)
    v0 = float(a.get("v0", 0.0))
    if ent["mode"] == "drop":
        s.init_drop(float(a["R"]), width=float(a["width"]), v0=v0)
    else:
        s.init_band(int(a["p"]), int(a["q"]), width=float(a["width"]))
    dumps = [int(x) for x in str(a["dump"]).split(",")]
    hashes, diags = {}, []
    for st in range(int(a["steps"])):
        s.step()
        if st in dumps:
            e = {"R": h(s.post_R), "B": h(s.post_B)}
            if s.F is not None:
                e["F"] = h(s.F)
            hashes[str(st)] = e
        if st % int(a["diag"]) == 0 or st == int(a["steps"]) - 1:
            diags.append((st, s.drop_diag(v0, st) if ent["mode"] == "drop" else s.incl_diag(int(a["p"]), int(a["q"]))))
    return hashes, diags


DROP_MAP = {"gi_" + k: k for k in ("mean_x mean_y mean_z drop_x drop_y max_ref rms_ref max_mean rms_mean max_lab "
                                   "rms_lab min_f min_rho min_phi max_phi axis_ratio center_dx center_dy "
                                   "x_unwrapped negative_total negative_color").split()}


def _val(d, name):
    return float(getattr(d, DROP_MAP.get(name, name)))


@pytest.mark.parametrize("idx", range(12))
def test_bitwise_vs_reference_build(golden, idx):
    ent = golden["ref"][idx]
    hashes, diags = _run_small(ent)
    assert hashes == ent["hash_nofma"], ent["label"]
    assert len(diags) == len(ent["diag_nofma"])
    for (st, d), ref in zip(diags, ent["diag_nofma"]):
        assert st == int(ref["step"])
        for k, v in ref.items():
            if k == "step":
                continue
            mine = _val(d, k)
            assert mine == v or (np.isnan(mine) and np.isnan(v)), (ent["label"], st, k, mine, v)


def solid_mask(nx, ny, solids):
    m = np.zeros(nx * ny, np.uint8)
    for x, y in solids:
        m[x + nx * y] = 1
    return m


def run_geo(ent):
    """The oracle on a golden box with solid sites; hashes of the fluid sites' populations (site order)."""
    a = ent["args"]
    nx, ny = int(a["nx"]), int(a["ny"])
    solid = solid_mask(nx, ny, a["solid"])
    cs = O.make_case(nx, ny, a["op"], float(a["tau"]), float(a["A"]), float(a["beta"]))
    s = O.Sim(cs, solid=solid, wett=float(a["wett"]))
    s.init_drop(float(a["R"]), width=float(a["width"]))
    fluid = solid == 0
    dumps = [int(x) for x in str(a["dump"]).split(",")]
    hashes = {}
    for st in range(int(a["steps"])):
        s.step()
        if st in dumps:
            e = {"R": h(s.post_R.reshape(-1, Q)[fluid]), "B": h(s.post_B.reshape(-1, Q)[fluid])}
            if s.F is not None:
                e["F"] = h(s.F.reshape(-1, 3)[fluid])
            hashes[str(st)] = e
    return s, hashes


@pytest.mark.parametrize("idx", range(7))
def test_solid_sites_bitwise_vs_reference_build(golden, idx):
    """Bounce-back, wall mediators of wettability wett and the force stencils' reflection at solid
    neighbours: bit for bit the reference's own functions on a geometry read by its read_geo."""
    ent = golden["ref_geo"][idx]
    _, hashes = run_geo(ent)
    assert hashes == ent["hash_nofma"], ent["label"]


def test_solid_sites_conserve_colour_mass_and_wall_changes_result(golden):
    """Bounce-back moves no population out of the fluid: each colour's mass over the fluid sites is
    conserved; and the wall's wettability reaches the fluid (a different wett changes the result)."""
    ent = golden["ref_geo"][2]
    a = ent["args"]
    nx, ny = int(a["nx"]), int(a["ny"])
    fluid = solid_mask(nx, ny, a["solid"]) == 0
    s, _ = run_geo(ent)
    mR0 = s.post_R.reshape(-1, Q)[fluid].sum()
    for _ in range(5):
        s.step()
    assert abs(s.post_R.reshape(-1, Q)[fluid].sum() - mR0) <= 1e-12 * mR0
    ent2 = {"args": dict(a, wett=-float(a["wett"]))}
    s2, _ = run_geo(ent2)
    assert not np.array_equal(s2.post_R.reshape(-1, Q)[fluid], s.post_R.reshape(-1, Q)[fluid])


# --------------------------------------------------------------------------- Lotes/ time series

def _lotes_case(ent):
    c = ent["case"]
    nx = c["nx"]
    cs = O.make_case(nx, nx, c["op"], c["tau"], c["A"], c["beta"])
    return c, cs


# columns whose value depends on a region mask or shell/side assignment that flips under round-off
FLIP_PRONE = {"p_out", "delta_p", "sigma", "raio_linha", "sigma_2pontos", "K_perfil", "r_perfil",
              "espessura_10_90", "residuo_ajuste", "n_camadas", "Ca", "sd_p_out", "|u-v0|max"}


# E1/F3 R=10 (40^2 box) drifts off-centre by amplified round-off only: there the reference's own
# FMA build (which wrote Lotes/) and its -ffp-contract=off build (bit-equal to the oracle) differ by
# up to ~1.5e-6 relative in velocity and extremum columns at step 500, so the case gets 1e-5.
@pytest.mark.parametrize("rel,loose", [("E2/P3 chi=0.5/tau=1/sigma=0.1", 1e-10), ("E2/F2/tau=1/sigma=0.1", 1e-10),
                                       ("E4v/P1/v0=0.1", 1e-10), ("E1/F3/R=10", 1e-5)])
def test_drop_against_lotes(golden, rel, loose):
    ent = golden["lotes"][rel]
    c, cs = _lotes_case(ent)
    s = O.Sim(cs)
    s.init_drop(0.25 * c["nx"], width=3.45, v0=c["v0"])
    gal = ent["tables"]["galilean.dat"]
    lap = ent["tables"]["laplace.dat"]
    iso = ent["tables"]["iso.dat"]
    steps = sorted(int(k) for k in gal["rows"])
    for st in range(max(steps) + 1):
        s.step()
        if st not in steps:
            continue
        d = s.drop_diag(c["v0"], st)
        row = dict(zip(gal["header"], [float(v) for v in gal["rows"][str(st)]]))
        # galilean.dat holds 17 digits: stable columns to 1e-10 relative (FMA-contracted build there)
        for name, mine in [("R", d.raio), ("mass_R", d.massa_R), ("mass_B", d.massa_B), ("p_in", d.p_in),
                           ("x0", d.x0), ("y0", d.y0), ("min_f", d.min_f), ("max_phi", d.max_phi),
                           ("umax_lab", d.max_lab), ("urms_ref", d.rms_ref), ("drop_ux", d.drop_x)]:
            ref = row[name]
            # velocities start at round-off level (~1e-17): absolute floor for them.  The centre of a
            # symmetric drop drifts only by amplified round-off (E1/F3 R=10: the reference's own FMA
            # and -ffp-contract=off builds differ by 7e-7 in x0 at step 500): absolute floor too.
            atol = {"umax_lab": 1e-15, "urms_ref": 1e-15, "drop_ux": 1e-15 if loose < 1e-9 else 1e-7, "x0": 1e-5, "y0": 1e-5}.get(name, 0.0)
            # site extrema and velocities follow the drift
            rtol = loose if name in ("min_f", "max_phi", "umax_lab", "urms_ref", "drop_ux") else 1e-10
            assert abs(mine - ref) <= rtol * abs(ref) + atol, (rel, st, name, mine, ref)
        # p_out / delta_p: the outer region is clamped to 0.45 nx, whose boundary holds lattice sites
        assert abs(d.p_out - row["p_out"]) <= 1e-6 * abs(row["p_out"])
        lrow = dict(zip(lap["header"], lap["rows"][str(st)]))
        for k, mine in [("2:raio_area", d.raio), ("3:raio_grad", d.raio_grad), ("5:p_in", d.p_in),
                        ("12:massa_R", d.massa_R), ("13:massa_B", d.massa_B)]:
            assert float("%.10g" % mine) == float(lrow[k]), (rel, st, k)
        irow = dict(zip(iso["header"], iso["rows"][str(st)]))
        for k, mine in [("2:R0", d.R0), ("6:rmin", d.rmin), ("7:rmax", d.rmax)]:
            assert abs(mine - float(irow[k])) <= max(2e-9, loose) * abs(mine), (rel, st, k)


@pytest.mark.parametrize("rel", ["E3/F1/pq=1,0"])
def test_incl_against_lotes(golden, rel):
    ent = golden["lotes"][rel]
    c, cs = _lotes_case(ent)
    p, q = c["pq"]
    s = O.Sim(cs)
    s.init_band(p, q, width=3.45)
    tab = ent["tables"]["inclinada.dat"]
    for st in range(1):
        s.step()
    d = s.incl_diag(p, q)
    row = dict(zip(tab["header"], [float(v) for v in tab["rows"]["0"]]))
    for k, mine in [("2:sigma_pop", d.sigma_pop), ("3:int_grad", d.int_grad), ("4:K", d.K),
                    ("12:massa_R", d.massa_R), ("13:massa_B", d.massa_B), ("18:min_f", d.min_f)]:
        # 10 printed digits; sigma_pop is round-off (~1e-16) for a force operator at theta = 0
        assert abs(mine - row[k]) <= 1e-9 * abs(row[k]) + 1e-14, (k, mine, row[k])


# --------------------------------------------------------------------------- mathematics

def test_equilibrium_is_fixed_point():
    cs = O.make_case(12, 10, "P3 0.5", 0.7, 0.225)
    s = O.Sim(cs)
    s.init_drop(-1.0, width=-1.0, v0=0.0)   # phi = 0 everywhere: pure blue at rest
    f0 = s.fR.copy(), s.fB.copy()
    for _ in range(3):
        s.step()
    # sum_i w_i rounds to 1 + 1 ulp, so f^eq = w_i rho' moves by round-off only
    assert np.array_equal(s.fR, f0[0]) and np.allclose(s.fB, f0[1], rtol=0, atol=1e-15)


def test_feq_moments():
    c = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                  [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
                  [0, -1, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1]], dtype=float)
    cs = O.make_case(16, 16, "P1", 1.0, 0.0)
    s = O.Sim(cs)
    s.init_drop(5.0, width=3.45, v0=0.07)
    f = (s.fR + s.fB).reshape(-1, Q)
    rho = f.sum(1)
    assert np.allclose(rho, 1.0, atol=1e-14)
    j = f @ c
    assert np.allclose(j[:, 0], 0.07 * rho, atol=1e-15) and np.allclose(j[:, 1:], 0, atol=1e-16)
    P = np.einsum("si,ia,ib->sab", f, c, c)
    assert np.allclose(P[:, 0, 0], rho * (1 / 3 + 0.07 ** 2), atol=1e-15)
    assert np.allclose(P[:, 1, 1], rho / 3, atol=1e-15)


def test_colour_mass_conserved_and_stream_is_permutation():
    cs = O.make_case(30, 26, "F3 1.569988", 0.8, 0.1)
    s = O.Sim(cs)
    s.init_drop(8.0)
    mR, mB = s.fR.sum(), s.fB.sum()
    for _ in range(20):
        s.step()
        # the stream is a permutation of the post-collision array: same multiset of values
        assert np.array_equal(np.sort(s.fR), np.sort(s.post_R))
    assert abs(s.fR.sum() - mR) <= 1e-12 * mR and abs(s.fB.sum() - mB) <= 1e-12 * mB


def test_mediator_gradient_is_central_difference():
    # phi depends on x only: the isotropic stencil reduces to  gm_x = (phi(x-1) - phi(x+1)) / 2
    nx, ny = 24, 8
    cs = O.make_case(nx, ny, "P1", 1.0, 0.0)
    s = O.Sim(cs)
    x = np.arange(nx)
    cR = 0.5 + 0.4 * np.sin(2 * np.pi * x / nx)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    fR = np.zeros((ny, nx, Q))
    fB = np.zeros((ny, nx, Q))
    fR[:] = cR[None, :, None] * w
    fB[:] = (1 - cR)[None, :, None] * w
    s.fR[:] = fR.ravel()
    s.fB[:] = fB.ravel()
    phi = 2 * cR - 1
    s.step()
    gm = s.gm.reshape(ny, nx, 3)
    expect = (np.roll(phi, 1) - np.roll(phi, -1)) / 2
    assert np.allclose(gm[:, :, 0], expect[None, :], atol=1e-15)
    assert np.allclose(gm[:, :, 1], 0, atol=1e-15) and np.allclose(gm[:, :, 2], 0, atol=1e-15)
