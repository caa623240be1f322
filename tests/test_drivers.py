"""The drop-in programs Imm_OPS / Imm_INCL: the reference's case-directory contract.

CPU (no device): input parsing, the dat_out.txt echo (text equal to the reference's own files for
the same data_in.txt) and the error exits (status 1 with the reference's messages) for invalid
inputs.  Without a GPU the programs stop at lbm_create, after the inputs have been read.

GPU: committed reference runs (tests/golden/driver_cases.json, copied from the reference's Lotes/
tree) are run again from the same inputs, whole (20000-30000 steps), and every output table is
compared twice:

* with the same case run through the reference programs built with -ffp-contract=off (the build
  whose arithmetic the FP64 library reproduces bit for bit; fixture "nofma"): same headers, same
  rows, every number within RTOL_NOFMA = 2e-9 relative (the 10 printed digits) plus ATOL = 1e-12.
  Only the diagnostics' summation order differs (tree reductions on the device), so columns that are
  cancellations of much larger sums get absolute floors (COLUMN_TOL);
* with the committed Lotes file itself (the reference's -O3 -march=native build, contracted
  multiply-adds): |ours - Lotes| <= 2 |nofma - Lotes| + RTOL_NOFMA |Lotes| + floor, i.e. the library
  is as close to the published outputs as the reference's own no-FMA build is.

Columns that are round-off noise in the reference itself (the centroid drift mean_x, momentum
components, |a2|/R0 of a centred drop, ~1e-17) fall under ATOL.  MLUPS is not compared.
"""
import math
import os
import sys

import pytest

import _cases as K

RTOL = 2e-9   # 10 printed significant digits
RTOL_NOFMA = RTOL
ATOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ CPU


def test_programs_are_built():
    K.program("Imm_OPS")
    K.program("Imm_INCL")


@pytest.mark.parametrize("c", K.golden_cases(), ids=lambda c: c["label"])
def test_dat_out_echo_matches_reference(tmp_path, c):
    K.make_golden_case(str(tmp_path), c)
    K.run(c["program"], [], str(tmp_path), timeout=120)
    with open(tmp_path / "dat_out.txt") as f:
        assert f.read() == c["outputs"]["dat_out.txt"]


def _stderr_out(r):
    return r.stdout + r.stderr


def test_invalid_inputs_exit_1(tmp_path):
    cases = {
        "no_operator": dict(),
        "bad_operator": dict(op="P2"),
        "negative_v0": dict(v0=-0.1),
        "tau_mismatch": dict(tau_B=0.7),
        "solid_site": dict(solid=(3, 4)),
        "bad_option": dict(options={"precisao": "f8"}),
    }
    msgs = {
        "no_operator": "operador.txt ausente ou invalido",
        "bad_operator": "operador.txt ausente ou invalido",
        "negative_v0": "velocity.txt invalido",
        "tau_mismatch": "tau_R e tau_B tem de ser iguais",
        "solid_site": "toda fluida",
        "bad_option": "itolbm.txt",
    }
    for name, kw in cases.items():
        d = tmp_path / name
        K.make_case(str(d), 16, kw.get("op", "P3 0.5"), 1.0, 0.225, 10, v0=kw.get("v0", 0.0),
                    tau_B=kw.get("tau_B"), options=kw.get("options"))
        if name == "no_operator":
            os.remove(d / "operador.txt")
        if "solid" in kw:
            K.write_box(str(d), 16, solid=kw["solid"])
        if name == "negative_v0":
            with open(d / "velocity.txt", "w") as f:
                f.write("-0.1\n")
        r = K.run("Imm_OPS", [], str(d), timeout=60)
        assert r.returncode == 1, name
        assert msgs[name] in _stderr_out(r), (name, _stderr_out(r))


def test_incl_rejects_zero_normal(tmp_path):
    K.make_case(str(tmp_path), 16, "F1", 1.0, 0.1, 10, pq=(0, 0))
    r = K.run("Imm_INCL", [], str(tmp_path), timeout=60)
    assert r.returncode == 1 and "inclinacao.txt" in _stderr_out(r)


def test_options_file(tmp_path):
    K.make_case(str(tmp_path), 16, "P1", 1.0, 0.225, 10, options={"precisao": "f32", "raio": 5, "espessura_ini": -1,
                                                               "passos_diag": 100, "digitos": 17, "referencia": 0,
                                                               "vtk": "binario"})
    r = K.run("Imm_OPS", [], str(tmp_path), timeout=60)
    out = _stderr_out(r)
    assert "raio inicial      = 5" in out and "precisao          = f32" in out
    with open(tmp_path / "itolbm.txt", "a") as f:
        f.write("vtk xml\n")
    r = K.run("Imm_OPS", [], str(tmp_path), timeout=60)
    assert r.returncode == 1 and "bad value for vtk" in _stderr_out(r)
    K.make_case(str(tmp_path), 16, "P1", 1.0, 0.225, 10, options={"vtk": "binario_double"})
    with open(tmp_path / "itolbm.txt", "a") as f:
        f.write("cor azul\n")
    r = K.run("Imm_OPS", [], str(tmp_path), timeout=60)
    assert r.returncode == 1 and "unknown key cor" in _stderr_out(r)


# ------------------------------------------------------------------ GPU


def _num(s):
    return float(s)   # "nan" / "-nan" / "inf" parse too


# Per-column tolerances (1-based columns) where the reference's own value is round-off dominated:
#   sd_p_in / sd_p_out: sqrt(<p^2> - <p>^2) cancels ~7 digits (and is pure noise at step 0);
#   min_f: smallest population, ~1e-11 in a pure phase (round-off of w_i - w_i);
#   d0: band displacement along the normal, ~1e-11 for a symmetric band.
#   qx, qy of a band: momentum sums of +-1e-3 terms over 40000 sites, ~1e-12.
COLUMN_TOL = {
    "laplace.dat": {15: (1e-6, 1e-7), 16: (1e-6, 1e-7)},
    "galilean.dat": {24: (RTOL, 1e-9)},
    "resumo.dat/Imm_OPS": {24: (RTOL, 1e-9)},
    "resumo.dat/Imm_INCL": {23: (RTOL, 1e-9), 24: (RTOL, 1e-9), 27: (RTOL, 1e-9)},
    "inclinada.dat": {7: (RTOL, 1e-9), 14: (RTOL, 1e-9), 15: (RTOL, 1e-9), 18: (RTOL, 1e-9)},
}


def compare_table(name, ours, ref, skip_cols=(), col_tol=None, row_skip=None):
    """Headers equal, same rows, numbers within tolerance.  Returns the largest relative deviation
    above ATOL; raises listing every violation."""
    col_tol = col_tol or {}
    ho, ro = ours
    hr, rr = ref
    assert ho == hr, f"{name}: header differs"
    assert len(ro) == len(rr), f"{name}: {len(ro)} rows, reference {len(rr)}"
    worst = (0.0, None)
    bad = []
    for a, b in zip(ro, rr):
        assert len(a) == len(b), f"{name}: row {b[0]} has {len(a)} columns, reference {len(b)}"
        for j, (x, y) in enumerate(zip(a, b)):
            col = j + 1
            if j in skip_cols or (row_skip and row_skip(b, col)):
                continue
            try:
                fx, fy = _num(x), _num(y)
            except ValueError:
                if x != y:
                    bad.append(f"row {b[0]} col {col}: {x!r} != {y!r}")
                continue
            if math.isnan(fy):
                if not math.isnan(fx):
                    bad.append(f"row {b[0]} col {col}: {x} != nan")
                continue
            rt, at = col_tol.get(col, (RTOL, ATOL))
            d = abs(fx - fy)
            if not d <= rt * abs(fy) + at:
                bad.append(f"row {b[0]} col {col}: {x} vs reference {y}")
            if d > at and d / max(abs(fy), 1e-300) > worst[0]:
                worst = (d / abs(fy), f"row {b[0]} col {col}")
    assert not bad, f"{name}: {len(bad)} values out of tolerance: " + "; ".join(bad[:25])
    return worst


def compare_bounded(name, ours, lotes, nofma, skip_cols=(), col_tol=None, row_skip=None):
    """|ours - Lotes| <= 2 |nofma - Lotes| + tolerance, value by value."""
    col_tol = col_tol or {}
    bad = []
    for a, b, n in zip(ours[1], lotes[1], nofma[1]):
        for j, (x, y, z) in enumerate(zip(a, b, n)):
            col = j + 1
            if j in skip_cols or (row_skip and row_skip(b, col)):
                continue
            try:
                fx, fy, fz = _num(x), _num(y), _num(z)
            except ValueError:
                continue
            if math.isnan(fy) or math.isnan(fz):
                continue
            rt, at = col_tol.get(col, (RTOL, ATOL))
            if not abs(fx - fy) <= 2.0 * abs(fz - fy) + rt * abs(fy) + at:
                bad.append(f"row {b[0]} col {col}: {x} vs Lotes {y} (no-FMA build {z})")
    assert not bad, f"{name}: {len(bad)} values further from Lotes than the no-FMA build: " + "; ".join(bad[:25])


def _ca_noise(row, col):
    # Ca = nu |u|max / sigma is 0 for sigma <= 0: at step 0 sigma is round-off (~1e-15) in both runs
    return col == 11 and abs(float(row[7])) < 1e-8


def _parse_text(text):
    lines = text.splitlines()
    head = lines[0] if lines and lines[0].startswith("#") else None
    return head, [ln.split() for ln in lines[1 if head else 0:] if ln.strip()]


@pytest.mark.gpu
@pytest.mark.parametrize("c", K.golden_cases(), ids=lambda c: c["label"])
def test_reference_run_reproduced(tmp_path, c):
    say(f"[gpu] {c['program']} {c['label']}")
    d = str(tmp_path)
    K.make_golden_case(d, c)
    r = K.run(c["program"], [], d, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    with open(os.path.join(d, "dat_out.txt")) as f:
        assert f.read() == c["outputs"]["dat_out.txt"]
    for name, text in c["outputs"].items():
        if name == "dat_out.txt":
            continue
        ours = K.read_table(os.path.join(d, name))
        lotes = _parse_text(text)
        nofma = _parse_text(c["nofma"]["outputs"][name])
        skip = (len(lotes[1][0]) - 1,) if name == "resumo.dat" else ()   # MLUPS: timing
        key = f"resumo.dat/{c['program']}" if name == "resumo.dat" else name
        tol = COLUMN_TOL.get(key)
        rskip = _ca_noise if name == "laplace.dat" else None
        worst = compare_table(name + " vs no-FMA reference build", ours, nofma, skip, tol, rskip)
        compare_table(name + " (layout) vs Lotes", (ours[0], ours[1]), (lotes[0], lotes[1]), tuple(range(99)))
        compare_bounded(name + " vs Lotes", ours, lotes, nofma, skip, tol, rskip)
        say(f"    {name}: {len(lotes[1])} rows; largest deviation from the no-FMA build {worst[0]:.2e} at {worst[1]}")
    if "perfil" in c:
        with open(os.path.join(d, "perfil.dat")) as f:
            lines = f.read().splitlines()
        for ref, label in ((c["nofma"]["perfil"], "no-FMA build"), (c["perfil"], "Lotes")):
            assert lines[0] == ref["lines"]["0"]
            if label == "no-FMA build":
                assert len(lines) == ref["n_lines"]
                ours = (None, [lines[int(i)].split() for i in ref["lines"] if int(i) > 0])
                theirs = (None, [ln.split() for i, ln in ref["lines"].items() if int(i) > 0])
                compare_table("perfil.dat vs " + label, ours, theirs)


@pytest.mark.gpu
def test_batch_of_case_dirs_equals_single_runs(tmp_path):
    """Several case directories in one invocation (one GPU context, batched launches): each writes the
    same files as when run alone."""
    say("[gpu] Imm_OPS batch of 4 uneven cases vs single runs")
    spec = [(40, "P3 0.5", 0.225, 1.0, 1200, 0.0), (48, "F2", 0.1, 0.8, 700, 0.05), (36, "F1", 0.1, 1.0, 1000, 0.0),
            (44, "P1", 0.3, 0.6, 500, 0.0)]
    dirs_b, dirs_s = [], []
    for k, (nx, op, A, tau, steps, v0) in enumerate(spec):
        for root, lst in (("b", dirs_b), ("s", dirs_s)):
            d = str(tmp_path / root / f"c{k}")
            K.make_case(d, nx, op, tau, A, steps, v0=v0)
            lst.append(d)
    r = K.run("Imm_OPS", dirs_b, str(tmp_path), timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    for d in dirs_s:
        r = K.run("Imm_OPS", [], d, timeout=600)
        assert r.returncode == 0, r.stdout + r.stderr
    for db, ds in zip(dirs_b, dirs_s):
        for name in ("laplace.dat", "galilean.dat", "iso.dat"):
            with open(os.path.join(db, name)) as f1, open(os.path.join(ds, name)) as f2:
                assert f1.read() == f2.read(), (db, name)
        rb = K.read_table(os.path.join(db, "resumo.dat"))[1][0]
        rs = K.read_table(os.path.join(ds, "resumo.dat"))[1][0]
        assert rb[:-1] == rs[:-1]
        steps = int(rs[12])
        for field in ("rho_R", "rho_B", "vel_"):
            with open(os.path.join(db, f"{field}{steps - 1}.vtk")) as f1, \
                    open(os.path.join(ds, f"{field}{steps - 1}.vtk")) as f2:
                assert f1.read() == f2.read()


@pytest.mark.gpu
def test_vtk_fields_match_oracle(tmp_path):
    """The final-step VTK fields against the oracle's post-collision populations (same formulas)."""
    import numpy as np
    import _oracle as O
    say("[gpu] Imm_OPS VTK fields vs oracle, 30x26 F2 v0=0.05, 40 steps")
    d = str(tmp_path)
    K.make_case(d, 30, "F2", 0.8, 0.1, 40, v0=0.05, ny=26)
    r = K.run("Imm_OPS", [], d, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    sim = O.Sim(O.make_case(30, 26, "F2", 0.8, 0.1, 0.8))
    sim.init_drop(7.5, width=3.45, x0=15, y0=13, v0=0.05)
    for _ in range(40):
        sim.step()
    fR = sim.post_R.reshape(-1, 19)
    fB = sim.post_B.reshape(-1, 19)

    def vtk_values(path):
        with open(path) as f:
            lines = f.read().splitlines()
        start = next(i for i, ln in enumerate(lines) if ln.startswith(("LOOKUP_TABLE", "VECTORS"))) + 1
        return np.array([float(v) for ln in lines[start:] for v in ln.split()])

    rho_R = vtk_values(os.path.join(d, "rho_R39.vtk"))
    assert np.allclose(rho_R, fR.sum(1), rtol=6e-6, atol=1e-12)   # printed with 6 significant digits
    vel = vtk_values(os.path.join(d, "vel_39.vtk")).reshape(-1, 3)
    c = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
                  [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1]], float)
    rho = fR.sum(1) + fB.sum(1)
    u = ((fR + fB) @ c.T - 0.5 * sim.F.reshape(-1, 3)[:, :2]) / rho[:, None]
    assert np.allclose(vel[:, :2], u, rtol=2e-5, atol=1e-10)
    assert np.abs(vel[:, 2]).max() <= 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("form,dt", [("binario", ">f4"), ("binario_double", ">f8")])
def test_vtk_binary_option(tmp_path, form, dt):
    """itolbm.txt "vtk binario[_double]": the same fields as the ASCII files of the same run, in the
    binary legacy form (big-endian float / double after the header, then a newline)."""
    import numpy as np
    say(f"[gpu] Imm_OPS vtk {form}, 24x24 P3 0.5, 20 steps")
    da, db = str(tmp_path / "a"), str(tmp_path / "b")
    for d, opt in ((da, {}), (db, {"vtk": form})):
        os.makedirs(d)
        K.make_case(d, 24, "P3 0.5", 1.0, 0.225, 20, options=opt)
        r = K.run("Imm_OPS", [], d, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
    for name, ncomp in (("rho_R19.vtk", 1), ("rho_B19.vtk", 1), ("vel_19.vtk", 3)):
        with open(os.path.join(da, name)) as f:
            lines = f.read().splitlines()
        start = next(i for i, ln in enumerate(lines) if ln.startswith(("LOOKUP_TABLE", "VECTORS"))) + 1
        asc = np.array([float(v) for ln in lines[start:] for v in ln.split()])
        raw = open(os.path.join(db, name), "rb").read()
        head = raw.split(b"\n")[:start]
        assert [h.decode() for h in head] == [ln.replace("ASCII", "BINARY").replace(
            "double", "float" if dt == ">f4" else "double") for ln in lines[:start]]
        body = raw[len(b"\n".join(head)) + 1:]
        assert body.endswith(b"\n") and len(body) - 1 == asc.size * np.dtype(dt).itemsize
        binv = np.frombuffer(body[:-1], dtype=dt).astype(float)
        assert np.allclose(binv, asc, rtol=6e-6, atol=1e-12), name   # ASCII keeps 6 digits
        assert asc.size == 24 * 24 * ncomp


@pytest.mark.gpu
def test_reduced_precision_forms_through_the_programs(tmp_path):
    """The drop-in program with itolbm.txt selecting FP32 / FP16 storage (and re-centring after each
    diagnostic): laplace.dat's last row (a 96^2 drop, R = 19.2, 6000 steps) against the FP64 run of
    the same case.  Tolerances as in test_gpu_precision.py: delta_p (column 7) 2e-4 / 3e-2 relative,
    radius (column 2) 1e-5 / 1e-3, colour masses (columns 12-13) 1e-5 / 5e-4.  The z-symmetric forms
    (f32z, f16z) must write the same bytes as f32 / f16."""
    rows = {}
    for prec in ("f64", "f32", "f16", "f32z", "f16z"):
        say(f"[gpu] Imm_OPS 96^2 P3 0.5 drop, precisao {prec}")
        d = str(tmp_path / prec)
        K.make_case(d, 96, "P3 0.5", 1.0, 0.225, 6000, options={"precisao": prec, "referencia": 1})
        r = K.run("Imm_OPS", [], d, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        rows[prec] = [float(v) for v in K.read_table(os.path.join(d, "laplace.dat"))[1][-1]]
        if prec.endswith("z"):
            # z-symmetric storage: every output file byte-equal to the full form's
            for name in ("laplace.dat", "galilean.dat", "iso.dat"):
                with open(os.path.join(d, name), "rb") as fz, open(os.path.join(str(tmp_path / prec[:-1]), name), "rb") as ff:
                    assert fz.read() == ff.read(), (prec, name)
    ref = rows["f64"]
    for prec, tol in (("f32", (2e-4, 1e-5, 1e-5)), ("f16", (3e-2, 1e-3, 5e-4))):
        got = rows[prec]
        assert got[0] == ref[0]   # same step
        assert abs(got[6] - ref[6]) <= tol[0] * abs(ref[6]), (prec, "delta_p", got[6], ref[6])
        assert abs(got[1] - ref[1]) <= tol[1] * ref[1], (prec, "radius", got[1], ref[1])
        for col in (11, 12):
            assert abs(got[col] - ref[col]) <= tol[2] * ref[col], (prec, "mass", col, got[col], ref[col])
