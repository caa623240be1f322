#!/usr/bin/env python3
"""Batches of drop / inclined-interface cases on one GPU (the GPU counterpart of rodar_lote.py).

Cases are case directories laid out exactly as the reference's batch script lays them out
(rodar_lote.py:74-223): <saida>/<batch>/<label>/... with data_in.txt, operador.txt, velocity.txt,
[inclinacao.txt] and an all-fluid meio.vtk, and the same resume rule: a case is done when its
resumo.dat holds a data line; an unfinished case is rebuilt from scratch.  analisar_lote.py reads
the results unchanged.

What changes is how the cases run: instead of one CPU process per case, all pending cases of a
program (Imm_OPS or Imm_INCL, built next to this file) go to ONE invocation of the GPU program with
every case directory as an argument; it advances them together in one GPU context (one kernel launch
per operator family and step, whatever the box sizes) and writes each case's files and
execucao.log.  --grupo splits a batch into invocations of at most that many cases.

Batches: the reference's E1, E2, E2chi, E4, E4v, E3, E3chi ("todos" = all seven) and S168, the
168-case static-drop sweep: tau in {0.51, 0.6, 0.8, 1, 1.5, 2} x sigma in {0.01, 0.05, 0.1, 0.5} x
the seven operators, radius R drawn uniformly from 10..50 with numpy.random.default_rng(2026) (one
integer per case, in that loop order), square box of side 4R (so the programs' default initial radius
nx / 4 is R), 20000 steps.

Usage
    python3 lote.py E1 --lista
    python3 lote.py E2 --precisao f32
    python3 lote.py S168 --precisao f32 --saida Lotes_gpu
    python3 lote.py E1 --passos 2000 --filtro F1 --saida Lotes_teste
"""

import argparse
import math
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
PROGRAMS = {"Imm_OPS": HERE / "Imm_OPS", "Imm_INCL": HERE / "Imm_INCL"}

BETA = 0.8            # Latva-Kokko recolouring, width 3.45 sites
XI_F3 = 1.569988      # 2 / K at beta = 0.8

# (folder label, operador.txt line, class) -- the reference's core operator set (rodar_lote.py:50-58)
OPERATORS = [
    ("P1", "P1", "P"),
    ("P3 chi=-0.4", "P3 -0.4", "P"),
    ("P3 chi=0.5", "P3 0.5", "P"),
    ("P3 chi=2", "P3 2", "P"),
    ("F1", "F1", "F"),
    ("F2", "F2", "F"),
    ("F3", f"F3 {XI_F3}", "F"),
]
FINE_CHI = [-2.0, -1.3, -1.0, -0.7, -0.55, -0.45, -0.4, -0.35, -0.3, -0.25, -0.15, 0.0, 0.5, 1.0, 2.0]
SLOPES = [(1, 0), (4, 1), (3, 1), (2, 1), (3, 2), (1, 1)]


def nu(tau):
    return (tau - 0.5) / 3.0


def amplitude(cls, sigma, tau):
    """fat_R_B: A = 9 sigma / (4 tau) for the perturbation operators, sigma for the forces."""
    return 9.0 * sigma / (4.0 * tau) if cls == "P" else sigma


def steps_to_converge(R, tau, low=30000, high=100000):
    """At least twice the viscous time R^2 / nu, rounded up to 500, within 30 000 .. 100 000."""
    return int(min(high, max(low, math.ceil(2.0 * R * R / nu(tau) / 500.0) * 500)))


def fmt(x):
    return f"{x:g}"


def case(folder, op, cls, sigma, tau, nx, steps, v0=0.0, batch="", pq=None):
    return dict(folder=Path(folder), op=op, cls=cls, sigma=sigma, tau=tau, nx=nx, steps=steps, v0=v0,
                batch=batch, A=amplitude(cls, sigma, tau), pq=pq, program="Imm_INCL" if pq else "Imm_OPS")


def batch_E1():
    return [case(Path("E1") / lab / f"R={R}", op, cl, 0.1, 1.0, 4 * R, 30000, batch="E1")
            for lab, op, cl in OPERATORS for R in (10, 15, 20, 25, 30, 40, 50)]


def batch_E2():
    return [case(Path("E2") / lab / f"tau={fmt(tau)}" / f"sigma={fmt(s)}", op, cl, s, tau, 100,
                 steps_to_converge(25, tau), batch="E2")
            for lab, op, cl in OPERATORS for tau in (0.6, 0.8, 1.0, 1.5) for s in (0.001, 0.003, 0.01, 0.03, 0.1, 0.3)]


def batch_E2chi():
    return [case(Path("E2chi") / f"tau={fmt(tau)}" / f"chi={fmt(chi)}", f"P3 {chi}", "P", 0.1, tau, 100,
                 steps_to_converge(25, tau), batch="E2chi")
            for tau in (0.6, 1.0) for chi in FINE_CHI]


def batch_E4():
    return [case(Path("E4") / lab / f"tau={fmt(tau)}" / f"sigma={fmt(s)}", op, cl, s, tau, 100, 30000, batch="E4")
            for lab, op, cl in OPERATORS for tau in (0.51, 0.52, 0.55, 0.6, 0.8, 1.0, 1.5, 2.0)
            for s in (0.1, 0.3, 0.5, 1.0, 2.0)]


def batch_E4v():
    return [case(Path("E4v") / lab / f"v0={fmt(v0)}", op, cl, 0.1, 1.0, 100, 30000, v0=v0, batch="E4v")
            for lab, op, cl in OPERATORS for v0 in (0.05, 0.1, 0.15, 0.2, 0.25)]


def batch_E3():
    return [case(Path("E3") / lab / f"pq={p},{q}", op, cl, 0.1, 1.0, 200, 20000, batch="E3", pq=(p, q))
            for lab, op, cl in OPERATORS for p, q in SLOPES]


def batch_E3chi():
    return [case(Path("E3chi") / f"chi={fmt(chi)}", f"P3 {chi}", "P", 0.1, 1.0, 200, 20000, batch="E3chi", pq=(2, 1))
            for chi in FINE_CHI]


def batch_S168():
    import numpy as np
    radii = np.random.default_rng(2026).integers(10, 51, 168)
    out, k = [], 0
    for tau in (0.51, 0.6, 0.8, 1.0, 1.5, 2.0):
        for s in (0.01, 0.05, 0.1, 0.5):
            for lab, op, cl in OPERATORS:
                R = int(radii[k])
                k += 1
                c = case(Path("S168") / lab / f"tau={fmt(tau)}" / f"sigma={fmt(s)}", op, cl, s, tau, 4 * R, 20000,
                         batch="S168")
                c["radius"] = R
                out.append(c)
    return out


BATCHES = {"E1": batch_E1, "E2": batch_E2, "E2chi": batch_E2chi, "E4": batch_E4, "E4v": batch_E4v,
           "E3": batch_E3, "E3chi": batch_E3chi, "S168": batch_S168}
ALL = ["E1", "E2", "E2chi", "E4", "E4v", "E3", "E3chi"]

DATA_IN = """meio.vtk

Dim_Pixel\t1.0

steps:\t    {steps}

files:\t    0

n_threads:  1

tau:\t    {tau}

tau_nh:     1.0

rho:\t    1.0

tau_R:\t    {tau}

tau_B:\t    {tau}

tau_m:\t    1.0

rho_R:\t    1.0

rho_B:\t    1.0

fat_R_B:\t{A:.12g}

fat_R_R:\t0.0

fat_B_B:    0.0

recoll:\t    {beta}

wett_R:\t    0.0
"""


def write_box(folder, nx):
    with open(folder / "meio.vtk", "w") as f:
        f.write("# vtk DataFile Version 2.0\nGeometria\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {nx} 1\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * nx}\n")
        f.write("SCALARS Geometria float\nLOOKUP_TABLE default\n")
        f.write((" ".join(["1"] * nx) + " \n") * nx)


def build(c, root, precision):
    """Writes the case directory (inputs only; an old resumo.dat is removed)."""
    d = root / c["folder"]
    d.mkdir(parents=True, exist_ok=True)
    (d / "resumo.dat").unlink(missing_ok=True)
    write_box(d, c["nx"])
    (d / "operador.txt").write_text(c["op"] + "\n")
    (d / "velocity.txt").write_text(fmt(c["v0"]) + "\n")
    if c["pq"]:
        (d / "inclinacao.txt").write_text(f"{c['pq'][0]} {c['pq'][1]}\n")
    (d / "data_in.txt").write_text(DATA_IN.format(steps=c["steps"], tau=c["tau"], A=c["A"], beta=BETA))
    options = []
    if precision != "f64":
        options.append(f"precisao {precision}")
    opt = d / "itolbm.txt"
    if options:
        opt.write_text("\n".join(options) + "\n")
    else:
        opt.unlink(missing_ok=True)
    return d


def done(c, root):
    r = root / c["folder"] / "resumo.dat"
    return r.exists() and len(r.read_text().strip().splitlines()) >= 2


def state_of(d):
    r = d / "resumo.dat"
    if not r.exists():
        return "?"
    lines = r.read_text().strip().splitlines()
    if len(lines) < 2:
        return "?"
    head = lines[0].lstrip("#").split()
    vals = lines[1].split()
    return vals[head.index("estado")] if "estado" in head else "?"


def cost(c):
    return c["nx"] * c["nx"] * c["steps"]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("lote", choices=list(BATCHES) + ["todos"])
    ap.add_argument("--saida", default="Lotes", help="root folder of the cases (default: Lotes)")
    ap.add_argument("--filtro", default="", help="only cases whose path contains this text")
    ap.add_argument("--passos", type=int, default=0, help="forces the number of steps (quick test)")
    ap.add_argument("--precisao", default="f64", choices=["f64", "f32", "f16", "f32z", "f16z"],
                    help="storage form on the GPU (f32z / f16z: z-symmetric storage)")
    ap.add_argument("--grupo", type=int, default=0, help="at most this many cases per GPU invocation (0: all)")
    ap.add_argument("--lista", action="store_true", help="only list the cases")
    ap.add_argument("--refaz", action="store_true", help="redo cases already done")
    a = ap.parse_args()

    names = ALL if a.lote == "todos" else [a.lote]
    cases = [c for n in names for c in BATCHES[n]()]
    if a.filtro:
        cases = [c for c in cases if a.filtro in str(c["folder"])]
    if a.passos:
        for c in cases:
            c["steps"] = a.passos
    root = Path(a.saida) if Path(a.saida).is_absolute() else Path.cwd() / a.saida
    pending = [c for c in cases if a.refaz or not done(c, root)]
    pending.sort(key=cost, reverse=True)
    total = sum(cost(c) for c in pending)
    print(f"{len(cases)} cases requested, {len(pending)} to run ({total / 1e9:.2f} G lattice updates)")
    if a.lista:
        for c in pending:
            print(f"   {str(c['folder']):45s}  '{c['op']}'  fat_R_B={c['A']:.6g}  nx={c['nx']}  passos={c['steps']}")
        return 0
    if not pending:
        print("nothing to do.")
        return 0
    for prog in sorted(set(c["program"] for c in pending)):
        if not PROGRAMS[prog].exists():
            sys.exit(f"{PROGRAMS[prog]} is missing: build with make -C {HERE}")
    root.mkdir(parents=True, exist_ok=True)
    rc_all = 0
    with open(root / "progresso.log", "a") as log:
        for prog in sorted(set(c["program"] for c in pending)):
            group = [c for c in pending if c["program"] == prog]
            size = a.grupo if a.grupo > 0 else len(group)
            for k in range(0, len(group), size):
                part = group[k:k + size]
                dirs = [build(c, root, a.precisao) for c in part]
                t0 = time.time()
                r = subprocess.run([str(PROGRAMS[prog])] + [str(d) for d in dirs], capture_output=True, text=True)
                dt = time.time() - t0
                lups = sum(cost(c) for c in part)
                states = [state_of(d) for d in dirs]
                msg = (f"{prog}: {len(part)} cases in {dt:.1f} s ({lups / max(dt, 1e-9) / 1e6:.0f} MLUPS), rc={r.returncode}, "
                       f"{states.count('estavel')} estavel, {states.count('instavel')} instavel, {states.count('?')} without resumo.dat")
                print(msg, flush=True)
                if r.returncode != 0:
                    print(r.stderr.strip()[-2000:], file=sys.stderr)
                    rc_all = 1
                log.write(time.strftime("%Y-%m-%d %H:%M:%S ") + msg + "\n")
                log.flush()
    return rc_all


if __name__ == "__main__":
    sys.exit(main())
