#!/usr/bin/env python3
"""Writes tests/golden/golden.json: the fixtures the oracle and the GPU library are pinned on.

Run in a container that holds the reference tree (default /root/reference) after
`make -C oracle ref`.  Three kinds of fixture are collected, all data (inputs + outputs):

  e0      the printed values of the reference's own E0 output (E0_resultado.txt)
  lotes   time-series rows of committed runs in the reference's Lotes/ tree (inputs parsed from
          the case directory: operador.txt, data_in.txt, velocity.txt, inclinacao.txt)
  ref     outputs of the reference's own functions, built from its sources by oracle/Makefile
          (oracle/_ref/harness_*): SHA-256 of the post-collision populations / force field at
          chosen steps and the full-precision diagnostics, for small boxes and all operators,
          with the "nofma" (-ffp-contract=off) build and with the optimised "fma" build.
  ref_geo the same hashes on a box with solid sites (geometry file read by the reference's read_geo;
          fluid sites only, in site order).
  read_geo the numbering read_geo gives the sites of geometry files (with added columns).

Usage:  python3 tests/golden/make_golden.py [--ref /root/reference]
"""
import argparse
import hashlib
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFBIN = os.path.join(ROOT, "oracle", "_ref")

OPS = ["P1", "P3 -0.4", "P3 0.5", "P3 2", "F1", "F2", "F3 1.569988"]

# small parity cases: (label, mode, args)
SMALL = []
for op in OPS:
    A = "0.225" if op[0] == "P" else "0.1"
    SMALL.append((f"drop24x20 {op}", "drop",
                  dict(nx=24, ny=20, R=6, width=3.45, op=op, tau=0.8, A=A, beta=0.8, steps=10, diag=4,
                       dump="0,1,9")))
SMALL.append(("drop sharp 26x22 P3 0.5 v0", "drop",
              dict(nx=26, ny=22, R=7, width=-1, op="P3 0.5", tau=1.0, A=0.225, beta=0.8, v0=0.05, steps=10,
                   diag=3, dump="0,9")))
SMALL.append(("drop sharp 26x22 F2 v0", "drop",
              dict(nx=26, ny=22, R=7, width=-1, op="F2", tau=0.6, A=0.1, beta=0.8, v0=0.05, steps=10,
                   diag=3, dump="0,9")))
SMALL.append(("incl 30 P3 0.5 (2,1)", "incl",
              dict(nx=30, ny=30, p=2, q=1, width=3.45, op="P3 0.5", tau=1.0, A=0.225, beta=0.8, steps=10,
                   diag=3, dump="0,9")))
SMALL.append(("incl 30 F1 (1,1)", "incl",
              dict(nx=30, ny=30, p=1, q=1, width=3.45, op="F1", tau=1.0, A=0.1, beta=0.8, steps=10,
                   diag=3, dump="0,9")))
# a normal that is not coprime: the reference's band period is N (Imm_INCL.cpp:314-315, "% N")
SMALL.append(("incl 30 P3 0.5 (2,2)", "incl",
              dict(nx=30, ny=30, p=2, q=2, width=3.45, op="P3 0.5", tau=1.0, A=0.225, beta=0.8, steps=10,
                   diag=3, dump="0,9")))


# boxes with solid sites (the reference library's bounce-back, wall mediators and force-stencil
# reflection; its drop and band programs never use them): a 24 x 20 box with a solid bottom row, a
# 3 x 2 block, single sites and one solid site inside the drop's interface; all operators, two wall
# wettabilities.  Populations are dumped for the fluid sites only (the reference stores no others).
GEO_SOLIDS = ([[x, 0] for x in range(24)] + [[x, y] for x in range(4, 7) for y in range(12, 14)]
              + [[18, 5], [20, 15], [21, 16], [16, 10]])
GEO = []
for k, op in enumerate(OPS):
    A = "0.225" if op[0] == "P" else "0.1"
    GEO.append((f"geo24x20 {op}", dict(nx=24, ny=20, R=5, width=3.45 if k % 2 == 0 else -1, op=op, tau=0.8, A=A,
                                         beta=0.8, wett=0.3 if k % 2 == 0 else -0.5, steps=8, diag=0, dump="0,7",
                                         solid=GEO_SOLIDS)))


def write_geometry(path, nx, ny, solids):
    """meio.vtk in the layout read_geo reads (LBM_functions.cpp:280-354): 10 header lines, then nx * ny
    integers in site order, 0 = solid."""
    m = [[1] * nx for _ in range(ny)]
    for x, y in solids:
        m[y][x] = 0
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\nGeometria\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {ny} 1\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * ny}\n")
        f.write("SCALARS Geometria float\nLOOKUP_TABLE default\n")
        for row in m:
            f.write(" ".join(str(v) for v in row) + " \n")


def write_values(path, dims, values):
    """A geometry file with arbitrary integer values (nx * ny * nz, site order)."""
    nx, ny, nz = dims
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\nGeometria\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {ny} {nz}\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * ny * nz}\n")
        f.write("SCALARS Geometria float\nLOOKUP_TABLE default\n")
        for r in range(ny * nz):
            f.write(" ".join(str(v) for v in values[r * nx:(r + 1) * nx]) + " \n")


# geometry files read by the reference's read_geo: its numbering of every site (with added columns)
def readgeo_cases():
    box = [1] * (24 * 20)
    for x, y in GEO_SOLIDS:
        box[x + 24 * y] = 0
    vol = [(3 * i + 1) % 4 for i in range(5 * 4 * 2)]   # 0 = solid; 1, 2, 3 all fluid
    return [("box 24x20 solids", (24, 20, 1), box, 0, 0), ("box 24x20 solids + 2/3 columns", (24, 20, 1), box, 2, 3),
            ("volume 5x4x2", (5, 4, 2), vol, 1, 1)]


def run_readgeo(dims, values, pin, pout):
    exe = os.path.join(REFBIN, "harness_drop_nofma")
    with tempfile.TemporaryDirectory() as d:
        geo = os.path.join(d, "meio.vtk")
        write_values(geo, dims, values)
        subprocess.run([exe, "mode=readgeo", f"geo={geo}", f"pts_in={pin}", f"pts_out={pout}", f"out={d}"],
                       check=True, stdout=subprocess.DEVNULL)
        n, nx, ny, nz = [int(v) for v in open(os.path.join(d, "n.txt")).read().split()]
        return {"n_fluid": n, "box": [nx, ny, nz], "meio_sha": sha(os.path.join(d, "meio.bin"))}


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def run_harness(kind, mode, args):
    exe = os.path.join(REFBIN, f"harness_{'incl' if mode == 'incl' else 'drop'}_{kind}")
    with tempfile.TemporaryDirectory() as d:
        extra = []
        if "solid" in args:
            geo = os.path.join(d, "meio.vtk")
            write_geometry(geo, int(args["nx"]), int(args["ny"]), args["solid"])
            extra = [f"geo={geo}"]
        cmd = ([exe, f"mode={mode}", f"out={d}"] + [f"{k}={v}" for k, v in args.items() if k != "solid"]
               + extra)
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
        lines = open(os.path.join(d, "diag.txt")).read().splitlines()
        names = lines[0][2:].split()
        diag = [dict(zip(names, [float(v) for v in l.split()])) for l in lines[1:]]
        hashes = {}
        for st in [int(s) for s in str(args.get("dump", "")).split(",") if s]:
            ent = {"R": sha(os.path.join(d, f"popR_{st}.bin")), "B": sha(os.path.join(d, f"popB_{st}.bin"))}
            fp = os.path.join(d, f"force_{st}.bin")
            if os.path.exists(fp):
                ent["F"] = sha(fp)
            hashes[str(st)] = ent
    return diag, hashes


def parse_e0(path):
    txt = open(path).read()
    errs = [m.group(1) for m in re.finditer(r"erro =\s+(\S+)\s+(ok|FALHOU)", txt)]
    m4 = []
    for line in txt.splitlines():
        if re.match(r"^\s+P[13].*\s+-?\d\.\d{6}\s", line) and "lambda" not in line:
            m4.append([float(v) for v in re.findall(r"-?\d+\.\d{6}", line)])
    return {"erro": errs, "m4": m4, "tudo_certo": "TUDO CERTO" in txt}


def read_case(d):
    dat = open(os.path.join(d, "data_in.txt")).read().split()
    vals = {dat[i].rstrip(":"): dat[i + 1] for i in range(1, len(dat) - 1, 2)}
    case = {
        "op": open(os.path.join(d, "operador.txt")).read().strip(),
        "tau": float(vals["tau_R"]), "A": float(vals["fat_R_B"]), "beta": float(vals["recoll"]),
    }
    vf = os.path.join(d, "velocity.txt")
    case["v0"] = float(open(vf).read().split()[0]) if os.path.exists(vf) else 0.0
    inc = os.path.join(d, "inclinacao.txt")
    if os.path.exists(inc):
        p, q = open(inc).read().split()[:2]
        case["pq"] = [int(p), int(q)]
    res = open(os.path.join(d, "resumo.dat")).read().splitlines()
    cab = res[0].lstrip("#").split()
    row = res[1].split()
    case["nx"] = int(float(row[cab.index("nx" if "nx" in cab else "N")]))
    return case


def table(path, steps):
    lines = open(path).read().splitlines()
    head = lines[0].lstrip("#").split()
    rows = {}
    for l in lines[1:]:
        v = l.split()
        if int(float(v[0])) in steps:
            rows[str(int(float(v[0])))] = v
    return {"header": head, "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    out = {"e0": parse_e0(os.path.join(a.ref, "E0_resultado.txt")), "lotes": {}, "ref": []}

    lotes = [
        ("E2/P3 chi=0.5/tau=1/sigma=0.1", ["galilean.dat", "laplace.dat", "iso.dat"], [0, 500, 1000]),
        ("E2/F2/tau=1/sigma=0.1", ["galilean.dat", "laplace.dat", "iso.dat"], [0, 500, 1000]),
        ("E4v/P1/v0=0.1", ["galilean.dat", "laplace.dat", "iso.dat"], [0, 500, 1000]),
        ("E1/F3/R=10", ["galilean.dat", "laplace.dat", "iso.dat"], [0, 500, 1000]),
        ("E3chi/chi=0.5", ["inclinada.dat"], [0, 500]),
        ("E3/F1/pq=1,0", ["inclinada.dat"], [0, 500]),
    ]
    for rel, files, steps in lotes:
        d = os.path.join(a.ref, "Lotes", rel)
        ent = {"case": read_case(d), "tables": {}}
        for f in files:
            ent["tables"][f] = table(os.path.join(d, f), set(steps))
        out["lotes"][rel] = ent

    for label, mode, args in SMALL:
        diag_n, hash_n = run_harness("nofma", mode, args)
        diag_f, _ = run_harness("fma", mode, args)
        out["ref"].append({"label": label, "mode": mode, "args": args, "diag_nofma": diag_n,
                           "hash_nofma": hash_n, "diag_fma": diag_f})

    out["read_geo"] = []
    for label, dims, values, pin, pout in readgeo_cases():
        out["read_geo"].append({"label": label, "dims": list(dims), "values": values, "pts_in": pin, "pts_out": pout,
                                **run_readgeo(dims, values, pin, pout)})
    out["ref_geo"] = []
    for label, args in GEO:
        _, hash_n = run_harness("nofma", "drop", args)
        out["ref_geo"].append({"label": label, "mode": "drop", "args": args, "hash_nofma": hash_n})

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
