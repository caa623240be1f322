#!/usr/bin/env python3
"""Writes tests/golden/driver_cases.json: complete input and output files of committed reference runs.

The reference tree holds, for every case of its batches (Lotes/<batch>/<case>/), the input files its
programs read and the output files they wrote.  The cases below are copied as data (file texts) so
that the GPU drop-in programs can be run on the same inputs and their outputs compared line by line:

  E1/F3/R=10                        40^2 drop, F3, 30000 steps
  E2/P3 chi=0.5/tau=1/sigma=0.1     100^2 drop, P3 chi = 0.5, 30000 steps
  E2/F2/tau=1/sigma=0.1             100^2 drop, F2, 30000 steps
  E4v/P1/v0=0.1                     100^2 drop translating at v0 = 0.1, P1, 30000 steps
  E4/F3/tau=0.51/sigma=2            unstable at step 500 (the stop rule and the nan means of resumo.dat)
  E3/F1/pq=1,0                      200^2 band, normal (1,0), F1, 20000 steps (perfil.dat sampled)
  E3chi/chi=0.5                     200^2 band, normal (2,1), P3 chi = 0.5, 20000 steps (perfil.dat sampled)

meio.vtk is not kept in the reference's case directories: rodar_lote.py writes it as an all-fluid
nx x nx x 1 box; the fixture records nx.

Each case is also run here through the reference programs built from the reference's sources with
-ffp-contract=off (oracle/Makefile: _ref/Imm_OPS_nofma, _ref/Imm_INCL_nofma), on the same inputs, and
their output files are kept under "nofma": the build whose arithmetic the FP64 library reproduces bit
for bit.  The committed Lotes files come from the reference's -O3 -march=native build (contracted
multiply-adds); the two builds' spread is what separates the library's outputs from the Lotes files.

Usage:  python3 tests/golden/make_driver_golden.py [--ref /root/reference]
"""
import argparse
import concurrent.futures as cf
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REFBIN = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref")
CASES = [("E1/F3/R=10", "Imm_OPS", 40), ("E2/P3 chi=0.5/tau=1/sigma=0.1", "Imm_OPS", 100),
         ("E2/F2/tau=1/sigma=0.1", "Imm_OPS", 100), ("E4v/P1/v0=0.1", "Imm_OPS", 100),
         ("E4/F3/tau=0.51/sigma=2", "Imm_OPS", 100), ("E3/F1/pq=1,0", "Imm_INCL", 200),
         ("E3chi/chi=0.5", "Imm_INCL", 200)]
INPUTS = ["data_in.txt", "operador.txt", "velocity.txt", "inclinacao.txt"]
OUTPUTS = ["dat_out.txt", "laplace.dat", "galilean.dat", "iso.dat", "resumo.dat", "inclinada.dat"]
PERFIL_EVERY = 7


def write_box(d, nx):
    with open(os.path.join(d, "meio.vtk"), "w") as f:
        f.write("# vtk DataFile Version 2.0\nGeometria\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {nx} 1\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * nx}\n")
        f.write("SCALARS Geometria float\nLOOKUP_TABLE default\n")
        f.write((" ".join(["1"] * nx) + " \n") * nx)


def collect(d, c):
    outs = {}
    for name in OUTPUTS:
        p = os.path.join(d, name)
        if os.path.exists(p):
            outs[name] = open(p).read()
    p = os.path.join(d, "perfil.dat")
    perfil = None
    if os.path.exists(p):
        lines = open(p).read().splitlines()
        perfil = {"n_lines": len(lines), "every": PERFIL_EVERY,
                  "lines": {str(i): lines[i] for i in range(0, len(lines), PERFIL_EVERY)}}
    return outs, perfil


def run_nofma(c):
    exe = os.path.join(REFBIN, c["program"] + "_nofma")
    with tempfile.TemporaryDirectory() as d:
        write_box(d, c["nx"])
        for name, text in c["inputs"].items():
            with open(os.path.join(d, name), "w") as f:
                f.write(text)
        subprocess.run([exe], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        return collect(d, c)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    out = []
    for label, prog, nx in CASES:
        d = os.path.join(a.ref, "Lotes", label)
        c = {"label": label, "program": prog, "nx": nx, "inputs": {}}
        for name in INPUTS:
            p = os.path.join(d, name)
            if os.path.exists(p):
                c["inputs"][name] = open(p).read()
        c["outputs"], perfil = collect(d, c)
        if perfil:
            c["perfil"] = perfil
        out.append(c)
    with cf.ProcessPoolExecutor(max_workers=min(len(out), os.cpu_count() or 1)) as ex:
        for c, (outs, perfil) in zip(out, ex.map(run_nofma, out)):
            c["nofma"] = {"outputs": outs}
            if perfil:
                c["nofma"]["perfil"] = perfil
    with open(os.path.join(HERE, "driver_cases.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {len(out)} cases")


if __name__ == "__main__":
    main()
