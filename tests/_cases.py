"""Case directories for the drop-in programs (Imm_OPS, Imm_INCL), laid out as the reference's batch
script does (rodar_lote.py:176-206): data_in.txt, operador.txt, velocity.txt, [inclinacao.txt] and an
all-fluid meio.vtk."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "interfacial-tension-operators-lbm_amd")

DATA_IN = """meio.vtk

Dim_Pixel\t1.0

steps:\t    {passos}

files:\t    0

n_threads:  1

tau:\t    {tau}

tau_nh:     1.0

rho:\t    1.0

tau_R:\t    {tau}

tau_B:\t    {tau_B}

tau_m:\t    1.0

rho_R:\t    1.0

rho_B:\t    1.0

fat_R_B:\t{A:.12g}

fat_R_R:\t0.0

fat_B_B:    0.0

recoll:\t    {beta}

wett_R:\t    0.0
"""


def program(name):
    path = os.path.join(PKG, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build with make -C {PKG}")
    return path


def write_box(d, nx, ny=None, solid=None):
    ny = nx if ny is None else ny
    with open(os.path.join(d, "meio.vtk"), "w") as f:
        f.write("# vtk DataFile Version 2.0\nGeometria\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {ny} 1\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * ny}\n")
        f.write("SCALARS Geometria float\nLOOKUP_TABLE default\n")
        for y in range(ny):
            f.write(" ".join("0" if solid == (x, y) else "1" for x in range(nx)) + " \n")


def make_case(d, nx, op, tau, A, passos, v0=0.0, beta=0.8, pq=None, ny=None, tau_B=None, options=None):
    os.makedirs(d, exist_ok=True)
    write_box(d, nx, ny)
    with open(os.path.join(d, "operador.txt"), "w") as f:
        f.write(op + "\n")
    with open(os.path.join(d, "velocity.txt"), "w") as f:
        f.write(f"{v0:g}\n")
    if pq:
        with open(os.path.join(d, "inclinacao.txt"), "w") as f:
            f.write(f"{pq[0]} {pq[1]}\n")
    with open(os.path.join(d, "data_in.txt"), "w") as f:
        f.write(DATA_IN.format(passos=passos, tau=tau, tau_B=tau if tau_B is None else tau_B, A=A, beta=beta))
    if options:
        with open(os.path.join(d, "itolbm.txt"), "w") as f:
            for k, v in options.items():
                f.write(f"{k} {v}\n")


def golden_cases():
    with open(os.path.join(HERE, "golden", "driver_cases.json")) as f:
        return json.load(f)


def make_golden_case(d, c, steps=None):
    """Writes a committed reference case's inputs into d (optionally with fewer steps)."""
    os.makedirs(d, exist_ok=True)
    write_box(d, c["nx"])
    for name, text in c["inputs"].items():
        if name == "data_in.txt" and steps is not None:
            lines = text.splitlines(keepends=True)
            text = "".join(f"steps:\t    {steps}\n" if ln.startswith("steps:") else ln for ln in lines)
        with open(os.path.join(d, name), "w") as f:
            f.write(text)


def run(prog, args, cwd, timeout=600):
    return subprocess.run([program(prog)] + list(args), cwd=cwd, capture_output=True, text=True, timeout=timeout)


def read_table(path):
    """(header line, rows as lists of strings)."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[0] if lines and lines[0].startswith("#") else None
    rows = [ln.split() for ln in lines[1 if head else 0:] if ln.strip()]
    return head, rows
