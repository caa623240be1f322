#!/usr/bin/env python3
"""Throughput of the colour-gradient LBM step on one MI355X.

Workload (the 512^2 static-drop configuration): a 512 x 512 x 1 periodic box holding a sharp
cylindrical drop of radius 100 at the centre, tau in {0.6, 1.0}, sigma = 0.1 (A = 9 sigma / 4 tau for
the perturbation operators), beta = 0.8, for all seven operators (P1, P3 chi = -0.4 / 0.5 / 2, F1,
F2, F3 xi = 1.569988): 14 cases resident on the GPU.  One bench step advances every case by one
time step.  Inputs are synthetic (generated on the host), resident in HBM before timing.

Printed: one JSON line with MLUPS (= sum over cases of sites x steps / wall / 1e6, the reference's
own definition, Imm_OPS.cpp:377), the roofline of the dominant step kernel (algorithmic bytes per
launch over its HIP-event-timed average duration) and a CPU baseline: the reference's own program
(built from its sources by oracle/Makefile) timed on this host's cores over a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

OPS = [("P1", "P"), ("P3 -0.4", "P"), ("P3 0.5", "P"), ("P3 2", "P"), ("F1", "F1"), ("F2", "F2"),
       ("F3 1.569988", "F3")]
TAUS = (0.6, 1.0)
SIGMA = 0.1
NX = NY = 512
R = 100.0
PEAK_HBM_GBS = 8000.0
KERNEL_NAMES = {"P": "step_ring1<false> (P1/P3)", "F3": "step_ring1<true> (F3)", "F1": "step_ring2<F1>",
                "F2": "step_ring2<F2>"}


def amplitude(kind, tau):
    return 9.0 * SIGMA / (4.0 * tau) if kind == "P" else SIGMA


def bytes_per_lu(prec):
    # 2 colours x 19 populations x (1 read + 1 write) x element size
    return {"f64": 608}[prec]


def cpu_baseline(max_threads):
    """The reference program (oracle/_ref/harness_drop_fma, compiled from the reference sources with its
    own optimisation flags) on this host's cores: 512^2 P3 chi = 0.5 drop, bounded number of steps."""
    exe = os.path.join(ROOT, "oracle", "_ref", "harness_drop_fma")
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    if os.path.exists(exe):
        steps = 200
        with tempfile.TemporaryDirectory() as d:
            t0 = time.perf_counter()
            r = subprocess.run([exe, "mode=drop", f"nx={NX}", f"ny={NY}", f"R={R}", "width=-1", "op=P3 0.5",
                                "tau=1.0", f"A={amplitude('P', 1.0)}", "beta=0.8", f"steps={steps}",
                                f"diag={10 * steps}", f"threads={threads}", f"out={d}"],
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            dt = time.perf_counter() - t0
        if r.returncode == 0:
            return {"value": NX * NY * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "reference",
                    "sample": f"reference program (OpenMP, {threads} threads): 512x512 sharp drop R=100, P3 chi=0.5, "
                              f"tau=1, {steps} steps incl. 2 diagnostics, wall {dt:.1f} s"}
    # fallback: the oracle (scalar C restatement), one core
    import numpy as np
    import _oracle as O
    cs = O.make_case(NX, NY, "P3 0.5", 1.0, amplitude("P", 1.0))
    s = O.Sim(cs)
    s.init_drop(R, width=-1.0)
    steps = 8
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    dt = time.perf_counter() - t0
    return {"value": NX * NY * steps / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": f"oracle C restatement, 1 thread: 512x512 P3 chi=0.5, {steps} steps, wall {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--prec", default="f64", choices=["f64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import _pkg
    m = _pkg.load()

    cases, kinds = [], []
    for tau in TAUS:
        for op, kind in OPS:
            cases.append(m.case_desc(NX, NY, op, tau, amplitude(kind, tau), 0.8))
            kinds.append(kind)
    ctx = m.Context(cases)
    for i in range(len(cases)):
        ctx.init_drop(R, width=-1.0, case=i)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)

    n_sites = NX * NY
    for i in range(len(cases)):
        ctx.step(a.warmup, case=i)
    ctx.sync()
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in cases]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(cases)):
        ev[i][0].record(stream)
        ctx.step(a.steps, case=i)
        ev[i][1].record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per_launch_ms = [ev[i][0].elapsed_time(ev[i][1]) / a.steps for i in range(len(cases))]
    kstats = {}
    for k, t in zip(kinds, per_launch_ms):
        e = kstats.setdefault(k, {"launches_per_step": 0, "ms_per_launch": [], "total_ms": 0.0})
        e["launches_per_step"] += 1
        e["ms_per_launch"].append(t)
        e["total_ms"] += t * a.steps
    dom = max(kstats, key=lambda k: kstats[k]["total_ms"])
    bpl = bytes_per_lu(a.prec) * n_sites
    kernels = {}
    for k, e in kstats.items():
        avg = sum(e["ms_per_launch"]) / len(e["ms_per_launch"])
        kernels[KERNEL_NAMES[k]] = {"launches_per_step": e["launches_per_step"], "avg_launch_us": avg * 1e3,
                                    "achieved_GBs": bpl / (avg * 1e-3) / 1e9,
                                    "share_of_time": e["total_ms"] / sum(x["total_ms"] for x in kstats.values())}
    dom_avg_ms = sum(kstats[dom]["ms_per_launch"]) / len(kstats[dom]["ms_per_launch"])
    achieved = bpl / (dom_avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{a.prec}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom)

    mlups = len(cases) * n_sites * a.steps / wall / 1e6
    out = {
        "metric": "MLUPS + % HBM roofline, 1 GPU; sweep cases/hour vs plain FP64 at same Δp error",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.prec,
        "data": "synthetic",
        "config": {"workload": "512x512x1 static sharp drop R=100; tau {0.6, 1.0} x 7 operators "
                               "(P1, P3 chi -0.4/0.5/2, F1, F2, F3 xi 1.569988), sigma 0.1, beta 0.8; "
                               f"{a.prec.upper()} one-case-per-launch path",
                   "cases": len(cases), "sites_per_case": n_sites, "bytes_per_lattice_update": bytes_per_lu(a.prec)},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic},
        "kernels": kernels,
    }
    if not a.no_cpu_baseline:
        cb = cpu_baseline(a.cpu_threads)
        out["cpu_baseline"] = cb
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
