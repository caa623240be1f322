#!/usr/bin/env python3
"""Throughput of the colour-gradient LBM step on one MI355X.

Workload (the 512^2 static-drop configuration): a 512 x 512 x 1 periodic box holding a sharp
cylindrical drop of radius 100 at the centre, tau in {0.6, 1.0}, sigma = 0.1 (A = 9 sigma / 4 tau for
the perturbation operators), beta = 0.8, for all seven operators (P1, P3 chi = -0.4 / 0.5 / 2, F1,
F2, F3 xi = 1.569988): 14 cases resident on the GPU, advanced together (one launch per operator
family per step).  A bench step advances every case by one time step.  Inputs are synthetic
(generated by the library's initial condition), resident in HBM before timing.

The headline form (--prec, default f32z) is timed over --steps steps after --warmup; the other
storage forms are timed the same way and reported beside it:

  f64    FP64 compute and storage (the drop-in form, bit-equal to the reference's arithmetic)
  f32    FP32 compute, FP32 storage of all 19 populations per site
  f16    FP32 compute, FP16 storage of all 19
  f32z   as f32, storing the 14 distinct populations of a state that is mirror-symmetric in z (with
         one node in z every state the programs create is; the results are bitwise those of f32)
  f16z   as f16, likewise

Also reported:

  roofline       the dominant kernel (the P1/P3 family launch, 8 of the 14 cases) timed alone with HIP
                 events on its stream: algorithmic bytes per launch (the form's stored bytes) / average
                 launch duration; frac_at_304B = the kernel's lattice updates per second x 304 B (the
                 full FP32 state's bytes per update) / peak, the MLUPS-based roofline fraction;
                 traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
                 (profiles/traffic_<form>.json), when present; copy_GBs = a plain 1 GiB device copy
                 (torch's copy kernel) timed the same way, the bandwidth a streaming kernel reaches
                 on this GPU, and frac_of_copy = achieved / copy_GBs
  large          the 8192 x 8192 translating drop (R = 1000, v0 = 0.1, P3 chi = 0.5) in the headline form
  sweep          the 168-case static-drop sweep (tau x sigma x 7 operators, R ~ U{10..50} from
                 numpy default_rng(2026), box 4R, initial width 3.45 as the reference programs'):
                 batched reduced-precision form vs the plain FP64
                 one-launch-per-case path, time per step and delta_p agreement after --sweep-steps
  cpu_baseline   the reference program (built from the reference's sources by oracle/Makefile) on
                 this host's cores, over a bounded sample of the same workload

MLUPS = sum over cases of sites x steps / wall / 1e6 (the reference's definition, Imm_OPS.cpp:377).
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
SWEEP_WARMUP = 400   # untimed sweep steps (the batched path tunes its tile shapes on them)
B_LU = {"f64": 608, "f32": 304, "f16": 152, "f32z": 224, "f16z": 112}
PREC_ID = {"f64": 0, "f32": 1, "f16": 2, "f32z": 3, "f16z": 4}
KERNEL = {"f64": "step_ring1<false> (P1/P3 family)", "f32": "lp_pipe<float,P,depth,false> (P1/P3 family)",
          "f16": "lp_pipe<_Float16,P,depth,false> (P1/P3 family)",
          "f32z": "lp_pipe<float,P,depth,true> (P1/P3 family)",
          "f16z": "lp_pipe<_Float16,P,depth,true> (P1/P3 family)"}
FORM_TEXT = {"f64": "f64 storage", "f32": "f32 storage of all 19 populations, fp32 compute",
             "f16": "f16 storage of all 19 populations, fp32 compute",
             "f32z": "f32 storage of the 14 distinct populations of the z-mirror-symmetric slab states "
                     "(bitwise equal to full f32 storage), fp32 compute",
             "f16z": "f16 storage of the 14 distinct populations of the z-mirror-symmetric slab states, fp32 compute"}


def amplitude(kind, tau, sigma=SIGMA):
    return 9.0 * sigma / (4.0 * tau) if kind == "P" else sigma


def say(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP events on the library's stream (a torch stream handed to the context)."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def run(self, fn):
        t = self.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        t.cuda.synchronize()
        w0 = time.perf_counter()
        e0.record(self.stream)
        fn()
        e1.record(self.stream)
        self.stream.synchronize()
        t.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3, time.perf_counter() - w0


def bench_form(m, timer, stream, form, steps, warmup):
    cases, kinds = [], []
    for tau in TAUS:
        for op, kind in OPS:
            cases.append(m.case_desc(NX, NY, op, tau, amplitude(kind, tau), 0.8))
            kinds.append(kind)
    ctx = m.Context(cases, precision=PREC_ID[form])
    ctx.set_stream(stream.cuda_stream)
    for i in range(len(cases)):
        ctx.init_drop(R, width=-1.0, case=i)
    ctx.step(warmup)
    gpu_s, wall_s = timer.run(lambda: ctx.step(steps))
    n_sites = NX * NY * len(cases)
    # the dominant kernel alone: the P1/P3 family launch (8 cases), same launch shape
    for i, k in enumerate(kinds):
        if k != "P":
            ctx.set_active(False, case=i)
    nP = sum(1 for k in kinds if k == "P")
    ctx.step(5)
    kern_s, _ = timer.run(lambda: ctx.step(steps))
    ctx.close()
    avg = kern_s / steps
    bytes_launch = B_LU[form] * NX * NY * nP
    return {"MLUPS": n_sites * steps / wall_s / 1e6, "ms_per_step": wall_s / steps * 1e3,
            "gpu_ms_per_step": gpu_s / steps * 1e3,
            "dominant_kernel": KERNEL[form], "kernel_cases_per_launch": nP,
            "kernel_avg_launch_us": avg * 1e6, "bytes_per_launch": bytes_launch,
            "achieved_GBs": bytes_launch / avg / 1e9, "frac": bytes_launch / avg / 1e9 / PEAK_HBM_GBS,
            "frac_at_304B": 304 * NX * NY * nP / avg / 1e9 / PEAK_HBM_GBS,
            "bytes_per_lattice_update": B_LU[form]}


def bench_large(m, timer, stream, form, steps):
    n = 8192
    ctx = m.Context(m.case_desc(n, n, "P3 0.5", 1.0, amplitude("P", 1.0), 0.8), precision=PREC_ID[form])
    ctx.set_stream(stream.cuda_stream)
    ctx.init_drop(1000.0, width=-1.0, v0=0.1)
    ctx.step(100)   # untimed: the library tunes the tile shape on these steps
    gpu_s, wall_s = timer.run(lambda: ctx.step(steps))
    ctx.close()
    avg = gpu_s / steps
    return {"workload": f"{n}x{n}x1 translating sharp drop R=1000 v0=0.1, P3 chi=0.5, tau=1, {form} storage",
            "steps": steps, "MLUPS": n * n * steps / wall_s / 1e6, "kernel_avg_launch_us": avg * 1e6,
            "achieved_GBs": B_LU[form] * n * n / avg / 1e9, "frac": B_LU[form] * n * n / avg / 1e9 / PEAK_HBM_GBS,
            "frac_at_304B": 304 * n * n / avg / 1e9 / PEAK_HBM_GBS}


def copy_ceiling(torch, timer, stream, reps=10):
    """Plain device copy of 1 GiB (torch's copy kernel) on the same stream: read + write bandwidth a
    simple streaming kernel reaches on this GPU, beside the nominal peak."""
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    with torch.cuda.stream(stream):
        for _ in range(3):
            y.copy_(x)

        def run():
            for _ in range(reps):
                y.copy_(x)
        gpu_s, _ = timer.run(run)
    nbytes = 2 * x.numel() * x.element_size()
    del x, y
    return {"bytes_per_copy": nbytes, "GBs": nbytes * reps / gpu_s / 1e9}


def sweep_cases(m):
    import numpy as np
    rng = np.random.default_rng(2026)
    radii = rng.integers(10, 51, 168)
    cases, k = [], 0
    for tau in (0.51, 0.6, 0.8, 1.0, 1.5, 2.0):
        for sigma in (0.01, 0.05, 0.1, 0.5):
            for op, kind in OPS:
                r = int(radii[k])
                cases.append((m.case_desc(4 * r, 4 * r, op, tau, amplitude(kind, tau, sigma), 0.8), float(r),
                              f"{op} tau={tau} sigma={sigma} R={r}"))
                k += 1
    return cases


def bench_sweep(m, timer, stream, form, steps):
    """168 cases: plain FP64 one launch per case and step vs the batched form, `steps` steps each."""
    import numpy as np
    cases = sweep_cases(m)
    descs = [c for c, _, _ in cases]
    sites = sum(d.nx * d.ny for d in descs)
    res = {}
    for label, prec in (("f64_per_case", 0), (f"{form}_batched", PREC_ID[form])):
        ctx = m.Context(descs, precision=prec)
        ctx.set_stream(stream.cuda_stream)
        for i, (d, r, _) in enumerate(cases):
            ctx.init_drop(r, width=3.45, case=i)
        # untimed warm-up (first launches, tile-shape tuning), the same steps on both paths
        if prec == 0:
            for i in range(len(descs)):
                ctx.step(SWEEP_WARMUP, case=i)
        else:
            ctx.step(SWEEP_WARMUP)
        ctx.sync()
        if prec == 0:
            # one launch per case and step (cases are independent, so case by case is the same work)
            def run():
                for i in range(len(descs)):
                    ctx.step(steps - 1, case=i)
                    ctx.step(1, record=True, case=i)
        else:
            def run():
                ctx.step(steps - 1)
                ctx.step(1, record=True)
        gpu_s, wall_s = timer.run(run)
        diags = [ctx.diag_drop(0.0, steps - 1, case=i) for i in range(len(descs))]
        dp = np.array([g.delta_p for g in diags])
        ok = np.array([g.ok for g in diags])
        ctx.close()
        res[label] = {"wall_s": wall_s, "ms_per_step": wall_s / steps * 1e3, "dp": dp, "ok": ok}
    a, b = res["f64_per_case"], res[f"{form}_batched"]
    good = (a["ok"] == 1) & (b["ok"] == 1) & np.isfinite(a["dp"]) & np.isfinite(b["dp"]) & (np.abs(a["dp"]) > 1e-9)
    err = np.abs(b["dp"][good] - a["dp"][good]) / np.abs(a["dp"][good])
    labels = [lab for (_, _, lab), g in zip(cases, good) if g]
    per_case_20k = lambda ms: 168 * 3600.0 / (ms * 1e-3 * 20000)
    return {"cases": len(descs), "sites": sites, "steps": steps, "warmup": SWEEP_WARMUP, "form": form,
            "f64_per_case_ms_per_step": a["ms_per_step"], "batched_ms_per_step": b["ms_per_step"],
            "speedup": a["ms_per_step"] / b["ms_per_step"],
            "cases_per_hour_20k_steps": {"f64_per_case": per_case_20k(a["ms_per_step"]),
                                         f"{form}_batched": per_case_20k(b["ms_per_step"])},
            "dp_compared_cases": int(good.sum()),
            "dp_rel_err_median": float(np.median(err)) if err.size else None,
            "dp_rel_err_max": float(err.max()) if err.size else None,
            "dp_within_1e-4": int((err <= 1e-4).sum()),
            "dp_worst_cases": [[labels[j], float(err[j])] for j in np.argsort(err)[::-1][:3]] if err.size else []}


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
    # the reference build is absent: the oracle (scalar C restatement), one core
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
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=450,
                    help="untimed steps first (the library tunes its tile shapes on the first ~360 of them)")
    ap.add_argument("--prec", default="f32z", choices=list(B_LU), help="headline storage form")
    ap.add_argument("--forms", default="f64,f32,f16,f16z", help="forms measured beside the headline")
    ap.add_argument("--sweep-steps", type=int, default=1000, help="0 skips the 168-case sweep")
    ap.add_argument("--large-steps", type=int, default=20, help="0 skips the 8192^2 case")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-only", action="store_true",
                    help="only the 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import _pkg
    m = _pkg.load()
    stream = torch.cuda.Stream()
    timer = Timer(torch, stream)
    if a.copy_only:
        print(json.dumps({"copy": copy_ceiling(torch, timer, stream, reps=3)}), flush=True)
        return

    forms = {}
    for form in [a.prec] + [f for f in a.forms.split(",") if f and f != a.prec]:
        say(f"bench: form {form}, 14 cases x 512^2, {a.steps} steps")
        forms[form] = bench_form(m, timer, stream, form, a.steps, a.warmup)
    head = forms[a.prec]

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{a.prec}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    say("bench: 1 GiB device copy")
    copy = copy_ceiling(torch, timer, stream)
    out = {
        "metric": "MLUPS + % HBM roofline, 1 GPU; sweep cases/hour vs plain FP64 at same Δp error",
        "value": head["MLUPS"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.prec == "f64" else "f32",
        "data": "synthetic",
        "config": {"workload": "512x512x1 static sharp drop R=100; tau {0.6, 1.0} x 7 operators "
                               "(P1, P3 chi -0.4/0.5/2, F1, F2, F3 xi 1.569988), sigma 0.1, beta 0.8; 14 cases "
                               f"batched; {FORM_TEXT[a.prec]}",
                   "cases": 14, "sites_per_case": NX * NY, "storage": a.prec,
                   "bytes_per_lattice_update": B_LU[a.prec]},
        "roofline": {"bound": "hbm", "kernel": head["dominant_kernel"], "achieved": head["achieved_GBs"],
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": head["frac"], "traffic": traffic,
                     "frac_at_304B": head["frac_at_304B"],
                     "algorithmic_bytes_per_launch": head["bytes_per_launch"],
                     "avg_launch_us": head["kernel_avg_launch_us"],
                     "copy_GBs": copy["GBs"], "frac_of_copy": head["achieved_GBs"] / copy["GBs"]},
        "forms": forms,
    }
    if a.large_steps > 0:
        say(f"bench: 8192^2 {a.prec}, {a.large_steps} steps")
        out["large"] = bench_large(m, timer, stream, a.prec if a.prec != "f64" else "f32z", a.large_steps)
    if a.sweep_steps > 0:
        sform = a.prec if a.prec != "f64" else "f32z"
        say(f"bench: 168-case sweep, {a.sweep_steps} steps, f64 per case vs {sform} batched")
        out["sweep"] = bench_sweep(m, timer, stream, sform, a.sweep_steps)
    if not a.no_cpu_baseline:
        say("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(a.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
