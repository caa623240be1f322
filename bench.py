#!/usr/bin/env python3
"""Throughput of the colour-gradient LBM step on one MI355X.

Workload (the 512^2 static-drop configuration): a 512 x 512 x 1 periodic box holding a sharp
cylindrical drop of radius 100 at the centre, tau in {0.6, 1.0}, sigma = 0.1 (A = 9 sigma / 4 tau for
the perturbation operators), beta = 0.8, for all seven operators (P1, P3 chi = -0.4 / 0.5 / 2, F1,
F2, F3 xi = 1.569988): 14 cases resident on the GPU, advanced together (in the launch form the
library chooses by timing: one launch per operator family per step, or, FP64 only, one launch for all
of them, reported as launch_form).  A bench step advances every case by one time step.  Inputs are synthetic
(generated by the library's initial condition), resident in HBM before timing.

Before anything is timed, 2 s of untimed streaming copies bring the GPU to its sustained clocks.
Every form first runs --tune-steps untimed steps (default 500, whatever --warmup is): the library
tunes its tile shapes on the first ~360-450 of them, and the shapes it chose are reported.  Then
--warmup untimed steps, then --steps timed steps.  The library replays runs of 32 steps from a
captured HIP graph; when --steps is shorter, the graph is sized to --steps and captured on one
untimed call of that length first, so the timed call replays it as a longer run does.  The headline (--prec, default f64) is the FP64
drop-in form, whose arithmetic is the reference's; the other storage forms are timed the same way
and reported under "forms", each with its delta_p difference from the FP64 form after the same steps:

  f64    FP64 compute and storage (the drop-in form, bit-equal to the reference's arithmetic)
  f32    FP32 compute, FP32 storage of all 19 populations per site
  f16    FP32 compute, FP16 storage of all 19
  f32z   as f32, storing the 14 distinct populations of a state that is mirror-symmetric in z (with
         one node in z every state the programs create is; the results are bitwise those of f32)
  f16z   as f16, likewise

Also reported:

  roofline       the dominant kernel (the P1/P3 family launch, 8 of the 14 cases) timed alone with HIP
                 events on its stream, at the tile shape the library tunes for it alone (after
                 --tune-steps more untimed steps): algorithmic bytes per launch (the form's stored
                 bytes) / average launch duration; traffic = HBM bytes per launch from rocprofv3
                 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic_<form>.json), when present;
                 stream_GBs = the library's own streaming kernel (read + write of 1 GiB, 16-byte
                 loads, nontemporal stores) timed the same way, the bandwidth a plain streaming
                 kernel reaches on this GPU
  large          the 8192 x 8192 translating drop (R = 1000, v0 = 0.1, P3 chi = 0.5) in FP32 (all 19
                 populations, 304 B per update), FP32 z-symmetric and FP64 storage
  shard_8gpu     the same forms on an 8192 x 1024 box (R = 300): one GPU's y-slab when the 8192^2 box
                 is split over 8 GPUs, the kernel at that shape (the slabs' halo-row exchange is not
                 part of it: this build runs on one GPU)
  sweep          the 168-case static-drop sweep (tau x sigma x 7 operators, R ~ U{10..50} from
                 numpy default_rng(2026), box 4R, initial width 3.45 as the reference programs'):
                 batched reduced-precision form vs the plain FP64 one-launch-per-case path over
                 --sweep-steps steps in all (the first 500 untimed), delta_p agreement at the end;
                 FP64 batched the same way (the same bits as one launch per case: delta_p equal);
                 FP16 z-symmetric storage batched the same way, for its delta_p agreement
  cpu_baseline   the reference program Imm_OPS (built from the reference's sources by oracle/Makefile)
                 on this host's cores, over a bounded sample of the same workload

MLUPS = sum over cases of sites x steps / wall / 1e6 (the reference's definition, Imm_OPS.cpp:377).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

OPS = [("P1", "P"), ("P3 -0.4", "P"), ("P3 0.5", "P"), ("P3 2", "P"), ("F1", "F1"), ("F2", "F2"),
       ("F3 1.569988", "F3")]
FAMILY_NAMES = ("P1/P3", "F3", "F1", "F2")
TAUS = (0.6, 1.0)
SIGMA = 0.1
NX = NY = 512
R = 100.0
PEAK_HBM_GBS = 8000.0
SWEEP_WARMUP = 500   # untimed sweep steps (the batched path tunes its tile shapes on them)
LARGE_TUNE = 120     # untimed steps of the 8192^2 case (tuning)
# storage forms; the bytes a step moves per lattice update come from the library for each case
# (lbm_stored_bytes): FP64 384 (the compact form of the P1/P3 and F3 cases) or 608, FP32 304, FP32
# z-symmetric 224, FP16 152, FP16 z-symmetric 112
FORMS = ("f64", "f32", "f16", "f32z", "f16z")
PREC_ID = {"f64": 0, "f32": 1, "f16": 2, "f32z": 3, "f16z": 4}
KERNEL = {"f64": "step_ring1c<false> (P1/P3 family, compact FP64 form)", "f32": "lp_pipe<float,P,depth,false> (P1/P3 family)",
          "f16": "lp_pipe<_Float16,P,depth,false> (P1/P3 family)",
          "f32z": "lp_pipe<float,P,depth,true> (P1/P3 family)",
          "f16z": "lp_pipe<_Float16,P,depth,true> (P1/P3 family)"}
FORM_TEXT = {"f64": "f64 compute and storage (the drop-in form, the reference's arithmetic)",
             "f32": "f32 storage of all 19 populations, fp32 compute",
             "f16": "f16 storage of all 19 populations, fp32 compute",
             "f32z": "f32 storage of the 14 distinct populations of the z-mirror-symmetric slab states "
                     "(bitwise equal to full f32 storage), fp32 compute",
             "f16z": "f16 storage of the 14 distinct populations of the z-mirror-symmetric slab states, fp32 compute"}
DTYPE = {"f64": "f64", "f32": "f32", "f32z": "f32", "f16": "f32 compute, f16 storage", "f16z": "f32 compute, f16 storage"}


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


class launch_form:
    """Fixes lbm_step's launch form (ITOLBM_FUSED: 0 one launch per family, 1 one launch, 2 one
    persistent launch) while the block runs; None leaves the choice to the library."""

    def __init__(self, mode):
        self.mode, self.old = mode, None

    def __enter__(self):
        self.old = os.environ.get("ITOLBM_FUSED")
        if self.mode is not None:
            os.environ["ITOLBM_FUSED"] = str(self.mode)

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("ITOLBM_FUSED", None)
        else:
            os.environ["ITOLBM_FUSED"] = self.old


def prime_graph(ctx, steps):
    """A call of fewer steps than the library's graph length (32) would launch step by step: size the
    graph to the timed call (its even part) and capture it on one untimed call of the same length, so
    the timed call replays it as a longer run does.  Returns the untimed steps run."""
    g = steps - (steps & 1)
    if 2 <= g < 32:
        ctx.set_graph(g)
        ctx.step(steps)
        return steps
    return 0


def shapes_of(ctx, kinds):
    out = {}
    for fam, name in enumerate(FAMILY_NAMES):
        for i, k in enumerate(kinds):
            if ("P1/P3" if k == "P" else k) == name:
                rows, depth, tuned = ctx.tile_shape(case=i)
                out[name] = {"rows_per_segment": rows, "depth": depth, "tuned": tuned}
                break
    return out


def bench_form(m, timer, stream, form, steps, warmup, tune_steps):
    cases, kinds = [], []
    for tau in TAUS:
        for op, kind in OPS:
            cases.append(m.case_desc(NX, NY, op, tau, amplitude(kind, tau), 0.8))
            kinds.append(kind)
    ctx = m.Context(cases, precision=PREC_ID[form])
    ctx.set_stream(stream.cuda_stream)
    for i in range(len(cases)):
        ctx.init_drop(R, width=-1.0, case=i)
    ctx.step(tune_steps)   # untimed: tile shapes are tuned on these steps
    ctx.sync()
    shapes = shapes_of(ctx, kinds)
    say(f"bench:   {form} tile shapes " + ", ".join(f"{k}: {v['rows_per_segment']} rows, depth {v['depth']}"
                                                    + ("" if v["tuned"] else " (default)") for k, v in shapes.items()))
    if warmup > 0:
        ctx.step(warmup)
    primed = prime_graph(ctx, steps)
    gpu_s, wall_s = timer.run(lambda: ctx.step(steps))
    ctx.step(1, record=True)
    total = tune_steps + warmup + primed + steps + 1
    dp = [ctx.diag_drop(0.0, total - 1, case=i).delta_p for i in range(len(cases))]
    n_sites = NX * NY * len(cases)
    mode = ctx.step_mode()   # the launch form the library chose for the 14 cases
    blu = {}
    for i, k in enumerate(kinds):
        blu["P1/P3" if k == "P" else k] = ctx.stored_bytes(case=i)
    # the dominant kernel alone: the P1/P3 family launch (8 cases), at the tile shape the library tunes
    # for it running alone (leaving a member out and back in re-tunes the family), in the launch form
    # of the 14-case step
    for i, k in enumerate(kinds):
        if k != "P":
            ctx.set_active(False, case=i)
    p0 = kinds.index("P")
    ctx.set_active(False, case=p0)
    ctx.set_active(True, case=p0)
    nP = sum(1 for k in kinds if k == "P")
    bytes_launch = sum(ctx.stored_bytes(case=i) for i, k in enumerate(kinds) if k == "P") * NX * NY
    ksteps = max(steps, 200)
    with launch_form(mode):
        ctx.step(tune_steps)
        ctx.sync()
        kshape = ctx.tile_shape(case=p0)
        kern_s, _ = timer.run(lambda: ctx.step(ksteps))
    ctx.close()
    avg = kern_s / ksteps
    return {"MLUPS": n_sites * steps / wall_s / 1e6, "ms_per_step": wall_s / steps * 1e3,
            "gpu_ms_per_step": gpu_s / steps * 1e3, "tile_shapes": shapes,
            "launch_form": mode,
            "dominant_kernel": KERNEL[form] if mode <= 0 else "step_fused (one launch; here the P1/P3 family's cases)",
            "kernel_cases_per_launch": nP,
            "kernel_tile_shape": {"rows_per_segment": kshape[0], "depth": kshape[1], "tuned": kshape[2]},
            "kernel_launches_timed": ksteps, "kernel_avg_launch_us": avg * 1e6, "bytes_per_launch": bytes_launch,
            "achieved_GBs": bytes_launch / avg / 1e9, "frac": bytes_launch / avg / 1e9 / PEAK_HBM_GBS,
            "bytes_per_lattice_update": blu, "steps_before_delta_p": total, "_dp": dp}


FAMILY_OPS = {"P": "P3 0.5", "F3": "F3 1.569988", "F1": "F1", "F2": "F2"}


def bench_family(m, timer, stream, family, form, steps, tune_steps):
    """One kernel family alone: eight 512^2 sharp drops R = 100 of that family's operator (tau
    alternating 0.6 / 1.0), so each launch is that family's kernel over 8 cases; tuned tile shape (or
    the one ITOLBM_LY / ITOLBM_DEPTH fix)."""
    op = FAMILY_OPS[family]
    kind = "P" if family == "P" else "F"
    descs = [m.case_desc(NX, NY, op, tau, amplitude(kind, tau), 0.8) for tau in TAUS * 4]
    ctx = m.Context(descs, precision=PREC_ID[form])
    ctx.set_stream(stream.cuda_stream)
    for i in range(len(descs)):
        ctx.init_drop(R, width=-1.0, case=i)
    with launch_form(0):   # the family's own kernel
        ctx.step(tune_steps)
        ctx.sync()
        rows, depth, tuned = ctx.tile_shape()
    gpu_s, _ = timer.run(lambda: ctx.step(steps))
    blu = ctx.stored_bytes()
    ctx.close()
    avg = gpu_s / steps
    nbytes = blu * NX * NY * len(descs)
    return {"family": family, "operator": op, "form": form, "cases": len(descs), "steps": steps,
            "tile_shape": {"rows_per_segment": rows, "depth": depth, "tuned": tuned},
            "bytes_per_lattice_update": blu,
            "kernel_avg_launch_us": avg * 1e6, "bytes_per_launch": nbytes, "achieved_GBs": nbytes / avg / 1e9,
            "frac": nbytes / avg / 1e9 / PEAK_HBM_GBS}


def dp_agreement(dp, ref):
    err = [abs(a - b) / abs(b) for a, b in zip(dp, ref) if b != 0 and a == a and b == b]
    err.sort()
    if not err:
        return None
    return {"cases": len(err), "median_rel": err[len(err) // 2], "max_rel": err[-1],
            "within_1e-4": sum(1 for e in err if e <= 1e-4)}


def bench_large(m, timer, stream, form, steps, nx=8192, ny=8192, R=1000.0):
    """One large case: the 8192^2 translating drop, or (nx = 8192, ny = 1024) one GPU's y-slab of it
    when the box is split over 8 GPUs (a periodic 8192 x 1024 box, R = 300: the kernel at the slab's
    shape; the slab's two halo rows and their exchange are not part of it)."""
    ctx = m.Context(m.case_desc(nx, ny, "P3 0.5", 1.0, amplitude("P", 1.0), 0.8), precision=PREC_ID[form])
    ctx.set_stream(stream.cuda_stream)
    ctx.init_drop(R, width=-1.0, v0=0.1)
    ctx.step(LARGE_TUNE)   # untimed: the library tunes the tile shape on these steps
    ctx.sync()
    rows, depth, tuned = ctx.tile_shape()
    prime_graph(ctx, steps)
    gpu_s, wall_s = timer.run(lambda: ctx.step(steps))
    mode = ctx.step_mode()
    blu = ctx.stored_bytes()
    ctx.close()
    avg = gpu_s / steps
    n = nx * ny
    return {"workload": f"{nx}x{ny}x1 translating sharp drop R={R:g} v0=0.1, P3 chi=0.5, tau=1, {form} storage",
            "steps": steps, "tile_shape": {"rows_per_segment": rows, "depth": depth, "tuned": tuned},
            "launch_form": mode,
            "MLUPS": n * steps / wall_s / 1e6, "kernel_avg_launch_us": avg * 1e6,
            "bytes_per_lattice_update": blu,
            "achieved_GBs": blu * n / avg / 1e9, "frac": blu * n / avg / 1e9 / PEAK_HBM_GBS}


def stream_ceiling(m, timer, stream, reps=10):
    """The library's own streaming kernel (lbm_stream_copy: 16-byte loads, nontemporal 16-byte stores)
    copying 1 GiB on the same stream: read + write bandwidth a plain streaming kernel reaches here."""
    import torch
    n = 1 << 30
    x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        m.stream_copy(y.data_ptr(), x.data_ptr(), n, stream.cuda_stream)

    def run():
        for _ in range(reps):
            m.stream_copy(y.data_ptr(), x.data_ptr(), n, stream.cuda_stream)
    gpu_s, _ = timer.run(run)
    ok = bool(torch.equal(x, y))
    del x, y
    return {"bytes_per_copy": 2 * n, "GBs": 2 * n * reps / gpu_s / 1e9, "copy_correct": ok}


def device_warmup(m, stream, seconds=2.0):
    """Untimed streaming copies for `seconds` before anything is timed.  The GPU raises its clocks
    over the first ~0.1-0.2 s of sustained load: eight 512^2 P-family cases timed after 20 steps of
    a fresh process ran at 113 us per launch, the same launches after 2 s of load at 97.5 us."""
    import time
    import torch
    n = 1 << 30
    x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    t0 = time.time()
    while time.time() - t0 < seconds:
        for _ in range(10):
            m.stream_copy(y.data_ptr(), x.data_ptr(), n, stream.cuda_stream)
        stream.synchronize()
    del x, y


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
    """168 cases: plain FP64 one launch per case and step vs the batched form, `steps` steps each in
    all; the first SWEEP_WARMUP (tile tuning, first launches) are not timed."""
    import numpy as np
    cases = sweep_cases(m)
    descs = [c for c, _, _ in cases]
    sites = sum(d.nx * d.ny for d in descs)
    timed = steps - SWEEP_WARMUP
    res = {}
    runs = [("f64_per_case", 0), ("f64_batched", 0), (f"{form}_batched", PREC_ID[form])]
    if form != "f16z":
        runs.append(("f16z_batched", PREC_ID["f16z"]))   # FP16 storage, for its delta_p agreement
    for label, prec in runs:
        say(f"bench:   sweep {label}")
        ctx = m.Context(descs, precision=prec)
        ctx.set_stream(stream.cuda_stream)
        for i, (d, r, _) in enumerate(cases):
            ctx.init_drop(r, width=3.45, case=i)
        per_case = label == "f64_per_case"
        if per_case:
            for i in range(len(descs)):
                ctx.step(SWEEP_WARMUP, case=i)
        else:
            ctx.step(SWEEP_WARMUP)
        ctx.sync()
        if per_case:
            # one launch per case and step (cases are independent, so case by case is the same work)
            def run():
                for i in range(len(descs)):
                    ctx.step(timed - 1, case=i)
                    ctx.step(1, record=True, case=i)
        else:
            def run():
                ctx.step(timed - 1)
                ctx.step(1, record=True)
        gpu_s, wall_s = timer.run(run)
        diags = [ctx.diag_drop(0.0, steps - 1, case=i) for i in range(len(descs))]
        dp = np.array([g.delta_p for g in diags])
        ok = np.array([g.ok for g in diags])
        ctx.close()
        res[label] = {"wall_s": wall_s, "ms_per_step": wall_s / timed * 1e3, "dp": dp, "ok": ok}
    a, b = res["f64_per_case"], res[f"{form}_batched"]

    def agreement(b):
        good = (a["ok"] == 1) & (b["ok"] == 1) & np.isfinite(a["dp"]) & np.isfinite(b["dp"]) & (np.abs(a["dp"]) > 1e-9)
        return good, np.abs(b["dp"][good] - a["dp"][good]) / np.abs(a["dp"][good])

    good, err = agreement(b)
    labels = [lab for (_, _, lab), g in zip(cases, good) if g]
    per_hour = lambda ms: 168 * 3600.0 / (ms * 1e-3 * 20000)
    over = [[labels[j], float(err[j])] for j in np.argsort(err)[::-1] if err[j] > 1e-4]
    fb = res["f64_batched"]
    f64_batched = {"ms_per_step": fb["ms_per_step"], "speedup_vs_per_case": a["ms_per_step"] / fb["ms_per_step"],
                   "dp_equal_to_per_case": bool(np.array_equal(fb["dp"], a["dp"], equal_nan=True))}
    fp16 = None
    if "f16z_batched" in res:
        g16, e16 = agreement(res["f16z_batched"])
        fp16 = {"batched_ms_per_step": res["f16z_batched"]["ms_per_step"], "dp_compared_cases": int(g16.sum()),
                "dp_rel_err_median": float(np.median(e16)) if e16.size else None,
                "dp_within_1e-4": int((e16 <= 1e-4).sum())}
    return {"cases": len(descs), "sites": sites, "steps": steps, "untimed_steps": SWEEP_WARMUP, "form": form,
            "f16z": fp16, "f64_batched": f64_batched,
            "f64_per_case_ms_per_step": a["ms_per_step"], "batched_ms_per_step": b["ms_per_step"],
            "speedup": a["ms_per_step"] / b["ms_per_step"],
            "cases_per_hour_20k_steps": {"f64_per_case": per_hour(a["ms_per_step"]),
                                         f"{form}_batched": per_hour(b["ms_per_step"])},
            "unstable_f64": int((a["ok"] != 1).sum()), "unstable_batched": int((b["ok"] != 1).sum()),
            "dp_compared_cases": int(good.sum()),
            "dp_rel_err_median": float(np.median(err)) if err.size else None,
            "dp_rel_err_max": float(err.max()) if err.size else None,
            "dp_within_1e-4": int((err <= 1e-4).sum()),
            "dp_beyond_1e-4": over}


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(max_threads):
    """The reference program itself (oracle/_ref/Imm_OPS, compiled from the reference's sources by
    oracle/Makefile with -O3 -march=x86-64-v3 -fopenmp) on this host's cores: a 512^2 P3 chi = 0.5
    case directory, bounded number of steps; MLUPS as the program reports it in resumo.dat."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    model, nproc = host_cpu()
    exe = os.path.join(ROOT, "oracle", "_ref", "Imm_OPS")
    threads = max(1, min(max_threads, nproc))
    if os.path.exists(exe):
        import _cases as K
        steps = 3001   # about 13 s at 57-59 MLUPS on 16 threads
        with tempfile.TemporaryDirectory() as d:
            K.make_case(d, NX, "P3 0.5", 1.0, amplitude("P", 1.0), steps)
            txt = open(os.path.join(d, "data_in.txt")).read().replace("n_threads:  1", f"n_threads:  {threads}")
            open(os.path.join(d, "data_in.txt"), "w").write(txt)
            t0 = time.perf_counter()
            r = subprocess.run([exe], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            dt = time.perf_counter() - t0
            mlups = None
            if r.returncode == 0 and os.path.exists(os.path.join(d, "resumo.dat")):
                mlups = float(open(os.path.join(d, "resumo.dat")).read().splitlines()[1].split()[-1])
        if mlups:
            return {"value": mlups, "unit": "MLUPS", "cores": threads, "kind": "reference",
                    "cpu_model": model, "nproc": nproc,
                    "sample": f"the reference program Imm_OPS (OpenMP, {threads} threads, built from its sources with "
                              f"-O3 -march=x86-64-v3) on a 512x512 P3 chi=0.5 tau=1 case directory (its own drop: "
                              f"R = nx/4, logistic profile; per-step work as the bench drop), {steps} steps with its "
                              f"diagnostics every 500; MLUPS as the program writes it to resumo.dat "
                              f"(process wall {dt:.1f} s)"}
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
            "cpu_model": model, "nproc": nproc,
            "sample": f"oracle C restatement, 1 thread: 512x512 P3 chi=0.5, {steps} steps, wall {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50, help="untimed steps after the tuning steps")
    ap.add_argument("--tune-steps", type=int, default=500,
                    help="untimed steps first, on which the library tunes its tile shapes (needs ~400)")
    ap.add_argument("--prec", default="f64", choices=FORMS, help="headline storage form")
    ap.add_argument("--forms", default="f32,f32z,f16,f16z", help="forms measured beside the headline")
    ap.add_argument("--sweep-steps", type=int, default=20000, help="0 skips the 168-case sweep")
    ap.add_argument("--large-steps", type=int, default=20, help="0 skips the 8192^2 cases")
    ap.add_argument("--large-forms", default="f32z,f32,f64")
    ap.add_argument("--shard-steps", type=int, default=20,
                    help="timed steps of the 8192 x 1024 slab in the --large-forms (0 skips it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-only", action="store_true",
                    help="only the 1 GiB streaming copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--family", choices=list(FAMILY_OPS),
                    help="only one kernel family, timed alone over eight 512^2 cases in form --prec")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if a.sweep_steps and a.sweep_steps <= SWEEP_WARMUP + 1:
        ap.error(f"--sweep-steps must exceed {SWEEP_WARMUP + 1} (or be 0)")

    import torch
    import _pkg
    m = _pkg.load()
    stream = torch.cuda.Stream()
    timer = Timer(torch, stream)
    if a.copy_only:
        print(json.dumps({"copy": stream_ceiling(m, timer, stream, reps=3)}), flush=True)
        return
    device_warmup(m, stream)
    if a.family:
        print(json.dumps(bench_family(m, timer, stream, a.family, a.prec, a.steps, a.tune_steps)), flush=True)
        return

    forms = {}
    for form in [a.prec] + [f for f in a.forms.split(",") if f and f != a.prec]:
        say(f"bench: form {form}, 14 cases x 512^2, {a.tune_steps} tuning + {a.warmup} warm-up + {a.steps} timed steps")
        forms[form] = bench_form(m, timer, stream, form, a.steps, a.warmup, a.tune_steps)
    ref_dp = forms["f64"]["_dp"] if "f64" in forms else None
    for form, r in forms.items():
        dp = r.pop("_dp")
        if ref_dp is not None and form != "f64":
            r["delta_p_vs_f64"] = dp_agreement(dp, ref_dp)
    head = forms[a.prec]

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{a.prec}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    say("bench: 1 GiB streaming copy")
    copy = stream_ceiling(m, timer, stream)
    out = {
        "metric": "MLUPS + % HBM roofline, 1 GPU; sweep cases/hour vs plain FP64 at same Δp error",
        "value": head["MLUPS"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "tune_steps": a.tune_steps,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[a.prec],
        "data": "synthetic",
        "config": {"workload": "512x512x1 static sharp drop R=100; tau {0.6, 1.0} x 7 operators "
                               "(P1, P3 chi -0.4/0.5/2, F1, F2, F3 xi 1.569988), sigma 0.1, beta 0.8; 14 cases "
                               f"batched; {FORM_TEXT[a.prec]}",
                   "cases": 14, "sites_per_case": NX * NY, "storage": a.prec,
                   "bytes_per_lattice_update": head["bytes_per_lattice_update"]},
        "roofline": {"bound": "hbm", "kernel": head["dominant_kernel"], "achieved": head["achieved_GBs"],
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": head["frac"], "traffic": traffic,
                     "algorithmic_bytes_per_launch": head["bytes_per_launch"],
                     "avg_launch_us": head["kernel_avg_launch_us"],
                     "stream_GBs": copy["GBs"], "frac_of_stream": head["achieved_GBs"] / copy["GBs"]},
        "forms": forms,
    }
    if a.large_steps > 0:
        out["large"] = {}
        for form in [f for f in a.large_forms.split(",") if f]:
            say(f"bench: 8192^2 {form}, {LARGE_TUNE} tuning + {a.large_steps} timed steps")
            out["large"][form] = bench_large(m, timer, stream, form, a.large_steps)
    if a.shard_steps > 0:
        out["shard_8gpu"] = {}
        for form in [f for f in a.large_forms.split(",") if f]:
            say(f"bench: 8192 x 1024 (one y-slab of 8192^2 over 8 GPUs) {form}, {LARGE_TUNE} tuning + "
                f"{a.shard_steps} timed steps")
            out["shard_8gpu"][form] = bench_large(m, timer, stream, form, a.shard_steps, ny=1024, R=300.0)
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
