"""Benchmark: dose fields/s for the 128^3-voxel, 67,240-spot forward dose on one GPU,
plus forward + adjoint time and the mixed-precision vs FP32 comparison.

Workload (one step = one forward dose of the whole field, inputs resident in HBM):
  water grid 128 x 128 x 128 at 2 mm; 41 x 41 spot lattice at 5 mm; 40 energy
  layers linspace(70, 180) MeV; spot weights lognormal(0, 0.5) from seed 0;
  per-layer IDD and sigma from the on-device FP64 LUT builder.

Prints ONE JSON line. ``value`` = fields/s of the selected arithmetic form
(``--precision``, default the BF16 split-product form); the FP32 form, the
forward + adjoint time (adjoint of the weighted-quadratic objective: dose
residual kernel + adjoint + depth reduction; FP32 adjoint for the FP32 form,
BF16 or BF16x3 adjoint for the 16-bit forms) and the error of each form
against the FP64 oracle are reported beside it. ``roofline`` is for the forward
call of the selected form (weight split, fused forward, layer-chunk combine),
timed per call with HIP events on the launch stream. Step and call times
(``ms_per_step``, ``adjoint_ms``) are wall time over back-to-back calls with a
synchronisation on both sides; the per-call events run in a pass of their own
(``launch_ms``, ``adjoint_launch_ms``), since recording them between calls adds
host time that the GPU would otherwise hide. ``cpu_baseline`` times the host restatement of the
reference's PyTorch batched-matmul engine (oracle/torch_cpu.py, FP64, all host
threads) on full forwards of the same field; its FP64 output is the reference
for ``accuracy_vs_fp64`` (errors and the 2 %/2 mm gamma pass rate of the
selected form and of the FP32 form), which is therefore reported only when the
CPU leg runs. ``robust_batch`` times the
robust-scenario workload (3 fields x 21 scenarios, 63 field-scenarios per graph
replay), forward only and forward + adjoint, with autotuned launch parameters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "gpu-proton-dose-engine_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

# HBM bytes of one forward call on this workload (the kernels bench times between its
# events: weight split, forward, layer-chunk combine), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch, summed over the call's kernels) scaled by the same
# counters on a 256 MiB copy (fetch x 2.000, write x 1.000):
# profiles/counters_forward_adjoint_128cube_67240spots.md (the shapes bench.py's autotuner picked
# there: 64-row tiles, layer chunk 14 for bf16x3 and fp16x3, 10 for bf16; 32 rows, chunk 28 for FP32)
TRAFFIC_KIB = {"fp32": (5026.4, 14848.0), "bf16x3": (9624.1, 22927.6), "bf16": (11477.6, 28559.6),
               "fp16x3": (9639.1, 22928.9)}
CALL_KERNELS = {"fp32": "fwd_fp32_kernel + combine_kernel"}
KERNEL_NAME = {"fp32": "fwd_fp32_kernel"}
# MFMA FLOP the contraction kernel issues per launch on this workload (not the
# algorithmic count: windows skip K steps, the x3 forms run three products), from
# the rocprofv3 counter SQ_INSTS_VALU_MFMA_MOPS_* (units of 512 FLOP) in the
# counter profile under profiles/; None where not profiled
MFMA_FLOP_ISSUED = {"fp32": 11348832 * 512.0, "bf16x3": 22984032 * 512.0, "bf16": 7661344 * 512.0,
                    "fp16x3": 22984032 * 512.0}
FETCH_SCALE, WRITE_SCALE = 2.000, 1.000
HBM_PEAK_GBS = 8000.0

# published dense MFMA peaks of MI355X (TFLOP/s) for the arithmetic each form runs on
PEAK = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0, "fp64": 78.6}
DTYPE = {"fp32": "f32", "bf16": "bf16", "fp16": "f16"}


def base_type(precision: str) -> str:
    return precision.split("x")[0].split("_")[0]


def field_flops(nx, ny, nz, nsx, nsy, L):
    """Algorithmic FLOP of one forward: 2 L nz min(nsy nsx ny + nx nsx ny, nsy nsx nx + ny nsy nx)."""
    a = nsy * nsx * ny + nx * nsx * ny
    b = nsy * nsx * nx + ny * nsy * nx
    return 2.0 * L * nz * min(a, b)


def traffic_bytes(precision):
    """Profiled HBM bytes per forward launch (None for forms not profiled)."""
    if precision not in TRAFFIC_KIB:
        return None
    f, w = TRAFFIC_KIB[precision]
    return (f * FETCH_SCALE + w * WRITE_SCALE) * 1024.0


def build_field(device):
    from pbdose import engines, model, physics
    grid = model.DoseGrid.centered(128, 0.2)
    energies = np.linspace(70.0, 180.0, 40)
    w = np.random.default_rng(seed=0).lognormal(0.0, 0.5, (40, 41, 41))
    idd, sig, _ = physics.build_luts(grid.nz, grid.oz, grid.dz, energies, device=device)
    lat = model.SpotLattice(41, 41, 40, 0.5, 0.5, energies=energies, weights=w)
    prob = model.PBProblem(grid, lat, idd.cpu().numpy(), sig.cpu().numpy())
    return prob, engines.DeviceProblem(prob, device, idd=idd, sigma=sig)


def call_events(fn, steps):
    """Mean duration of fn() from events recorded around each call on the launch
    stream (a pass of its own: event records between calls add host time, so they
    stay out of the timed steps)."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for s, e in ev:
        s.record()
        fn()
        e.record()
    torch.cuda.synchronize()
    return float(np.mean([s.elapsed_time(e) for s, e in ev]))


def kernel_events(dp, precision, steps, out):
    """Mean duration of the fused contraction kernel alone: pbh_forward_profiled
    records a pair of events on the launch stream right around it, per call."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for s, e in ev:                       # create the events before handing them over
        s.record()
        e.record()
    for s, e in ev:
        dp.forward(precision, out=out, kernel_events=(s, e))
    torch.cuda.synchronize()
    return float(np.mean([s.elapsed_time(e) for s, e in ev]))


def time_forward(dp, precision, steps, warmup):
    """Returns (dose, wall ms per step over `steps` back-to-back steps, ms per call
    from per-call events, ms of the contraction kernel alone)."""
    out = torch.empty(dp.shape, dtype=torch.float32, device=dp.dev)
    for _ in range(warmup):
        dp.forward(precision, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dp.forward(precision, out=out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = call_events(lambda: dp.forward(precision, out=out), steps)
    kernel_ms = kernel_events(dp, precision, steps, out)
    return out, wall * 1e3 / steps, launch_ms, kernel_ms


def time_forward_fp64(dp, steps, warmup):
    """The FP64 form (torch_separable(dtype=torch.float64)): (dose, wall ms per step over
    back-to-back steps, ms per call from per-call events; the call is one kernel). Its
    weights are the FP32 weights every other form sees, held in FP64, so that the CPU
    leg's FP64 dose is its reference too."""
    out = torch.empty(dp.shape, dtype=torch.float64, device=dp.dev)
    w = dp.w.double().contiguous()
    for _ in range(warmup):
        dp.forward_fp64(out=out, w=w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dp.forward_fp64(out=out, w=w)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = call_events(lambda: dp.forward_fp64(out=out, w=w), steps)
    return out, wall * 1e3 / steps, launch_ms


def accuracy(ref64, D):
    """Errors of a device dose against the FP64 reference dose (checker only):
    global max / RMS (engines.py:364-378), mean and max pointwise relative error
    on voxels >= 10 % of max, and gamma 2 %/2 mm (10 % threshold) on the GPU."""
    from oracle import pb_oracle as O
    from pbdose import gamma
    Dh = D.double().cpu().numpy()
    r = O.relative_error(Dh, ref64)
    mean_pt, max_pt = O.mean_rel_error(Dh, ref64, 0.10)
    pr, n_eval, gmax = gamma.pass_rate(torch.as_tensor(ref64, device=D.device), D, (0.2, 0.2, 0.2), 2.0, 2.0, 10.0)
    return {"max_rel_err_global": r["max_rel_err_global"], "rms_rel_err_global": r["rms_rel_global"],
            "mean_rel_err_ge10pct": mean_pt, "max_rel_err_ge10pct": max_pt,
            "gamma_2pct_2mm_pass_rate": pr, "gamma_n_eval": n_eval, "gamma_max": gmax}


def time_torch_separable(prob, precision, reps):
    """The drop-in call as the reference's callers make it: host arrays copied to
    the device on every call, then the forward, then a synchronisation."""
    from pbdose import engines
    for _ in range(2):
        engines.torch_separable(prob, "cuda", precision=precision)
    ts = []
    for _ in range(reps):
        _, dt = engines.torch_separable(prob, "cuda", precision=precision, return_timing=True)
        ts.append(dt)
    return float(np.median(ts)) * 1e3


def adjoint_form(precision):
    """Arithmetic of the adjoint that goes with a forward form."""
    if precision == "fp32":
        return "fp32"
    t = "fp16" if precision.startswith("fp16") else "bf16"
    return t if precision in ("bf16", "fp16") else t + "x3"


def time_adjoint(dp, D, steps, warmup, precision="fp32"):
    """J and dJ/dw of the weighted-quadratic objective per call (objective residual,
    adjoint and depth reduction in one pbh_objective_adjoint call): wall ms per call
    over `steps` back-to-back calls, ms per call from per-call events, dJ/dw of the
    last call."""
    from pbdose import planning
    Dp, wv = planning.central_cube(dp.problem.grid, device=dp.dev)
    g = None

    def call():
        nonlocal g
        _, g = planning.objective_gradient(dp, D, Dp, wv, precision=precision)

    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return wall, call_events(call, steps), g


def time_planning_step(dp, steps, warmup, fwd="bf16x3", adj="bf16x3"):
    """One inverse-planning step per call (pbh_forward_objective_adjoint: forward dose,
    J and dJ/dw, with the objective formed in the forward's combine): wall ms per call
    over `steps` back-to-back calls, and dJ/dw of the last call."""
    from pbdose import planning
    Dp, wv = planning.central_cube(dp.problem.grid, device=dp.dev)
    out = torch.empty(dp.shape, dtype=torch.float32, device=dp.dev)
    g = None

    def call():
        nonlocal g
        _, _, g = planning.forward_objective_gradient(dp, Dp, wv, precision=adj, forward_precision=fwd, out=out)

    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, g


def time_robust(precision, steps, objective):
    """Robust-scenario batch (3 fields x 21 scenarios on 128^3), one graph replay
    per step; launch parameters from the measured autotuner."""
    from pbdose import robust
    b = robust.RobustBatch(robust.robust_fields(), robust.standard_scenarios(), precision=precision,
                           objective=objective, tuning="auto")
    b.capture()
    for _ in range(3):
        b.run()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        b.run()
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / steps
    flop = b.flops() * (2.0 if objective else 1.0)
    out = {"items": len(b.items), "ms": ms, "field_scenarios_per_s": len(b.items) * 1e3 / ms,
           "algorithmic_tflop_per_s": flop / (ms * 1e-3) / 1e12, "kernel_launches": b.launch_count,
           "tuning": b.tuning}
    b.close()
    return out


def host_cores():
    """(logical CPUs, physical cores, CPUs this process may run on)."""
    logical = os.cpu_count() or 1
    try:
        import psutil
        physical = psutil.cpu_count(logical=False) or logical
    except Exception:
        physical = logical
    try:
        allowed = len(os.sched_getaffinity(0))
    except Exception:
        allowed = logical
    return logical, physical, allowed


def cpu_leg(prob, budget_s=12.0, max_reps=15):
    """CPU baseline: the host restatement of the reference's batched-matmul engine
    (oracle/torch_cpu.py: broadcast exp factors, two torch.bmm, layer sum) on full
    forwards of the field, in FP64 (the reference's CPU default) and FP32, with the
    reference's protocol (2 warm-up runs, median of up to 15, time.perf_counter).
    Threads: the host share of one GPU on this machine, min(16, OMP_NUM_THREADS,
    CPUs), whatever os.cpu_count() reports; both counts are stated. Its FP64 output
    is the reference dose the GPU forms are checked against."""
    from oracle.torch_cpu import torch_separable_cpu
    logical, physical, allowed = host_cores()
    nthreads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS") or 16), allowed))
    torch.set_num_threads(nthreads)
    g, lat = prob.grid, prob.lattice
    w32 = lat.weights.astype(np.float32).astype(np.float64)     # the weights the GPU sees
    rows, ref = {}, None
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        args = (g.x, g.y, lat.spot_x, lat.spot_y, w32, prob.idd, prob.sigma)
        for _ in range(2):
            out = torch_separable_cpu(*args, dtype=dt)
        if dt == torch.float64:
            ref = out
        ts = []
        t_start = time.perf_counter()
        while len(ts) < max_reps and (time.perf_counter() - t_start) < budget_s:
            t0 = time.perf_counter()
            torch_separable_cpu(*args, dtype=dt)
            ts.append(time.perf_counter() - t0)
        rows[name] = {"median_ms": float(np.median(ts)) * 1e3, "min_ms": min(ts) * 1e3, "reps": len(ts)}
    med = rows["fp64"]["median_ms"] * 1e-3
    base = {"value": 1.0 / med, "unit": "fields/s", "cores": nthreads, "kind": "port",
            "sample": f"{rows['fp64']['reps']} full forwards of the same 128^3 / 67,240-spot field, FP64 "
                      f"torch bmm on the host (median {rows['fp64']['median_ms']:.1f} ms, min "
                      f"{rows['fp64']['min_ms']:.1f} ms; 2 warm-up runs) on {nthreads} threads",
            "host_logical_cpus": logical, "host_physical_cores": physical, "cpus_allowed": allowed,
            "threads_note": "threads = the host share of one GPU on this machine (16), not all its CPUs",
            "fp32": {"value": 1e3 / rows["fp32"]["median_ms"], "unit": "fields/s", **rows["fp32"]},
            "fp64": rows["fp64"]}
    return base, ref.numpy()


def profile_kernels(forms, reps=5):
    """Short run for rocprofv3: forwards of the given forms (at bench.py's autotuned
    tile shapes), objective + adjoint, and three 256 MiB device copies (known bytes,
    to calibrate HBM counters)."""
    from pbdose import planning
    dev = torch.device("cuda", 0)
    prob, dp = build_field(dev)
    out = torch.empty(dp.shape, dtype=torch.float32, device=dev)
    for f in forms:
        if f == "fp64":
            out64 = torch.empty(dp.shape, dtype=torch.float64, device=dev)
            w64 = dp.w.double().contiguous()
            for _ in range(reps):
                dp.forward_fp64(out=out64, w=w64)
            torch.cuda.synchronize()
            print(f"forward fp64: {reps} launches", flush=True)
            continue
        # the shape bench.py times (measured, or read from $PBDOSE_AUTOTUNE_CACHE when a
        # previous run of this mode filled it, so that a profiled run times nothing else)
        chunk, rows = dp.autotune(f)
        for _ in range(reps):
            dp.forward(f, out=out)
        torch.cuda.synchronize()
        print(f"forward {f}: {reps} launches ({rows}-row tiles, layer chunk {chunk})", flush=True)
    Dp, wv = planning.central_cube(prob.grid, device=dev)
    D = dp.forward("fp32")
    for aform in sorted({adjoint_form(f) for f in forms if f != "fp64"} | {"fp32"}):
        dp.autotune_split(aform)
        for _ in range(reps):
            planning.objective_gradient(dp, D, Dp, wv, precision=aform)
        torch.cuda.synchronize()
        print(f"objective + {aform} adjoint: {reps} calls", flush=True)
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy 256 MiB: 3 calls", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-robust", action="store_true", help="skip the robust-scenario batch leg")
    ap.add_argument("--profile-kernels", metavar="FORMS", default=None,
                    help="comma-separated forms: run a short kernel sequence for rocprofv3 and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the engine computes the "
              "whole field on a single MI355X and does not shard it over GPUs", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.profile_kernels:
        profile_kernels(args.profile_kernels.split(","))
        return
    prob, dp = build_field(dev)
    g, lat = prob.grid, prob.lattice
    flops = field_flops(g.nx, g.ny, g.nz, lat.n_sx, lat.n_sy, lat.n_layers)

    # every arithmetic form: back-to-back steps, per-call events, the contraction
    # kernel alone; the selected form is the headline
    forms = ["fp32", "bf16", "fp16", "bf16x3", "fp16x3", "bf16x3_peak", "fp16x3_peak"]
    if args.precision not in forms:
        raise SystemExit(f"unknown precision {args.precision}")
    per_form, doses = {}, {}
    from pbdose import _lib
    for f in forms:
        # measured tile height and layer chunk for this field (cached; outside the timing)
        dp.autotune(f)
        steps = args.steps if f in (args.precision, "fp32") else max(20, args.steps // 4)
        D, ms, call_ms, kern_ms = time_forward(dp, f, steps, args.warmup)
        doses[f] = D
        chunk, rows = dp._tuned[_lib.PREC[f]]
        per_form[f] = {"ms_per_step": ms, "fields_per_s": 1e3 / ms, "call_us": call_ms * 1e3,
                       "tile": f"{rows} x 128 voxels, layer chunk {chunk} (autotuned)",
                       "kernel_us": kern_ms * 1e3, "steps": steps,
                       "kernel_achieved_tflop_s": flops / (kern_ms * 1e-3) / 1e12,
                       "kernel_frac_of_peak": flops / (kern_ms * 1e-3) / 1e12 / PEAK[base_type(f)],
                       "accuracy_vs_fp64": None}
    # the FP64 form: FP64 weights, tables, arithmetic and dose (one kernel: FP64 MFMA on factor tables)
    steps64 = max(20, args.steps // 4)
    D64, ms64, call64 = time_forward_fp64(dp, steps64, args.warmup)
    doses["fp64"] = D64
    per_form["fp64"] = {"ms_per_step": ms64, "fields_per_s": 1e3 / ms64, "call_us": call64 * 1e3,
                        "tile": "16 x 128 voxels, all layers (FP64 MFMA 16x16x4 on LDS factor tables)",
                        "kernel_us": call64 * 1e3, "steps": steps64,
                        "kernel_achieved_tflop_s": flops / (call64 * 1e-3) / 1e12,
                        "kernel_frac_of_peak": flops / (call64 * 1e-3) / 1e12 / PEAK["fp64"],
                        "accuracy_vs_fp64": None}
    ms_step, launch_ms = per_form[args.precision]["ms_per_step"], per_form[args.precision]["call_us"] * 1e-3
    kernel_ms = per_form[args.precision]["kernel_us"] * 1e-3
    ms32 = per_form["fp32"]["ms_per_step"]
    D, D32 = doses[args.precision], doses["fp32"]
    aform = adjoint_form(args.precision)
    # measured layer splits of the adjoint (cached; outside the timing)
    dp.autotune_split("fp32")
    if aform != "fp32":
        dp.autotune_split(aform)
    adj32_ms, adj32_ev, g32 = time_adjoint(dp, D32, max(5, args.steps // 4), 3)
    if aform != "fp32":
        adj_ms, adj_ev, gmix = time_adjoint(dp, D32, max(5, args.steps // 4), 3, aform)
        adj_err = float((gmix.double() - g32.double()).abs().max() / g32.double().abs().max())
    else:
        adj_ms, adj_ev, adj_err = adj32_ms, adj32_ev, 0.0
    # both split-product adjoints side by side (the FP16 one scales R per wave)
    adj_forms = {}
    for af in ("bf16x3", "fp16x3"):
        if af == aform:
            ams, aerr = adj_ms, adj_err
        else:
            dp.autotune_split(af)
            ams, _, ga = time_adjoint(dp, D32, max(5, args.steps // 4), 3, af)
            aerr = float((ga.double() - g32.double()).abs().max() / g32.double().abs().max())
        adj_forms[af] = {"ms": ams, "max_rel_diff_vs_fp32_adjoint": aerr}
    # forward + objective + adjoint as one planning-step call, for the headline form and FP32
    step_ms, _ = time_planning_step(dp, max(5, args.steps // 4), 3, args.precision, aform)
    step32_ms, _ = time_planning_step(dp, max(5, args.steps // 4), 3, "fp32", "fp32")
    pcie_ms = time_torch_separable(prob, args.precision, max(5, args.steps // 10))
    bt = base_type(args.precision)
    peak = PEAK[bt]
    achieved_k = flops / (kernel_ms * 1e-3) / 1e12
    achieved_c = flops / (launch_ms * 1e-3) / 1e12
    issued = MFMA_FLOP_ISSUED.get(args.precision)
    line = {
        "metric": "dose fields/s at 128\u00b3\u00d767,240 spots; fwd+adjoint ms; mixed vs FP32 @\u22641e-5 rel",
        "value": 1e3 / ms_step,
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[bt],
        "data": "synthetic (lognormal spot weights, seed 0; on-device FP64 LUT)",
        "config": {"workload": "water 128x128x128 @ 2 mm, 41x41 spots @ 5 mm x 40 layers "
                               "70-180 MeV (67,240 spots), forward dose",
                   "precision": args.precision},
        "accuracy_vs_fp64": None,
        "fp32": {"fields_per_s": 1e3 / ms32, "ms": ms32, "launch_ms": per_form["fp32"]["call_us"] * 1e-3,
                 "accuracy_vs_fp64": None},
        "speedup_vs_fp32": ms32 / ms_step,
        "forms": per_form,
        "adjoint_form": aform,
        "adjoint_ms": adj_ms,
        "adjoint_ms_fp32": adj32_ms,
        "adjoint_launch_ms": adj_ev,
        "adjoint_launch_ms_fp32": adj32_ev,
        "adjoint_max_rel_diff_vs_fp32_adjoint": adj_err,
        "adjoint_forms": adj_forms,
        "fwd_adj_ms": ms_step + adj_ms,
        "fwd_adj_ms_fp32": ms32 + adj32_ms,
        "fwd_adj_speedup_vs_fp32": (ms32 + adj32_ms) / (ms_step + adj_ms),
        "planning_step": {"call": "pbh_forward_objective_adjoint (forward, J and dJ/dw in one call; the "
                                  "objective formed in the forward's layer-chunk combine)",
                          "forms": f"{args.precision} forward, {aform} adjoint",
                          "ms": step_ms, "ms_fp32": step32_ms, "speedup_vs_fp32": step32_ms / step_ms},
        "torch_separable_ms_incl_host_copies": pcie_ms,
        "roofline": {"bound": "mfma",
                     "kernel": KERNEL_NAME.get(args.precision, "fwd_mixed_kernel"),
                     "achieved": achieved_k, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved_k / peak, "traffic": traffic_bytes(args.precision),
                     "algorithmic_flop_per_launch": flops, "kernel_ms": kernel_ms,
                     "mfma_flop_issued_per_launch": issued,
                     "mfma_issued_tflop_s": issued / (kernel_ms * 1e-3) / 1e12 if issued else None,
                     "note": "achieved = algorithmic FLOP / the contraction kernel's mean duration "
                             "(events right around it on its stream); mfma_flop_issued = MFMA "
                             "instructions the kernel issues (windows skip some, split forms run "
                             "three products) from the rocprofv3 counter profile; traffic = the "
                             "forward call's HBM bytes from the same profile"},
        "roofline_call": {"kernels": CALL_KERNELS.get(args.precision,
                                                      "wsplit_kernel + fwd_mixed_kernel + combine_kernel"),
                          "achieved": achieved_c, "peak": peak, "unit": "TFLOP/s", "frac": achieved_c / peak,
                          "launch_ms": launch_ms},
    }
    # HBM rate of the forward call (profiled bytes per call / event-timed call) against the
    # 8 TB/s peak, and MFMA utilisation (MFMA work issued / the contraction kernel's time
    # against the dense peak of its type)
    tb = traffic_bytes(args.precision)
    line["hbm"] = {"bytes_per_call": tb, "gb_s": tb / (launch_ms * 1e-3) / 1e9 if tb else None,
                   "peak_gb_s": HBM_PEAK_GBS,
                   "frac": tb / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if tb else None,
                   # what a call must move: dose (FP32), spot weights (FP32), the two (L, nz) FP64 tables
                   "algorithmic_bytes_per_call": g.nx * g.ny * g.nz * 4 + lat.n_layers * lat.n_sy * lat.n_sx * 4 +
                                                 2 * lat.n_layers * g.nz * 8}
    line["mfma_utilisation"] = issued / (kernel_ms * 1e-3) / 1e12 / peak if issued else None
    if not args.no_robust:
        rs = max(5, args.steps // 10)
        line["robust_batch"] = {
            "workload": "fields 31x31x25, 41x41x40, 27x35x45 (nsx x nsy x layers) x 21 scenarios "
                        "(nominal, +-3 mm x/y/z, range x {1, 0.965, 1.035}) on water 128^3 @ 2 mm; "
                        "one HIP graph replay per step",
            "forward": time_robust(args.precision, rs, objective=False),
            "forward_adjoint": time_robust(args.precision, rs, objective=True)}
    if not args.no_cpu:
        line["cpu_baseline"], ref64 = cpu_leg(prob)
        for f in forms + ["fp64"]:
            per_form[f]["accuracy_vs_fp64"] = accuracy(ref64, doses[f])
        line["accuracy_vs_fp64"] = per_form[args.precision]["accuracy_vs_fp64"]
        line["fp32"]["accuracy_vs_fp64"] = per_form["fp32"]["accuracy_vs_fp64"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
