"""Benchmark: dose fields/s for the 128^3-voxel, 67,240-spot forward dose on one GPU.

Workload (one step = one forward dose of the whole field, inputs resident in HBM):
  water grid 128 x 128 x 128 at 2 mm; 41 x 41 spot lattice at 5 mm; 40 energy
  layers linspace(70, 180) MeV; spot weights lognormal(0, 0.5) from seed 0;
  per-layer IDD and sigma from the on-device FP64 LUT builder.

Prints ONE JSON line. ``value`` = fields/s of the selected arithmetic form
(``--precision``). ``roofline`` is for the dominant kernel (the fused forward),
timed per launch with HIP events on the launch stream. ``cpu_baseline`` times
the host restatement of the reference's PyTorch batched-matmul engine
(oracle/torch_cpu.py, FP64, all host threads) on full forwards of the same field.
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

PEAK = {"fp32": (157.3, "mfma"), "bf16": (2500.0, "mfma"), "fp16": (2500.0, "mfma")}
KERNEL = {"fp32": "fwd_fp32_kernel"}


def field_flops(nx, ny, nz, nsx, nsy, L):
    """Algorithmic FLOP of one forward: 2 L nz min(nsy nsx ny + nx nsx ny, nsy nsx nx + ny nsy nx)."""
    a = nsy * nsx * ny + nx * nsx * ny
    b = nsy * nsx * nx + ny * nsy * nx
    return 2.0 * L * nz * min(a, b)


def build_field(device):
    from pbdose import engines, model, physics
    grid = model.DoseGrid.centered(128, 0.2)
    energies = np.linspace(70.0, 180.0, 40)
    w = np.random.default_rng(seed=0).lognormal(0.0, 0.5, (40, 41, 41))
    idd, sig, _ = physics.build_luts(grid.nz, grid.oz, grid.dz, energies, device=device)
    lat = model.SpotLattice(41, 41, 40, 0.5, 0.5, energies=energies, weights=w)
    prob = model.PBProblem(grid, lat, idd.cpu().numpy(), sig.cpu().numpy())
    return prob, engines.DeviceProblem(prob, device, idd=idd, sigma=sig)


def time_forward(dp, precision, steps, warmup):
    out = torch.empty(dp.shape, dtype=torch.float32, device=dp.dev)
    for _ in range(warmup):
        dp.forward(precision, out=out)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record()
        dp.forward(precision, out=out)
        e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))
    return out, wall * 1e3 / steps, launch_ms


def parity(prob, D):
    from oracle import pb_oracle as O
    idd, sig = O.layer_luts(prob.grid.z, prob.lattice.energies)
    zs = np.arange(0, prob.grid.nz, 8)
    ref = O.separable_dose(prob.grid.x, prob.grid.y, prob.lattice.spot_x, prob.lattice.spot_y,
                           prob.lattice.weights, idd, sig, z_idx=zs)
    r = O.relative_error(D.double().cpu().numpy()[zs], ref)
    return r["max_rel_err_global"], r["rms_rel_global"], len(zs)


def cpu_baseline(prob, budget_s=20.0, max_reps=15):
    from oracle.torch_cpu import torch_separable_cpu
    nthreads = os.cpu_count() or 1
    torch.set_num_threads(nthreads)
    g, lat = prob.grid, prob.lattice
    args = (g.x, g.y, lat.spot_x, lat.spot_y, lat.weights, prob.idd, prob.sigma)
    for _ in range(2):
        torch_separable_cpu(*args)
    ts = []
    t_start = time.perf_counter()
    while len(ts) < max_reps and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        torch_separable_cpu(*args)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    return {"value": 1.0 / med, "unit": "fields/s", "cores": nthreads, "kind": "port",
            "sample": f"{len(ts)} full forwards of the same 128^3 / 67,240-spot field, FP64 "
                      f"torch bmm on the host (median {med * 1e3:.1f} ms, min "
                      f"{min(ts) * 1e3:.1f} ms; 2 warm-up runs)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build runs the "
              "whole field on a single MI355X", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    prob, dp = build_field(dev)
    g, lat = prob.grid, prob.lattice
    flops = field_flops(g.nx, g.ny, g.nz, lat.n_sx, lat.n_sy, lat.n_layers)

    D, ms_step, launch_ms = time_forward(dp, args.precision, args.steps, args.warmup)
    err_max, err_rms, n_sl = parity(prob, D)
    peak, bound = PEAK.get(args.precision.split("x")[0].replace("_peak", ""), PEAK["fp32"])
    achieved = flops / (launch_ms * 1e-3) / 1e12
    line = {
        "metric": "dose fields/s at 128^3 x 67,240 spots (forward)",
        "value": 1e3 / ms_step,
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32"}.get(args.precision, args.precision),
        "data": "synthetic (lognormal spot weights, seed 0; on-device FP64 LUT)",
        "config": {"workload": "water 128x128x128 @ 2 mm, 41x41 spots @ 5 mm x 40 layers "
                               "70-180 MeV (67,240 spots), forward dose",
                   "precision": args.precision},
        "max_rel_err_vs_fp64": err_max,
        "rms_rel_err_vs_fp64": err_rms,
        "parity_slices": n_sl,
        "roofline": {"bound": bound, "kernel": KERNEL.get(args.precision, "fwd_mixed_kernel"),
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "algorithmic_flop_per_launch": flops, "launch_ms": launch_ms},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(prob)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
