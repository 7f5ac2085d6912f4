"""Benchmark of the batched TDSE propagation on one MI355X.

Workload (the headline configuration): 4096 TDSE instances (64 r x 64 z of a focused
Gaussian beam, w0 = 100 um, 1.5e14 W/cm^2, 800 nm, 10-cycle sin^2 pulse), Ar soft-Coulomb
atom, x in [-800, 800] a.u. with dx = 0.1 (Nx = 16001), dt = 0.25 a.u., complex128.
Each instance runs N_E + 2 = 4416 Crank-Nicolson steps.  One bench "step" is one
propagation of the whole batch (one persistent kernel launch).

value = total grid-point-steps per second (sum over instances of Nx * steps / wall time),
inputs resident in HBM.  The JSON line also carries the HBM roofline of the propagation
kernel (32 B of psi traffic per grid-point-step, HIP events on the launch stream) and a
CPU baseline: the FP64 oracle (one instance per thread, OpenMP) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TDSE grid-pt-steps/s batched vs per-instance @≤1e-8 err; adjoint/fwd time ratio"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PT_STEP = 32.0  # psi read + write, complex128


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(target_s=12.0):
    """FP64 oracle on the host cores: a bounded sample of the same workload (Gaussian-beam
    members, full grid, truncated step count), one instance per thread."""
    from oracle import ctdse as O
    from mma_amd import workloads as W
    threads = max(1, min(os.cpu_count() or 1, 16))
    g = W.GRID
    x, psi0, e, it, st = O.ground_state(g["num_r"], g["dx"], g["a"], g["e_guess"], g["cv"])
    t, fields = W.gaussian_beam_batch(n_r=threads, n_z=1)
    nx = g["num_r"] + 1
    # calibrate on one thread
    cal_steps = 60
    t0 = time.perf_counter()
    O.propagate_batch(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, fields[:1], cal_steps, n_threads=1)
    rate1 = nx * cal_steps / (time.perf_counter() - t0)
    steps = int(min(fields.shape[1] - 1, max(50, target_s * rate1 / nx)))
    t0 = time.perf_counter()
    O.propagate_batch(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, fields, steps, n_threads=threads)
    dt = time.perf_counter() - t0
    return {"value": threads * nx * steps / dt, "unit": "grid-pt-steps/s", "cores": threads, "kind": "port",
            "sample": "%d Gaussian-beam instances x %d steps x Nx=%d, FP64 oracle (C, -O2 -ffp-contract=off), "
                      "OpenMP one instance per thread, %.1f s" % (threads, steps, nx, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-r", type=int, default=64)
    ap.add_argument("--n-z", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the TDSE instances are independent and would be sharded "
              "across GPUs without communication; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from mma_amd import tdse, workloads as W

    g = W.GRID
    log("setup: ground state + factor tables")
    geo = tdse.Geometry(g["num_r"], g["dx"], 0.25, g["a"], g["e_guess"], g["cv"], g["x_int"])
    t, fields = W.gaussian_beam_batch(n_r=args.n_r, n_z=args.n_z)
    n_inst, n_e = fields.shape
    n_steps = n_e + 2
    nx = g["num_r"] + 1
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    out = torch.empty((n_inst, 5, n_steps), dtype=torch.float64, device="cuda")
    pt_steps = float(n_inst) * nx * n_steps
    for w in range(args.warmup):
        log("warmup %d" % w)
        tdse.propagate_batch(geo, dev, out=out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        tdse.propagate_batch(geo, dev, out=out)
        ev[k][1].record(stream)
        log("step %d done" % k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = wall / args.steps * 1e3
    value = pt_steps / (wall / args.steps)
    achieved = BYTES_PER_PT_STEP * pt_steps / (kern_ms * 1e-3) / 1e9
    src = out[:, 0, :].cpu().numpy()
    finite = bool(np.all(np.isfinite(src)))

    single = None
    if not args.no_single:
        # per-instance form: the same propagation for one instance alone
        one = dev[:1].contiguous()
        o1 = torch.empty((1, 5, n_steps), dtype=torch.float64, device="cuda")
        tdse.propagate_batch(geo, one, out=o1)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        tdse.propagate_batch(geo, one, out=o1)
        torch.cuda.synchronize()
        single = nx * n_steps / (time.perf_counter() - s0)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline()

    line = {
        "metric": METRIC, "value": value, "unit": "grid-pt-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "tdse-batch-gaussian-beam", "instances": n_inst, "grid_r_x_z": [args.n_r, args.n_z],
                   "Nx": nx, "dx": g["dx"], "dt": 0.25, "time_steps": n_steps, "atom": "Ar (a=1.1893)",
                   "field": "800nm 10-cycle sin2, 1.5e14 W/cm2 peak, w0=100um", "precision": "complex128"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
        "per_instance_pt_steps_per_s": single,
        "batched_over_per_instance": (value / single) if single else None,
        "adjoint_over_forward": None,
        "kernel_ms": kern_ms,
        "outputs_finite": finite,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
