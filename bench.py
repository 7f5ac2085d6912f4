"""Benchmark of the batched TDSE propagation on one MI355X.

Workload (the headline configuration): 4096 TDSE instances (64 r x 64 z of a focused
Gaussian beam, w0 = 100 um, 1.5e14 W/cm^2, 800 nm, 10-cycle sin^2 pulse), Ar soft-Coulomb
atom, x in [-800, 800] a.u. with dx = 0.1 (Nx = 16001), dt = 0.25 a.u., complex128.
Each instance runs N_E + 2 = 4416 Crank-Nicolson steps.  One bench "step" is one
propagation of the whole batch (one persistent kernel launch).

value = total grid-point-steps per second (sum over instances of Nx * steps / wall time),
inputs resident in HBM.  The JSON line also carries the HBM roofline of the propagation
kernel (32 B of psi traffic per grid-point-step, HIP events on the launch stream), the
per-instance rate (one instance alone), the gradient workload's adjoint/forward time ratio
(1024 instances, Nx = 8001, dt = 0.1, 40000 field samples) and a CPU baseline: the
reference's own CTDSE sources compiled in oracle/_ref (else the FP64 oracle) on a bounded
sample of whole members, one per host thread.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-single]
                       [--no-adjoint] [--only-adjoint] [--copy-probe MIB]
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
# HBM bytes per launch of tdse_propagate_kernel on the default 64 x 64 workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE counters scaled by their readings for a plain 1 GiB device copy
# (x2.0 and x1.0); commands and readings in profiles/tdse_propagate_gaussian_beam_4096.md
MEASURED_TRAFFIC_BYTES = 5.147e12 + 4.656e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(target_s=8.0):
    """The CPU path on the host cores, on a bounded sample of the same workload: whole
    Gaussian-beam members (all steps, full grid), one member per thread.  When the reference's
    own CTDSE sources were compiled into oracle/_ref (gcc -O2), that build is timed (kind
    "reference", one call1DTDSE-equivalent per member from a thread pool; the reference keeps no
    mutable globals); otherwise the FP64 C oracle with OpenMP (kind "port").  The member count
    is a multiple of the thread count chosen from a short one-thread calibration."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import ctdse as O
    from mma_amd import workloads as W
    threads = max(1, min(os.cpu_count() or 1, 16))
    g = W.GRID
    nx = g["num_r"] + 1
    t, fields = W.gaussian_beam_batch(n_r=64, n_z=64)
    n_out = fields.shape[1] + 2  # exported samples; the reference runs n_out - 1 steps for them
    use_ref = os.path.exists(O.REF_PATH)
    if use_ref:
        x, psi0, e = O.ref_ground_state(g["num_r"], g["dx"], g["a"], g["e_guess"], g["cv"])
        t0 = time.perf_counter()
        O.ref_call(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, t[:200], fields[0, :200])
        per_inst = (time.perf_counter() - t0) * n_out / 202.0
    else:
        x, psi0, e, it, st = O.ground_state(g["num_r"], g["dx"], g["a"], g["e_guess"], g["cv"])
        padded = np.concatenate([fields, np.zeros((fields.shape[0], 3))], axis=1)
        t0 = time.perf_counter()
        O.propagate_batch(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, padded[:1], 200, n_threads=1)
        per_inst = (time.perf_counter() - t0) * n_out / 200.0
    reps = max(1, int(round(target_s / per_inst)))
    n_inst = threads * reps
    pick = np.linspace(0, fields.shape[0] - 1, n_inst).astype(int)
    t0 = time.perf_counter()
    if use_ref:
        with ThreadPoolExecutor(max_workers=threads) as pool:
            list(pool.map(lambda i: O.ref_call(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, t, fields[i]),
                          pick))
    else:
        O.propagate_batch(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, padded[pick], n_out - 1,
                          n_threads=threads)
    dt = time.perf_counter() - t0
    what = ("reference CTDSE sources (prop.c, tridiag.c, tools.c; gcc -O2 -ffp-contract=off), thread pool"
            if use_ref else "FP64 C oracle (gcc -O2 -ffp-contract=off), OpenMP")
    return {"value": n_inst * nx * n_out / dt, "unit": "grid-pt-steps/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": "%d of the 4096 Gaussian-beam members, %d samples each (Nx=%d), %s, one member per "
                      "thread, %.1f s wall" % (n_inst, n_out, nx, what, dt)}


def adjoint_measure(n_inst, n_t):
    """Gradient workload: n_inst instances, Nx = 8001, dt = 0.1, n_t field samples; loss = harmonic
    yield over harmonics 21-31.  Times one plain forward propagation (no checkpoints), the
    checkpointing forward of TDSEDipole and its backward (recompute + adjoint sweep)."""
    import torch
    from mma_amd import adjoint as AD, tdse, workloads as W
    g = W.ADJOINT_GRID
    geo = tdse.Geometry(g["num_r"], g["dx"], g["dt"], g["a"], g["e_guess"], g["cv"], g["x_int"])
    t, f = W.adjoint_batch(n_inst=n_inst, n_t=n_t)
    dev = torch.tensor(tdse.pad_fields(f), device="cuda")
    n = dev.shape[1] - 1
    win = AD.harmonic_window(n, g["dt"], n * g["dt"], W.omega_au(800e-9), 21, 31).cuda()
    # load the adjoint code path once on a tiny problem
    warm = dev[:2, :201].contiguous().requires_grad_(True)
    AD.harmonic_yield(AD.TDSEDipole.apply(warm, geo, 200), g["dt"], win[:101]).sum().backward()
    torch.cuda.synchronize()
    log("adjoint workload: %d instances, Nx %d, %d steps" % (n_inst, g["num_r"] + 1, n))
    out = torch.empty((n_inst, 5, n), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    tdse.propagate_batch(geo, dev, out=out)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    log("plain forward %.3f s" % t_fwd)
    del out
    runs = []
    for it in range(2):  # the first pass also pays the device allocations of checkpoints and blocks
        E = dev.clone().requires_grad_(True)
        t0 = time.perf_counter()
        s = AD.TDSEDipole.apply(E, geo, n)
        loss = AD.harmonic_yield(s, g["dt"], win).sum()
        torch.cuda.synchronize()
        t_ck = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        log("pass %d: checkpointing forward %.3f s, backward %.3f s" % (it, t_ck, t_bwd))
        runs.append((t_ck, t_bwd, bool(torch.isfinite(E.grad).all())))
        del E, s, loss
    t_ck, t_bwd, ok = runs[-1]
    return {"instances": n_inst, "Nx": g["num_r"] + 1, "dt": g["dt"], "steps": n, "forward_s": t_fwd,
            "checkpointing_forward_s": t_ck, "backward_s": t_bwd, "adjoint_over_forward": t_bwd / t_fwd,
            "forward_plus_backward_over_forward": (t_ck + t_bwd) / t_fwd, "gradient_finite": ok,
            "first_pass_s": list(runs[0][:2]), "checkpoint_every": int(math.ceil(math.sqrt(n)))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-r", type=int, default=64)
    ap.add_argument("--n-z", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-inst", type=int, default=1024)
    ap.add_argument("--adj-samples", type=int, default=40000)
    ap.add_argument("--only-adjoint", action="store_true", help="time only the gradient workload")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB three times and exit (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d MiB three times" % args.copy_probe)
        return
    if args.only_adjoint:
        import torch
        print(json.dumps({"adjoint": adjoint_measure(args.adj_inst, args.adj_samples)}), flush=True)
        return
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

    adj = None
    if not args.no_adjoint:
        del out, dev
        torch.cuda.empty_cache()
        adj = adjoint_measure(args.adj_inst, args.adj_samples)

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
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC_BYTES if (args.n_r, args.n_z) == (64, 64) else None},
        "cpu_baseline": cpu,
        "per_instance_pt_steps_per_s": single,
        "batched_over_per_instance": (value / single) if single else None,
        "adjoint_over_forward": adj["adjoint_over_forward"] if adj else None,
        "adjoint": adj,
        "kernel_ms": kern_ms,
        "outputs_finite": finite,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
