"""Benchmark of the batched TDSE propagation on one MI355X.

Workload (the headline configuration): 4096 TDSE instances (64 r x 64 z of a focused
Gaussian beam, w0 = 100 um, 1.5e14 W/cm^2, 800 nm, 10-cycle sin^2 pulse), Ar soft-Coulomb
atom, x in [-800, 800] a.u. with dx = 0.1 (Nx = 16001), dt = 0.25 a.u., complex128.
Each instance runs N_E + 2 = 4416 Crank-Nicolson steps.  One bench "step" is one
propagation of the whole batch (one persistent kernel launch).

value = total grid-point-steps per second (sum over instances of Nx * steps / wall time),
inputs resident in HBM.  The JSON line also carries the HBM roofline of the propagation
kernel (32 B of psi traffic per grid-point-step, HIP events on the launch stream), the
per-instance rate (one instance alone, as call1DTDSE runs it), the 8192-member uneven batch,
the one-launch-per-step form of the headline batch, the gradient workload's adjoint/forward
time ratio (1024 instances, Nx = 8001, dt = 0.1, 40000 field samples), the far-field
contraction (256 planes x 2048 frequencies x 512 radii -> 1024 screen radii) and a CPU
baseline: the reference's own CTDSE sources compiled in oracle/_ref (else the FP64 oracle)
on a bounded sample of whole members, one per host thread.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-single]
                       [--no-adjoint] [--no-uneven] [--no-stepwise] [--no-hankel]
                       [--only-single | --only-uneven | --only-adjoint | --only-hankel]
                       [--launch-config W,G,S] [--stream-probe GIB] [--copy-probe MIB]
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
# HBM bytes per launch on the default 64 x 64 workload, by steps per pass of the launch shape
# (1: tdse_propagate_kernel at 4 x 3, 2: tdse_propagate_pair_stash_kernel at 4 x 2), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE counters scaled by their readings for a plain 1 GiB device
# copy (x2.0 and x1.0); commands and readings in profiles/tdse_propagate_gaussian_beam_4096.md
MEASURED_TRAFFIC_BYTES = {1: 5.105e12 + 4.685e12, 2: 2.636e12 + 2.394e12}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class stdout_to_stderr:
    """Sends file descriptor 1 to stderr for the duration (the reference's C code prints its
    progress with printf; stdout carries only the result line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def host_info():
    """The host's CPUs as this process sees them: the machine's count (os.cpu_count, what nproc
    reports without an affinity mask), this process's affinity mask, the cgroup CPU quota, the
    OpenMP thread setting and the CPU model; `threads` is what the CPU baselines use: every CPU
    the affinity mask, the quota and OMP_NUM_THREADS leave to this process."""
    info = {"cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, p = f.read().split()[:2]
        if q != "max":
            quota = int(q) / int(p)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    omp = os.environ.get("OMP_NUM_THREADS")
    info["omp_num_threads"] = int(omp) if omp and omp.isdigit() else None
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["model"] = model
    threads = info["affinity"]
    if quota:
        threads = min(threads, max(1, int(quota)))
    if info["omp_num_threads"]:
        threads = min(threads, info["omp_num_threads"])
    info["threads"] = max(1, threads)
    return info


def cpu_baseline(target_s=8.0, gpu_rows=None):
    """The CPU path on the host cores, on a bounded sample of the same workload: whole
    Gaussian-beam members (all steps, full grid), one member per thread.  When the reference's
    own CTDSE sources were compiled into oracle/_ref (gcc -O2), that build is timed (kind
    "reference", one call1DTDSE-equivalent per member from a thread pool; the reference keeps no
    mutable globals); otherwise the FP64 C oracle with OpenMP (kind "port").  The member count
    is a multiple of the thread count chosen from a short one-thread calibration.  gpu_rows
    (n_inst, 4, n_out): the GPU's sourceterm, PopTot, PopInt, expval of the same batch; the CPU
    members are compared with them (max |a - b| / max |b| per row, and the L2 ratio)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import ctdse as O
    from mma_amd import workloads as W
    host = host_info()
    threads = host["threads"]
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
            res = list(pool.map(lambda i: O.ref_call(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, t,
                                                     fields[i]), pick))
        cpu_rows = [np.stack([r[k][:n_out] for k in ("sourceterm", "PopTot", "PopInt", "expval")]) for r in res]
    else:
        res = O.propagate_batch(g["num_r"], g["dx"], g["a"], g["x_int"], 0.25, psi0, x, padded[pick], n_out - 1,
                                n_threads=threads)
        cpu_rows = None
    dt = time.perf_counter() - t0
    parity = None
    if gpu_rows is not None and cpu_rows is not None:
        worst = np.zeros(4)
        l2 = np.zeros(4)
        for i, c in zip(pick, cpu_rows):
            gr = gpu_rows[i, :, :n_out]
            worst = np.maximum(worst, np.max(np.abs(gr - c), axis=1) / np.max(np.abs(c), axis=1))
            l2 = np.maximum(l2, np.linalg.norm(gr - c, axis=1) / np.linalg.norm(c, axis=1))
        parity = {"members": int(len(pick)), "against": "reference CTDSE sources on the CPU",
                  "max_abs_over_max": dict(zip(("sourceterm", "PopTot", "PopInt", "expval"), worst.tolist())),
                  "l2_rel": dict(zip(("sourceterm", "PopTot", "PopInt", "expval"), l2.tolist())),
                  "within_1e-8": bool(np.all(worst <= 1e-8))}
    what = ("reference CTDSE sources (prop.c, tridiag.c, tools.c; gcc -O2 -ffp-contract=off), thread pool"
            if use_ref else "FP64 C oracle (gcc -O2 -ffp-contract=off), OpenMP")
    return {"value": n_inst * nx * n_out / dt, "unit": "grid-pt-steps/s", "cores": threads, "host": host,
            "kind": "reference" if use_ref else "port",
            "sample": "%d of the 4096 Gaussian-beam members, %d samples each (Nx=%d), %s, one member per "
                      "thread, %.1f s wall" % (n_inst, n_out, nx, what, dt)}, parity


def adjoint_measure(n_inst, n_t):
    """Gradient workload: n_inst instances, Nx = 8001, dt = 0.1, n_t field samples; loss = harmonic
    yield over harmonics 21-31.  Times one plain forward propagation (no checkpoints), the
    checkpointing forward of TDSEDipole and its backward (states rebuilt next to the adjoint
    sweep), and once the backward that recomputes and stores each checkpoint block instead."""
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
    E = dev.clone().requires_grad_(True)
    loss = AD.harmonic_yield(AD.TDSEDipole.apply(E, geo, n, None, "stored"), g["dt"], win).sum()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t_stored = time.perf_counter() - t0
    log("backward with stored blocks %.3f s" % t_stored)
    del E, loss
    return {"instances": n_inst, "Nx": g["num_r"] + 1, "dt": g["dt"], "steps": n, "forward_s": t_fwd,
            "checkpointing_forward_s": t_ck, "backward_s": t_bwd, "adjoint_over_forward": t_bwd / t_fwd,
            "forward_plus_backward_over_forward": (t_ck + t_bwd) / t_fwd, "gradient_finite": ok,
            "backward_mode": "reconstruct", "stored_blocks_backward_s": t_stored,
            "first_pass_s": list(runs[0][:2]), "checkpoint_every": int(math.ceil(math.sqrt(n)))}


def hankel_cpu_baseline(zg, rg, og, rho, src, pre, distance, threads, target_s=6.0):
    """The FP64 oracle (numpy + scipy J0, the restatement of HankelTransform) on a slice of the
    far-field workload: a few whole planes, an omega subset, all rho and r; one omega per thread
    task.  Returns elements/s (element = one (plane, omega, rho, r) term)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import hankel as OH
    nz, no, nr = src.shape
    nrho = len(rho)
    pre_h = pre.cpu().numpy()

    def one(args):
        k, o = args
        plane = src[k, o:o + 1].cpu().numpy()
        OH.hankel_plane(og[o:o + 1], rg, plane, distance - zg[k], rho, True, np.outer(np.ones(nr), pre_h[k, o:o + 1]))
        return 1

    t0 = time.perf_counter()
    one((0, 0))
    per = time.perf_counter() - t0
    n_tasks = max(threads, int(round(target_s * threads / max(per, 1e-6))))
    tasks = [(k % nz, (k * 37) % no) for k in range(n_tasks)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        list(pool.map(one, tasks))
    dt = time.perf_counter() - t0
    elems = n_tasks * nrho * nr
    return {"value": elems / dt, "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": "%d (plane, omega) pairs x %d rho x %d r: FP64 oracle (numpy, scipy jv) one pair per task "
                      "on a %d-thread pool, %.1f s; C3 (%.3g elements) extrapolates to %.0f s"
                      % (n_tasks, nrho, nr, threads, dt, float(nz) * no * nrho * nr, float(nz) * no * nrho * nr * dt / elems)}


def single_measure(geo, field, n_steps, nx, reps=1):
    """Per-instance form: one instance alone, as the reference-compatible call1DTDSE runs it
    (ctdse_propagate_one: the segments of a step side by side, one launch per step); best of
    `reps` timed calls after one untimed call, allocation and the final wait included."""
    import torch
    from mma_amd import tdse
    one = field.contiguous()
    o1 = torch.empty((5, n_steps), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    tdse.propagate_one(geo, one, n_steps, o1, stream)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        s0 = time.perf_counter()
        tdse.propagate_one(geo, one, n_steps, o1, stream)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - s0)
    return {"pt_steps_per_s": nx * n_steps / best, "seconds": best, "Nx": nx, "steps": n_steps,
            "launches": n_steps, "form": "ctdse_propagate_one"}


def call1dtdse_measure():
    """The reference-compatible entry points end to end on the single-instance case (Nx 16001,
    10-cycle 800 nm pulse, 4416 exported samples), host arrays in and out as the reference's
    ctypes callers use them (TDSE_DLL): Initialise_grid_and_ground_state, a first call1DTDSE
    (geometry tables built, propagation, spectra, copies to host memory) and a second call on
    the same inputs (geometry reused).  Wall times in seconds."""
    from mma_amd import workloads as W
    from mma_amd.ctdse import TDSE_DLL
    from mma_amd.ctdse.structures import inputs_def, outputs_def
    g = W.GRID
    dll = TDSE_DLL()
    t, f = W.single_instance()
    inputs = inputs_def()
    inputs.init_default_inputs(Eguess=g["e_guess"], num_r=g["num_r"], dx=g["dx"], dt=0.25, trg_a=g["a"], CV=g["cv"],
                               x_int=g["x_int"])
    inputs.init_time_and_field(dll, t=t, E=f)
    t0 = time.perf_counter()
    dll.init_GS(inputs)
    t_gs = time.perf_counter() - t0
    calls = []
    for _ in range(3):
        outputs = outputs_def()
        t0 = time.perf_counter()
        dll.call1DTDSE(inputs, outputs)
        calls.append(time.perf_counter() - t0)
        n_out = outputs.Nt
        outputs.delete(dll)
    inputs.delete(dll)
    log("call1DTDSE: ground state %.1f ms, first call %.1f ms, then %.1f ms" % (t_gs * 1e3, calls[0] * 1e3,
                                                                              min(calls[1:]) * 1e3))
    return {"Nx": g["num_r"] + 1, "samples": n_out, "init_ground_state_s": t_gs, "first_call_s": calls[0],
            "repeat_call_s": min(calls[1:]),
            "repeat_call_pt_steps_per_s": (g["num_r"] + 1) * n_out / min(calls[1:])}


def hankel_measure(cpu=True):
    """Far-field workload: 256 source planes (5 mm Ar cell, z-modulated pressure, NIST f1 /
    Henke f2) x 2048 frequencies (harmonics 11-61 of 800 nm) x 512 radii (0-200 um) onto 1024
    screen radii (0-7 mm) at 1 m.  Times the device work of each contraction (pre-factor tables
    included) on complex128 sources, and the matrix-core ones on complex64 sources too, and states
    every result's error against the FP64 direct contraction."""
    import torch
    from mma_amd import hankel as H, workloads as W
    zall, rg, og, rho, src, distance = W.hankel_workload()
    pressure = W.hankel_pressure()
    zg = zall[:-1]
    nz, no, nr = src.shape
    elems = float(nz) * no * len(rho) * nr
    dz = zg[1:] - zg[:-1]
    w = np.zeros(nz)
    w[:-1] += 0.5 * dz
    w[1:] += 0.5 * dz
    Wt = np.stack([w, np.eye(nz)[0], np.eye(nz)[-1]])

    tabs = []

    def run(kind, sources=src):
        # the pre-factor tables are queued with the contraction; their table-error flag is read
        # after the pass (check_tables), as Hankel_long does before it hands results over
        tab = H._PreFactor(zg, rg, og, "Ar", pressure, "Henke", True, "NIST", True, 1.0, defer=True)
        tabs.append(tab)
        if kind == "fp64":
            return H.direct(sources, zg, og, rg, rho, distance, weights=Wt, pre=tab.pre)
        return H.gridded(sources, zg, og, rg, rho, distance, Wt, pre=tab.pre, kind=kind)

    def check_tables():
        for t in tabs:
            t.check()
        tabs.clear()

    out, res = {}, {}
    for kind in ("fp64", "fp64-grid", "fp32-mfma", "bf16x3-mfma", "bf16x6-mfma", "fp16x3-mfma"):
        run(kind)  # warm-up (code objects, allocations)
        torch.cuda.synchronize()
        reps = 1 if kind == "fp64" else 5
        check_tables()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = run(kind)
        torch.cuda.synchronize()
        check_tables()
        out[kind] = (time.perf_counter() - t0) / reps
        res[kind] = r
        log("hankel %s: %.2f ms" % (kind, out[kind] * 1e3))
    # the same passes on complex64 sources (half the bytes for the spread to read)
    src64 = src.to(torch.complex64)
    out64, res64 = {}, {}
    for kind in ("fp32-mfma", "bf16x3-mfma", "bf16x6-mfma", "fp16x3-mfma"):
        run(kind, src64)
        torch.cuda.synchronize()
        check_tables()
        t0 = time.perf_counter()
        for _ in range(5):
            r = run(kind, src64)
        torch.cuda.synchronize()
        check_tables()
        out64[kind] = (time.perf_counter() - t0) / 5
        res64[kind] = r
        log("hankel %s, complex64 sources: %.2f ms" % (kind, out64[kind] * 1e3))
    del src64
    ref = res["fp64"]
    errors = {}
    # each output against its own scale: the integrated field (a sum over planes weighted by the
    # plane spacing) is ~1/Nz of a plane's magnitude, so a max over all three would hide its error
    for kind in ("fp64-grid", "fp32-mfma", "bf16x3-mfma", "bf16x6-mfma", "fp16x3-mfma"):
        errors[kind] = {}
        for i, name in enumerate(("FF_integrated", "entry_plane_transform", "exit_plane_transform")):
            d = res[kind][i] - ref[i]
            errors[kind][name] = {"max_abs_over_max": float(d.abs().max() / ref[i].abs().max()),
                                  "frobenius": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref[i]))}
    errors64 = {}
    for kind in out64:
        d = res64[kind][0] - ref[0]
        errors64[kind] = {"FF_integrated": {
            "max_abs_over_max": float(d.abs().max() / ref[0].abs().max()),
            "frobenius": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref[0]))}}
    nq = int(H.grid_size_of(zg, og, rg, rho, distance)[0])
    gemm_flops = 2.0 * len(rho) * (3 * 2 * no) * nq
    info = {
        "workload": "far field: %d planes x %d omega x %d r -> %d rho, complex128 sources, Ar z-modulated cell"
                    % (nz, no, nr, len(rho)),
        "elements": elems,
        "seconds": out,
        "errors_vs_fp64_direct": errors,
        "speedup_vs_fp64_direct": {k: out["fp64"] / out[k] for k in out if k != "fp64"},
        "seconds_complex64_sources": out64,
        "errors_complex64_sources_vs_fp64_direct": errors64,
        "fp64_direct_j0_per_s": elems / out["fp64"],
        "fp64_direct_algorithmic_tflops": 4.0 * elems / out["fp64"] / 1e12,
        "grid_points_q": nq,
        "gemm_flops": gemm_flops,
    }
    if cpu:
        threads = host_info()["threads"]
        tab = H._PreFactor(zg, rg, og, "Ar", pressure, "Henke", True, "NIST", True, 1.0)
        info["cpu_baseline"] = hankel_cpu_baseline(zg, rg, og, rho, src, tab.pre, distance, threads)
    return info


def uneven_measure(steps=2, config=None):
    """Uneven workload: 8192 instances drawn from a seeded PCG64 generator, each with its own
    grid (Nx in [4001, 32001]), duration (6-20 cycles), peak intensity (5e13-3e14 W/cm^2,
    log-uniform) and wavelength (800-1800 nm), dt = 0.25.  All ground states and tables are
    built in grouped launches (timed as setup), the launch shape comes from the autotuner,
    and the batch runs as one persistent launch replayed from a HIP graph."""
    import torch
    from mma_amd import autotune, tdse, workloads as W
    U = W.UNEVEN
    members = W.uneven_batch(**U)
    fields, n_steps = W.stack_fields([m["field"] for m in members])
    num_r = [m["num_r"] for m in members]
    del members
    t0 = time.perf_counter()
    geoms = tdse.Geometry.create_many(num_r, 0.1, U["dt"], W.GRID["a"], W.GRID["e_guess"], W.UNEVEN_CV,
                                      W.GRID["x_int"])
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log("uneven: %d geometries built in %.2f s" % (len(geoms), t_setup))
    dev = torch.tensor(fields, device="cuda")
    del fields
    out = torch.zeros((len(geoms), 5, max(n_steps)), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    res = autotune.tune(geoms, dev, n_steps=n_steps) if config is None else autotune.TuneResult(
        tdse.LaunchConfig(*config), [], False, None)
    t_tune = time.perf_counter() - t0
    log("uneven: tuned %r in %.2f s (cached %s)" % (res.config, t_tune, res.cached))
    plan = tdse.BatchPlan(geoms, dev, n_steps=n_steps, out=out, config=res.config)
    default = tdse.default_config(geoms)
    stream = torch.cuda.current_stream()
    plan.replay(stream)  # captures the graph; warm-up
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        plan.replay(stream)
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kern = float(np.mean([a.elapsed_time(b) for a, b in ev])) * 1e-3
    ok = bool(torch.isfinite(out).all()) and float(out[:, 1, :].max()) <= 1.0 + 1e-12
    pt = float(plan.pt_steps)
    log("uneven: %.3f s per batch, %.3g pt-steps/s" % (wall, pt / wall))
    work = np.array([(r + 1) * s for r, s in zip(num_r, n_steps)], dtype=float)
    info = {"workload": "tdse-batch-uneven", "instances": len(geoms), "seed": U["seed"],
            "Nx_range": [int(min(num_r)) + 1, int(max(num_r)) + 1], "distinct_Nx": len(set(num_r)),
            "steps_range": [int(min(n_steps)), int(max(n_steps))], "pt_steps": pt,
            "max_over_min_instance_work": float(work.max() / work.min()),
            "value": pt / wall, "unit": "grid-pt-steps/s", "seconds": wall, "kernel_s": kern,
            "hbm_frac_psi_only": BYTES_PER_PT_STEP * pt / kern / 1e9 / HBM_PEAK_GBS,
            "setup_s": t_setup, "gs_cv": W.UNEVEN_CV,
            "gs_iterations_max": max(g.gs_iterations for g in geoms), "tune_s": t_tune, "launch_config": list(res.config.as_tuple()),
            "default_config": list(default.as_tuple()), "tune_table": [list(r) for r in res.table[:6]],
            "workgroups": plan.workgroups, "launch": "hipGraph replay of one persistent launch",
            "outputs_finite_and_norm_conserved": ok}
    plan.close()
    del out, dev, geoms
    return info


def stepwise_measure(geo, dev, n_steps, ref):
    """The plain form of the headline batch: one kernel launch per time step over all
    instances, states in HBM between launches (the same segment solve, in the default launch
    shape).  Returns the wall time and max |dipole difference| / max |dipole| against ref."""
    import torch
    from mma_amd import tdse
    out = torch.empty((dev.shape[0], 5, n_steps), dtype=torch.float64, device="cuda")
    tdse.propagate_stepwise(geo, dev[:, :3].contiguous(), out=out[:, :, :2].contiguous())  # load paths
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tdse.propagate_stepwise(geo, dev, out=out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    diff = float((out[:, 0] - ref[:, 0]).abs().max() / ref[:, 0].abs().max())
    del out
    return wall, diff


PROBE_KINDS = {"copy": 0, "read": 1, "write": 2}


def stream_rate(kind, gib, groups_per_cu, reps=5):
    """Rate (TB/s of bytes moved) of a bandwidth probe over gib GiB: 'copy' (read + write, both
    counted), 'read' (read only) or 'write' (write only), from mma_amd/lib/libmmaprobe.so (built
    with the package, loaded only here); HIP events on the launch stream.  The copy rate is the
    practical ceiling of the propagation kernel's read + write traffic."""
    import ctypes
    import torch
    from mma_amd import HERE
    L = ctypes.CDLL(os.path.join(HERE, "lib", "libmmaprobe.so"))
    L.probe_stream.restype = ctypes.c_int
    L.probe_stream.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int,
                               ctypes.c_void_p]
    nbytes = int(gib) << 30
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a) if kind != "read" else torch.zeros(2, dtype=torch.float64, device="cuda")
    k = PROBE_KINDS[kind]
    st = torch.cuda.current_stream()

    def run():
        code = L.probe_stream(k, a.data_ptr(), b.data_ptr(), nbytes, groups_per_cu, st.cuda_stream)
        assert code == 0, code

    run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        run()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    if kind == "copy":
        assert torch.equal(a, b)
    elif kind == "write":
        assert bool((b == 1.0).all())
    moved = (2.0 if kind == "copy" else 1.0) * nbytes
    rate = moved / (ms * 1e-3) / 1e12
    log("stream %s %d GiB, %d workgroups/CU: %.3f ms, %.3f TB/s" % (kind, gib, groups_per_cu, ms, rate))
    del a, b
    return rate


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
    ap.add_argument("--no-uneven", action="store_true", help="skip the 8192-instance uneven workload")
    ap.add_argument("--no-stepwise", action="store_true", help="skip the one-launch-per-step form")
    ap.add_argument("--only-uneven", action="store_true", help="time only the uneven workload")
    ap.add_argument("--only-single", action="store_true", help="time only one instance alone (call1DTDSE's form)")
    ap.add_argument("--launch-config", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    metavar="W,G,S", help="launch shape (waves/workgroup, workgroups/CU, steps/pass) instead "
                                          "of the autotuner's (profiling runs: no trial launches)")
    ap.add_argument("--adj-inst", type=int, default=1024)
    ap.add_argument("--adj-samples", type=int, default=40000)
    ap.add_argument("--only-adjoint", action="store_true", help="time only the gradient workload")
    ap.add_argument("--no-hankel", action="store_true")
    ap.add_argument("--only-hankel", action="store_true", help="time only the far-field workload")
    ap.add_argument("--stream-probe", type=int, default=0, metavar="GIB",
                    help="time the copy, read-only and write-only bandwidth probes over GIB GiB and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB three times and exit (HBM counter calibration)")
    args = ap.parse_args()
    if args.stream_probe:
        res = {kind: {g: stream_rate(kind, args.stream_probe, g) for g in (4, 8, 16, 32)} for kind in PROBE_KINDS}
        print(json.dumps({"stream_TBps": res, "GiB": args.stream_probe}), flush=True)
        return
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d MiB three times" % args.copy_probe)
        return
    if args.only_hankel:
        with stdout_to_stderr():
            hk = hankel_measure(cpu=not args.no_cpu)
        print(json.dumps({"hankel": hk}), flush=True)
        return
    if args.only_uneven:
        print(json.dumps({"uneven": uneven_measure(config=args.launch_config)}), flush=True)
        return
    if args.only_single:
        import torch
        from mma_amd import tdse, workloads as W
        g = W.GRID
        geo = tdse.Geometry(g["num_r"], g["dx"], 0.25, g["a"], g["e_guess"], g["cv"], g["x_int"])
        t, fields = W.gaussian_beam_batch(n_r=1, n_z=1)
        dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
        one = single_measure(geo, dev[0], fields.shape[1] + 2, g["num_r"] + 1, reps=3)
        with stdout_to_stderr():
            dropin = call1dtdse_measure()
        print(json.dumps({"single": one, "call1dtdse": dropin}), flush=True)
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
    from mma_amd import autotune, tdse, workloads as W

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
    # launch shape from the autotuner (a few steps of every instance per candidate), then one
    # persistent launch per bench step, replayed from a HIP graph
    t_tune = time.perf_counter()
    tuned = autotune.tune(geo, dev, n_steps=[n_steps] * n_inst) if args.launch_config is None else \
        autotune.TuneResult(tdse.LaunchConfig(*args.launch_config), [], False, None)
    t_tune = time.perf_counter() - t_tune
    log("launch shape %r (cached %s, %.2f s)" % (tuned.config, tuned.cached, t_tune))
    plan = tdse.BatchPlan(geo, dev, n_steps=[n_steps] * n_inst, out=out, config=tuned.config)
    stream = torch.cuda.current_stream()
    for w in range(args.warmup):
        log("warmup %d" % w)
        plan.replay(stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        plan.replay(stream)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    log("%d steps done" % args.steps)
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = wall / args.steps * 1e3
    value = pt_steps / (wall / args.steps)
    achieved = BYTES_PER_PT_STEP * pt_steps / (kern_ms * 1e-3) / 1e9
    gpu_rows = out[:, :4, :].cpu().numpy()  # for the parity check against the CPU reference members
    finite = bool(np.all(np.isfinite(gpu_rows)))

    stepwise = None
    if not args.no_stepwise:
        log("stepwise form")
        sw_wall, sw_diff = stepwise_measure(geo, dev, n_steps, out)
        stepwise = {"pt_steps_per_s": pt_steps / sw_wall, "seconds": sw_wall, "launches": n_steps,
                    "dipole_max_rel_diff_vs_persistent": sw_diff}
        log("stepwise: %.3f s" % sw_wall)
    plan.close()

    single, dropin = None, None
    if not args.no_single:
        single = single_measure(geo, dev[0], n_steps, nx)["pt_steps_per_s"]
        with stdout_to_stderr():
            dropin = call1dtdse_measure()

    uneven = None
    if not args.no_uneven:
        out = None
        torch.cuda.empty_cache()
        uneven = uneven_measure()  # tuned on its own (another batch signature)
        torch.cuda.empty_cache()

    adj = None
    if not args.no_adjoint:
        del out, dev
        torch.cuda.empty_cache()
        adj = adjoint_measure(args.adj_inst, args.adj_samples)

    hk = None
    if not args.no_hankel:
        torch.cuda.empty_cache()
        hk = hankel_measure(cpu=not args.no_cpu)
        torch.cuda.empty_cache()

    # the box's practical ceilings, same units as achieved: read + write (the propagation's
    # traffic mix), read only and write only
    copy_gbs, read_gbs, write_gbs = (1e3 * stream_rate(kind, 16, 32) for kind in ("copy", "read", "write"))
    torch.cuda.empty_cache()

    cpu, parity = None, None
    if not args.no_cpu:
        log("cpu baseline")
        with stdout_to_stderr():
            cpu, parity = cpu_baseline(gpu_rows=gpu_rows)

    line = {
        "metric": METRIC, "value": value, "unit": "grid-pt-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "tdse-batch-gaussian-beam", "instances": n_inst, "grid_r_x_z": [args.n_r, args.n_z],
                   "Nx": nx, "dx": g["dx"], "dt": 0.25, "time_steps": n_steps, "atom": "Ar (a=1.1893)",
                   "field": "800nm 10-cycle sin2, 1.5e14 W/cm2 peak, w0=100um", "precision": "complex128"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "copy_ceiling": copy_gbs, "frac_of_copy_ceiling": achieved / copy_gbs if copy_gbs else None,
                     "read_only_ceiling": read_gbs, "write_only_ceiling": write_gbs,
                     "traffic": MEASURED_TRAFFIC_BYTES.get(tuned.config.as_tuple()[2])
                     if (args.n_r, args.n_z) == (64, 64) else None,
                     "kernel": "tdse_propagate_pair_stash_kernel" if tuned.config.as_tuple()[2] == 2
                     else "tdse_propagate_kernel"},
        "cpu_baseline": cpu,
        "gpu_over_cpu_baseline": (value / cpu["value"]) if cpu else None,
        "parity_vs_reference": parity,
        "launch": {"form": "hipGraph replay of one persistent launch", "config": list(tuned.config.as_tuple()),
                   "tune_table": [list(r) for r in tuned.table[:6]], "tune_s": t_tune, "cached": tuned.cached},
        "stepwise_form": stepwise,
        "persistent_over_stepwise": (value / stepwise["pt_steps_per_s"]) if stepwise else None,
        "uneven": uneven,
        "per_instance_pt_steps_per_s": single,
        "call1dtdse": dropin,
        "batched_over_per_instance": (value / single) if single else None,
        "adjoint_over_forward": adj["adjoint_over_forward"] if adj else None,
        "adjoint": adj,
        "hankel": hk,
        "kernel_ms": kern_ms,
        "outputs_finite": finite,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
