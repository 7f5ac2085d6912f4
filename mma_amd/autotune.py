"""Launch-shape autotuner of the persistent TDSE propagator.

The propagation kernel's launch parameters trade against each other:

* waves per workgroup sets the partition of the parallel tridiagonal solve: a segment is
  64 * waves chunks of 8 points, of which the truncation halo (H chunks, fixed by the
  geometry's decay rate) is recomputed by the next segment, so wider segments waste fewer
  lanes but synchronise more waves at every scan;
* workgroups per compute unit sets how many instances share a CU and hide each other's
  load and barrier stalls (the one-step kernel's 168-register budget allows 12 resident
  waves per CU, the two-step kernel's 256-register budget 8);
* steps per pass: the two-step kernel keeps the intermediate state on chip and streams the
  state through HBM half as often, at lower occupancy.

The best pair depends on the grid sizes, the halo and the batch size, so it is measured:
every candidate runs the caller's own batch (same geometries, fields and queue order)
for a few time steps, timed with HIP events on the launch stream, and the fastest in
grid-point-steps per second wins.  Results are cached per device and batch signature in a
JSON file ($MMA_AMD_TUNE_CACHE, default ~/.cache/mma_amd/tdse_launch.json).
"""
import hashlib
import json
import os

import torch

from . import tdse

MAX_WAVES = 8            # TDSE_MAX_WAVES of the kernel
RESIDENT_WAVES_PER_CU = {1: 12, 2: 8}  # waves per CU the register budget of each kernel allows
CHUNK = 8


def candidates(halo_lanes, max_waves=MAX_WAVES, resident=RESIDENT_WAVES_PER_CU, steps=(1, 2)):
    """(waves_per_group, groups_per_cu, steps_per_pass) shapes worth timing: segments that
    keep at least 64 own chunks after the halo, and no more workgroups per CU than can be
    resident."""
    out = []
    for sp in steps:
        for nw in (2, 3, 4, 5, 6, 8):
            if nw > max_waves or nw * 64 - halo_lanes < 64:
                continue
            for per_cu in range(1, max(1, resident[sp] // nw) + 1):
                out.append((nw, per_cu, sp))
    return out


def signature(device_name, cus, nx, n_steps):
    """Cache key of a batch: device, CU count, batch size and the spread of its work."""
    nx = sorted(int(v) for v in nx)
    st = sorted(int(v) for v in n_steps)
    q = lambda v: [v[0], v[len(v) // 2], v[-1]] if v else []
    blob = json.dumps({"dev": device_name, "cus": int(cus), "n": len(nx), "nx": q(nx),
                       "steps": [int(round(s, -2)) for s in q(st)]}, sort_keys=True)
    return hashlib.sha1(blob.encode()).hexdigest()[:16], blob


def _cache_path():
    return os.environ.get("MMA_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mma_amd",
                                                             "tdse_launch.json"))


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, entry):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_cache()
        data[key] = entry
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


class TuneResult:
    def __init__(self, config, table, cached, key):
        self.config = config      # tdse.LaunchConfig
        self.table = table        # [(waves, groups_per_cu, steps_per_pass, pt_steps_per_s)], fastest first
        self.cached = cached
        self.key = key

    def __repr__(self):
        return "TuneResult(%r, cached=%s, %d candidates)" % (self.config, self.cached, len(self.table))


def tune(geometries, fields, n_steps=None, trial_steps=48, repeats=2, use_cache=True, stream=None, pairs=None):
    """Pick the launch shape for a batch.

    geometries: one tdse.Geometry per row of `fields` (or one for all);
    fields:     the batch's (n_inst, L) float64 CUDA field tensor;
    n_steps:    the steps each instance will run (default L - 1).
    Each candidate runs min(n_steps[i], trial_steps) steps of every instance into a
    scratch output; the best of `repeats` timings counts.  pairs: time only these
    (waves_per_group, groups_per_cu, steps_per_pass) shapes (neither read from nor written to
    the cache).
    Returns a TuneResult."""
    n_inst, L_f = fields.shape
    geoms = geometries if isinstance(geometries, (list, tuple)) else [geometries] * n_inst
    if n_steps is None:
        n_steps = [L_f - 1] * n_inst
    dev = fields.device
    props = torch.cuda.get_device_properties(dev)
    key, blob = signature(props.name, props.multi_processor_count, [g.num_r + 1 for g in geoms], n_steps)
    if use_cache and pairs is None:
        hit = _load_cache().get(key)
        if hit:
            return TuneResult(tdse.LaunchConfig(*hit["config"]), [tuple(r) for r in hit["table"]], True, key)
    halo = max(g.halo_points // CHUNK for g in geoms)
    full_search = pairs is None
    pairs = list(pairs) if pairs is not None else candidates(halo)
    trial = [max(1, min(int(s), int(trial_steps))) for s in n_steps]  # even: whole pairs
    width = max(trial)
    scratch = torch.empty((n_inst, 5, width), dtype=torch.float64, device=dev)
    stream = stream if stream is not None else torch.cuda.current_stream(dev)
    table = []
    for shape in pairs:
        plan = tdse.BatchPlan(geoms, fields, n_steps=trial, out=scratch, config=shape)
        best = None
        with torch.cuda.stream(stream):
            plan.enqueue(stream)  # untimed first run
            for _ in range(max(1, repeats)):
                t0 = torch.cuda.Event(enable_timing=True)
                t1 = torch.cuda.Event(enable_timing=True)
                t0.record(stream)
                plan.enqueue(stream)
                t1.record(stream)
                t1.synchronize()
                ms = t0.elapsed_time(t1)
                best = ms if best is None else min(best, ms)
        plan.close()
        table.append(tuple(plan.config.as_tuple()) + (plan.pt_steps / (best * 1e-3),))
    table.sort(key=lambda r: -r[-1])
    cfg = tdse.LaunchConfig(*table[0][:3])
    if use_cache and full_search:
        _store_cache(key, {"config": list(cfg.as_tuple()), "table": [list(r) for r in table], "batch": blob})
    return TuneResult(cfg, table, False, key)


def tuned_plan(geometries, fields, n_steps=None, out=None, **kw):
    """tune() then a tdse.BatchPlan with the chosen shape."""
    res = tune(geometries, fields, n_steps=n_steps, **kw)
    return tdse.BatchPlan(geometries, fields, n_steps=n_steps, out=out, config=res.config), res


__all__ = ["candidates", "signature", "tune", "tuned_plan", "TuneResult"]
