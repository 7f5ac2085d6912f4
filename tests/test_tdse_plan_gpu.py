"""GPU checks of the prepared-batch path: geometries built many at a time, the plan's
enqueue and its HIP-graph replay against the immediate entry point, the launch-shape
autotuner, and members of the uneven workload (Nx, duration, intensity and wavelength drawn
per instance) in one launch against the CPU oracle."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from mma_amd import autotune, tdse, workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-8
ROWS = ["sourceterm", "PopTot", "PopInt", "expval", "Efield"]


def _grid_geometry(num_r, dt=0.25):
    return tdse.Geometry(num_r, 0.1, dt, W.GRID["a"], W.GRID["e_guess"], 1e-25, 2.0)


def test_create_many_equals_one_at_a_time():
    """Ground states run side by side give the same bits as separate launches."""
    sizes = [1200, 2000, 3602, 800]
    many = tdse.Geometry.create_many(sizes, 0.1, 0.25, W.GRID["a"], W.GRID["e_guess"], 1e-25, 2.0)
    for n, g in zip(sizes, many):
        one = _grid_geometry(n)
        assert g.status == 0
        assert g.e_init == one.e_init and g.gs_iterations == one.gs_iterations
        assert g.halo_points == one.halo_points
        xa, pa = g.ground_state()
        xb, pb = one.ground_state()
        np.testing.assert_array_equal(xa, xb)
        np.testing.assert_array_equal(pa, pb)


def test_shared_lattice_tables_match_own_tables():
    """Geometries built together on one lattice (equal dx, dt, a and num_r parity) read g, V'
    and 1/beta of the widest one, each from the alignment that puts its chunks on shared chunk
    boundaries, with its own pivots in its first H chunks; the results equal those of the same
    geometries built one at a time (own tables) to rounding.  Offsets 0..7 into the widest even
    grid (every alignment), an odd group, one and two steps per pass."""
    import os
    even = [3000 - 2 * k for k in range(8)]
    odd = [2001, 1999, 1501]
    sizes = even + odd
    shared = tdse.Geometry.create_many(sizes, 0.1, 0.25, W.GRID["a"], W.GRID["e_guess"], 1e-25, 2.0)
    own = [_grid_geometry(n) for n in sizes]
    rng = np.random.default_rng(21)
    t = np.arange(420) * 0.25
    f = np.stack([W.sin2_pulse(t, rng.uniform(0.04, 0.09), W.omega_au(800e-9), 3, phase=rng.uniform(0, 3))
                  for _ in sizes])
    dev = torch.tensor(tdse.pad_fields(f), device="cuda")
    steps = [dev.shape[1] - 1 - (i % 3) for i in range(len(sizes))]
    for spp in ("1", "2"):
        os.environ["MMA_TDSE_STEPS_PER_PASS"] = spp
        try:
            a = tdse.propagate_batch(shared, dev, n_steps=steps).cpu().numpy()
            b = tdse.propagate_batch(own, dev, n_steps=steps).cpu().numpy()
        finally:
            del os.environ["MMA_TDSE_STEPS_PER_PASS"]
        for i, n in enumerate(sizes):
            for r, name in enumerate(ROWS[:4]):
                e = rel_err(a[i, r, :steps[i]], b[i, r, :steps[i]])
                assert e < 1e-11, (spp, n, name, e)
        print("steps/pass", spp, "shared vs own tables: max %.2e" % max(
            rel_err(a[i, 0, :steps[i]], b[i, 0, :steps[i]]) for i in range(len(sizes))), flush=True)


def test_instance_queues_give_every_instance_its_own_bits():
    """The persistent launch takes instances from one queue per XCD: members of a shared lattice
    by their alignment (here all in one alignment, so the other queues' workgroups take them
    from that queue), the rest dealt round-robin.  Every instance of a batch mixing both gets
    the bits it gets alone (same dx and dt, so the same halo and launch shape)."""
    import os
    lattice = tdse.Geometry.create_many([3000, 2984, 2968, 2952], 0.1, 0.25, W.GRID["a"], W.GRID["e_guess"],
                                        1e-25, 2.0)
    own = [_grid_geometry(n) for n in (2001, 1600)]
    geoms = [lattice[i % 4] for i in range(12)] + [own[i % 2] for i in range(6)]
    assert len({g.halo_points for g in geoms}) == 1
    rng = np.random.default_rng(33)
    t = np.arange(300) * 0.25
    f = np.stack([W.sin2_pulse(t, rng.uniform(0.04, 0.09), W.omega_au(800e-9), 3, phase=rng.uniform(0, 3))
                  for _ in geoms])
    dev = torch.tensor(tdse.pad_fields(f), device="cuda")
    for spp in ("1", "2"):
        os.environ["MMA_TDSE_STEPS_PER_PASS"] = spp
        try:
            batch = tdse.propagate_batch(geoms, dev).cpu().numpy()
            for i, g in enumerate(geoms):
                alone = tdse.propagate_batch([g], dev[i:i + 1]).cpu().numpy()
                np.testing.assert_array_equal(batch[i], alone[0], err_msg="instance %d, steps/pass %s" % (i, spp))
        finally:
            del os.environ["MMA_TDSE_STEPS_PER_PASS"]


def test_create_many_reports_bad_members():
    with pytest.raises(tdse.CTDSEError):
        tdse.Geometry.create_many([1000, 1], 0.1, 0.25, W.GRID["a"])


def test_plan_enqueue_and_graph_replay_match_immediate():
    """The plan's enqueue and its graph give the immediate entry point's bits; replays are
    interleaved with other device work and a refilled field buffer, and every replay is
    checked against the eager result for the fields it read."""
    t, fields = W.gaussian_beam_batch(n_r=4, n_z=3, n_cycles=2)
    geo = _grid_geometry(3000)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    ref = tdse.propagate_batch(geo, dev).clone()
    alt = torch.tensor(tdse.pad_fields(0.7 * fields[::-1].copy()), device="cuda")
    ref_alt = tdse.propagate_batch(geo, alt).clone()

    buf = dev.clone()
    plan = tdse.BatchPlan(geo, buf)
    out = plan.enqueue()
    torch.cuda.synchronize()
    torch.testing.assert_close(out, ref, rtol=0, atol=0)
    for rep in range(4):
        src = alt if rep % 2 else dev
        want = ref_alt if rep % 2 else ref
        buf.copy_(src)
        plan.out.fill_(float("nan"))
        junk = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        plan.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        torch.testing.assert_close(plan.out, want, rtol=0, atol=0)
    plan.close()


def test_plan_rejects_bad_launch_shapes():
    geo = _grid_geometry(1000)
    dev = torch.zeros((2, 50), dtype=torch.float64, device="cuda")
    with pytest.raises(tdse.CTDSEError):
        tdse.BatchPlan(geo, dev, config=(9, 1, 1))
    with pytest.raises(tdse.CTDSEError):
        tdse.BatchPlan(geo, dev, config=(4, 0, 1))
    with pytest.raises(tdse.CTDSEError):
        tdse.BatchPlan(geo, dev, config=(4, 2, 3))


def test_launch_shapes_agree_with_oracle():
    """Every launch shape the autotuner may choose gives the oracle's result (the segment
    partition moves the truncation points, so agreement is to rounding, not bitwise)."""
    from oracle import ctdse as O
    t, fields = W.gaussian_beam_batch(n_r=2, n_z=2, n_cycles=2)
    geo = _grid_geometry(4000)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    n_steps = dev.shape[1] - 1
    x, psi0 = geo.ground_state()
    ref = O.propagate(4000, 0.1, W.GRID["a"], 2.0, 0.25, psi0, x, tdse.pad_fields(fields[3])[0], n_steps)
    for shape in autotune.candidates(geo.halo_points // 8):
        plan = tdse.BatchPlan(geo, dev, config=shape)
        assert plan.config.as_tuple() == shape
        o = plan.enqueue().cpu().numpy()
        torch.cuda.synchronize()
        for r, k in enumerate(ROWS):
            assert rel_err(o[3, r], ref[k][:n_steps]) < TOL, (shape, k)
        plan.close()


def test_autotune_returns_a_timed_candidate(tmp_path, monkeypatch):
    monkeypatch.setenv("MMA_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    t, fields = W.gaussian_beam_batch(n_r=8, n_z=8, n_cycles=2)
    geo = _grid_geometry(4000)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    pairs = [(4, 3, 1), (4, 2, 1), (8, 1, 1), (2, 6, 1), (4, 2, 2)]
    res = autotune.tune(geo, dev, trial_steps=16, pairs=pairs)
    assert not res.cached
    assert res.config.as_tuple() in pairs
    assert [r[:3] for r in res.table if r[:3] == res.config.as_tuple()]
    assert all(r[3] > 0 for r in res.table) and res.table[0][3] == max(r[3] for r in res.table)
    again = autotune.tune(geo, dev, trial_steps=16)
    assert again.cached is False  # explicit pairs are not cached as the full search
    res2 = autotune.tune(geo, dev, trial_steps=16)
    assert res2.cached and res2.config.as_tuple() == again.config.as_tuple()
    plan, _ = autotune.tuned_plan(geo, dev)
    ref = tdse.propagate_batch(geo, dev)
    o = plan.replay()
    torch.cuda.synchronize()
    assert float((o - ref).abs().max()) <= 1e-10 * float(ref.abs().max())


def test_uneven_workload_members_match_oracle():
    """Members of the seeded uneven workload (own Nx, cycles, intensity, wavelength) in one
    launch through a plan with the shape the autotuner picks for the whole workload (4 waves x
    2 workgroups per CU, two steps per pass) and the batch's shared lattice tables: the three
    smallest against the oracle over their whole run, the widest (Nx 31997) and the one with the
    most work (Nx 31723, 19477 steps) over their first 1500 steps, and every member conserves the
    norm."""
    from oracle import ctdse as O
    num_r, n_e = W.uneven_work()
    work = (num_r + 1) * (n_e + 2)
    pick = [int(i) for i in np.argsort(work)[:12]] + [int(np.argmax(num_r)), int(np.argmax(work))]
    sel = W.uneven_batch(indices=pick)
    geoms = tdse.Geometry.create_many([m["num_r"] for m in sel], 0.1, 0.25, W.GRID["a"], W.GRID["e_guess"],
                                      W.UNEVEN_CV, 2.0)
    fields, steps = W.stack_fields([m["field"] for m in sel])
    dev = torch.tensor(fields, device="cuda")
    plan = tdse.BatchPlan(geoms, dev, n_steps=steps, config=(4, 2, 2))
    out = plan.replay().cpu().numpy()
    for i, s in enumerate(steps):  # instance i owns columns [0, steps[i])
        assert np.all(out[i, 1, :s] <= 1.0 + 1e-12) and np.all(np.isfinite(out[i, :, :s])), pick[i]
    widest, longest = len(sel) - 2, len(sel) - 1
    assert sel[widest]["num_r"] + 1 == 31997 and steps[longest] == 19477
    for i, window in [(0, None), (1, None), (2, None), (widest, 1500), (longest, 1500)]:
        m, g = sel[i], geoms[i]
        s = steps[i] if window is None else window
        x, psi0 = g.ground_state()
        ref = O.propagate(m["num_r"], 0.1, W.GRID["a"], 2.0, 0.25, psi0, x, fields[i], s)
        for r, k in enumerate(ROWS):
            e = rel_err(out[i, r, :s], ref[k][:s])
            print(pick[i], m["num_r"], s, k, "%.2e" % e, flush=True)
            assert e < TOL, (pick[i], k)
    plan.close()


def test_stepwise_form_equals_persistent_launch():
    """One launch per time step (states stepped in place in HBM) gives the persistent
    launch's bits."""
    t, fields = W.gaussian_beam_batch(n_r=3, n_z=2, n_cycles=1)
    geo = _grid_geometry(2400)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    ref = tdse.propagate_batch(geo, dev)
    got = tdse.propagate_stepwise(geo, dev)
    torch.cuda.synchronize()
    torch.testing.assert_close(got, ref, rtol=0, atol=0)


def _single_vs_pair(monkeypatch, run):
    monkeypatch.setenv("MMA_TDSE_STEPS_PER_PASS", "1")
    a = run()
    monkeypatch.setenv("MMA_TDSE_STEPS_PER_PASS", "2")
    b = run()
    monkeypatch.delenv("MMA_TDSE_STEPS_PER_PASS")
    return a, b


def test_two_step_passes_give_single_step_bits(monkeypatch):
    """Two steps per pass (the intermediate state kept on chip) give the bits of one step
    per pass: the same segment arithmetic runs on the same values."""
    t, fields = W.gaussian_beam_batch(n_r=3, n_z=3, n_cycles=2)
    geo = _grid_geometry(5000)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    a, b = _single_vs_pair(monkeypatch, lambda: tdse.propagate_batch(geo, dev).clone())
    torch.testing.assert_close(b, a, rtol=0, atol=0)
    # odd and even step counts (a final single step), several geometries in one launch
    geoms = tdse.Geometry.create_many([3000, 2001 * 2, 1600], 0.1, 0.25, W.GRID["a"], W.GRID["e_guess"], 1e-25, 2.0)
    f3 = dev[:3].contiguous()
    a, b = _single_vs_pair(monkeypatch,
                           lambda: tdse.propagate_batch(geoms, f3, n_steps=[f3.shape[1] - 1, f3.shape[1] - 2, 301]).clone())
    torch.testing.assert_close(b, a, rtol=0, atol=0)


def test_two_step_passes_with_saves_partial_runs_and_snapshots(monkeypatch, dll):
    """Saves after odd steps, partial runs of odd length from a saved state, and snapshots
    every 81 steps force single steps in between pairs; every output agrees bit for bit."""
    from mma_amd import adjoint as AD
    from conftest import golden
    t, fields = W.gaussian_beam_batch(n_r=2, n_z=2, n_cycles=2)
    geo = _grid_geometry(3000)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    F = AD.gate(dev)[0].contiguous()
    n = F.shape[1] - 1
    S = AD.state_length(geo)

    def partial():
        save = torch.zeros((4, 9, S, 2), dtype=torch.float64, device="cuda")
        out = torch.zeros((4, 5, n), dtype=torch.float64, device="cuda")
        AD.run_states(geo, F, n, 0, 63, None, save, 7, out=out)
        save2 = torch.zeros((4, 3, S, 2), dtype=torch.float64, device="cuda")
        AD.run_states(geo, F, n, 63, 41, save, save2, 13, out=out, init_slot=8)
        torch.cuda.synchronize()
        return torch.cat([save.flatten(), save2.flatten(), out.flatten()])

    a, b = _single_vs_pair(monkeypatch, partial)
    torch.testing.assert_close(b, a, rtol=0, atol=0)

    g = golden("ref_short_odd.npz")
    from test_tdse_gpu import run_case

    def snaps():
        inputs, outputs = run_case(dll, g, writewft=1, tprint=20.25)
        wf = outputs.get_wavefunction(inputs, grids=False)
        res = np.concatenate([outputs.get_sourceterm(), outputs.get_PopTot(), wf.ravel().view(np.float64)])
        dll.free_mtrx(outputs.psi, len(wf))
        outputs.delete(dll)
        inputs.delete(dll)
        return res

    a, b = _single_vs_pair(monkeypatch, snaps)
    np.testing.assert_array_equal(a, b)


def test_plan_enqueue_captured_by_torch_graph():
    """The plan's enqueue (no allocation, copy or wait) recorded into a torch.cuda.CUDAGraph:
    every replay, with eager work and a refilled field buffer in between, equals the eager
    result for the fields it read; destroying the plan after an unsynchronised run is safe."""
    t, fields = W.gaussian_beam_batch(n_r=3, n_z=2, n_cycles=2)
    geo = _grid_geometry(2400)
    dev = torch.tensor(tdse.pad_fields(fields), device="cuda")
    alt = torch.tensor(tdse.pad_fields(0.5 * fields), device="cuda")
    ref = tdse.propagate_batch(geo, dev).clone()
    ref_alt = tdse.propagate_batch(geo, alt).clone()
    buf = dev.clone()
    plan = tdse.BatchPlan(geo, buf)
    plan.enqueue()  # warm-up outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.enqueue()
    for rep in range(4):
        buf.copy_(alt if rep % 2 else dev)
        plan.out.zero_()
        torch.cuda.synchronize()
        g.replay()
        junk = (torch.arange(1000, device="cuda", dtype=torch.float64) ** 2).sum()
        torch.cuda.synchronize()
        assert float(junk) > 0
        torch.testing.assert_close(plan.out, ref_alt if rep % 2 else ref, rtol=0, atol=0)
    del g
    plan.replay()
    plan.close()  # waits for the replay still in flight
