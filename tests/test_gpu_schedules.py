"""Schedules that must not change a single bit: split steps replayed as captured
hipGraphs, the fused kernel with 128 or 256 threads per fibre, and the
autotuner (which runs trial steps on the batch and must leave its states,
stimulation state machine, log and spike counters as they were).

Each graph test replays the graph many times across several advance calls with
host reads in between and checks every call against the eagerly launched run.
"""
import numpy as np
import pytest

from cases import uneven, vc_lane_stim_range, firing_table

pytestmark = pytest.mark.gpu


def _batch(mesh, **kw):
    from opendihu_amd.fibers import FiberBatch
    return FiberBatch(**mesh, **kw)


def _firing(batch, mesh):
    n = mesh["n_nodes"]
    F = n.size
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    sb, se = vc_lane_stim_range(off, n // 2, n, 4)
    fe = firing_table(40, 10, 4.0, 0.1, seed=2)
    mu = (np.arange(F) * 7) % 10 + 1
    jit = [[0.1, -0.1, 0.05][: (f % 4)] for f in range(F)]
    batch.set_firing_table_stimulation(sb, se, fe, mu, call_frequency=1.0, repeat_after_first_call=0.1,
                                       call_enable_begin=0.2 * (np.arange(F) % 3), jitter_lists=jit,
                                       log_capacity=8192)


@pytest.mark.parametrize("pair", [1, 0])
@pytest.mark.parametrize("thomas", ["serial", "partitioned"])
def test_graph_replay_equals_eager(thomas, pair):
    mesh, _ = uneven(12, lo=200, hi=600, seed=9)
    runs = {}
    for mode in ("eager", "graph"):
        g = _batch(mesh, scheme="IE", kernels="split", thomas=thomas)
        g.set_tuning(pair_halves=pair)
        if mode == "graph":
            g.set_tuning(graph_steps=8)
        _firing(g, mesh)
        g.enable_spike_detection(0.0)
        t, dt, per_call = 0.0, 2e-3, []
        for k in (37, 8, 50, 3, 400, 1002):
            g.advance(t, dt, k)
            t += k * dt
            per_call.append(g.get_states())
        runs[mode] = (per_call, g.spikes(), g.stimulation_log(), g.graph_counts())
    for c, (a, b) in enumerate(zip(runs["eager"][0], runs["graph"][0])):
        np.testing.assert_array_equal(a, b, err_msg=f"advance call {c}")
    for a, b in zip(runs["eager"][1], runs["graph"][1]):
        np.testing.assert_array_equal(a, b)
    ga = sorted(zip(runs["graph"][2][1].tolist(), runs["graph"][2][0].tolist()))
    ea = sorted(zip(runs["eager"][2][1].tolist(), runs["eager"][2][0].tolist()))
    assert ga == ea and len(ea) > 0
    counts = runs["graph"][3]
    print("graph counts", counts)
    assert counts["captures"] == 1, "the captured graph is reused across advance calls"
    # paired: the K-1 bodies between the call's first and last half are replayed
    assert counts["replays"] == sum((k - pair) // 8 for k in (37, 8, 50, 3, 400, 1002))
    assert runs["eager"][3]["replays"] == 0


@pytest.mark.parametrize("reaction_threads", [64, 256])
def test_graph_replay_equals_eager_large_model(reaction_threads):
    """the 57-state TK2014 library (per-fibre reaction kernel, paired halves):
    graph replays bit-identical to eager launches across advance calls"""
    import os
    from opendihu_amd.build import LIB_DIR, TK2014_LIBRARY
    if not os.path.exists(os.path.join(LIB_DIR, TK2014_LIBRARY)):
        pytest.skip("TK2014 library not built (needs the reference checkout at build time)")
    mesh, _ = uneven(8, lo=200, hi=600, seed=11)
    calls = (37, 8, 50, 3, 300)
    runs = {}
    for mode in ("eager", "graph"):
        g = _batch(mesh, scheme="IE", kernels="split", thomas="partitioned", library=TK2014_LIBRARY)
        g.set_tuning(reaction_per_fibre=1, reaction_threads=reaction_threads, graph_steps=8 if mode == "graph" else 0)
        _firing(g, mesh)
        g.enable_spike_detection(-40.0)
        t, dt, per_call = 0.0, 2.5e-5, []
        for k in calls:
            g.advance(t, dt, k)
            t += k * dt
            per_call.append(g.get_states())
        runs[mode] = (per_call, g.spikes(), g.graph_counts())
        g.close()
    for c, (a, b) in enumerate(zip(runs["eager"][0], runs["graph"][0])):
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, b, err_msg=f"advance call {c}")
    for a, b in zip(runs["eager"][1], runs["graph"][1]):
        np.testing.assert_array_equal(a, b)
    assert runs["graph"][2]["replays"] == sum((k - 1) // 8 for k in calls)


def test_graph_follows_stimulation_and_spike_changes():
    """new stimulation values are read from device memory by the same graph;
    switching the spike detector on changes the launch sequence and recaptures"""
    mesh, stim = uneven(6, lo=300, hi=500, seed=3)
    out = []
    for gs in (0, 16):
        g = _batch(mesh, scheme="CN", kernels="split", thomas="partitioned")
        g.set_tuning(graph_steps=gs)
        g.set_current_stimulation(**stim)
        g.advance(0.0, 1e-3, 64)
        st2 = dict(stim, onset=stim["onset"] + 0.5, amplitude=stim["amplitude"] * 0.8)
        g.set_current_stimulation(**st2)
        g.advance(0.064, 1e-3, 640)
        c1 = g.graph_counts()
        g.enable_spike_detection(0.0)
        g.advance(0.704, 1e-3, 300)
        out.append((g.get_states(), g.spikes(), c1, g.graph_counts()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)
    assert out[1][2]["captures"] == 1 and out[1][3]["captures"] == 2


@pytest.mark.parametrize("thomas", ["serial", "partitioned"])
@pytest.mark.parametrize("resident", [0, 1])
def test_fused_128_threads_equals_256(thomas, resident):
    """serial solve: bit-identical; partitioned solve: the partition follows the
    thread count (another elimination order), so the bar is the partitioned one"""
    mesh, stim = uneven(10, lo=300, hi=1000 if thomas == "serial" or resident else 2000, seed=8)
    runs = []
    for threads in (256, 128):
        g = _batch(mesh, scheme="CN", thomas=thomas, kernels="fused")
        g.set_tuning(threads=threads, resident_gates=resident, steps_per_launch=77)
        g.set_current_stimulation(**stim)
        g.enable_spike_detection(0.0)
        g.advance(0.0, 1e-3, 700)
        runs.append((g.get_states(), g.spikes()))
    if thomas == "serial":
        np.testing.assert_array_equal(runs[0][0], runs[1][0])
    else:
        err = np.linalg.norm(runs[0][0][0] - runs[1][0][0]) / np.linalg.norm(runs[0][0][0])
        print(f"128 vs 256 threads, partitioned: Vm rel L2 {err:.3e}")
        assert err <= 1e-10
    np.testing.assert_array_equal(runs[0][1][0], runs[1][1][0])


def test_autotune_keeps_state_and_bits():
    mesh, _ = uneven(16, lo=300, hi=900, seed=5)
    dt = 2e-3
    trial = np.stack([np.arange(40) * dt + 1.0, np.full(40, dt), np.full(40, dt / 2), np.full(40, dt)], 1)
    ref = _batch(mesh, scheme="IE", kernels="split", thomas="serial")
    tuned = _batch(mesh, scheme="IE", kernels="split", thomas="serial")
    for b in (ref, tuned):
        _firing(b, mesh)
        b.enable_spike_detection(0.0)
        b.advance(0.0, dt, 300)
    before = (tuned.get_states(), tuned.spikes(), tuned.stimulation_log())
    choice = tuned.autotune(trial, allow_split=True)
    print("autotune choice", choice)
    after = (tuned.get_states(), tuned.spikes(), tuned.stimulation_log())
    np.testing.assert_array_equal(before[0], after[0])
    for a, b in zip(before[1], after[1]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(before[2], after[2]):
        np.testing.assert_array_equal(a, b)
    assert len(choice["candidates_ms"]) >= 3 and all(np.isfinite(choice["candidates_ms"]))
    assert choice["rejected"] == 0 and min(choice["candidates_ms"]) > 0, choice
    for b in (ref, tuned):
        b.advance(300 * dt, dt, 700)
    np.testing.assert_array_equal(ref.get_states(), tuned.get_states())
    for a, b in zip(ref.spikes(), tuned.spikes()):
        np.testing.assert_array_equal(a, b)
    ra, ta = ref.stimulation_log(), tuned.stimulation_log()
    assert sorted(zip(ra[1].tolist(), ra[0].tolist())) == sorted(zip(ta[1].tolist(), ta[0].tolist()))


@pytest.mark.parametrize("thomas", ["serial", "partitioned"])
def test_paired_halves_equal_unpaired(thomas):
    """the second reaction half of step q and the first of step q+1 in one
    launch: states, spikes, stimulation log and algebraics bit-identical to one
    launch per half, over calls of 1..400 steps, with n0D = 1 and 2, eager and
    replayed as graphs"""
    mesh, _ = uneven(10, lo=200, hi=700, seed=12)
    runs = {}
    for pair in (0, 1):
        for gs in (0, 4):
            g = _batch(mesh, scheme="CN", kernels="split", thomas=thomas)
            g.set_tuning(pair_halves=pair, graph_steps=gs)
            _firing(g, mesh)
            g.enable_spike_detection(-20.0)
            g.set_algebraics_for_transfer([0, 3])
            t, dt, out = 0.0, 2e-3, []
            for k, n0D in ((1, 1), (9, 2), (2, 1), (400, 1), (1, 2), (57, 1)):
                g.advance(t, dt, k, dt0D=dt / 2 / n0D if n0D > 1 else None, n0D=n0D)
                t += k * dt
                out.append((g.get_states(), g.get_algebraics()))
            runs[(pair, gs)] = (out, g.spikes(), g.stimulation_log())
            g.close()
    ref = runs[(0, 0)]
    assert (ref[1][1] > 0).sum() > 0, "the stimulus must trigger action potentials"
    for key, run in runs.items():
        for c, ((a, aa), (b, bb)) in enumerate(zip(ref[0], run[0])):
            np.testing.assert_array_equal(a, b, err_msg=f"{key} call {c}")
            np.testing.assert_array_equal(aa, bb, err_msg=f"{key} call {c} algebraics")
        for a, b in zip(ref[1], run[1]):
            np.testing.assert_array_equal(a, b, err_msg=f"{key} spikes")
        ea = sorted(zip(ref[2][1].tolist(), ref[2][0].tolist()))
        ga = sorted(zip(run[2][1].tolist(), run[2][0].tolist()))
        assert ea == ga and len(ea) > 0, key


def test_autotune_rejects_schedules_with_other_bits():
    """the autotuner compares every candidate's end state with the solver's own
    schedule's: a 128-thread fused schedule with the partitioned solve
    partitions each fibre differently from every 256-thread candidate (a few ulp
    apart), so all candidates are rejected and the solver keeps its schedule"""
    mesh, stim = uneven(16, lo=300, hi=900, seed=5)
    g = _batch(mesh, scheme="CN", kernels="fused", thomas="partitioned")
    g.set_tuning(threads=128)
    g.set_current_stimulation(**stim)
    g.advance(0.0, 1e-3, 400)
    dt = 1e-3
    trial = np.stack([np.arange(40) * dt + 0.4, np.full(40, dt), np.full(40, dt / 2), np.full(40, dt)], 1)
    choice = g.autotune(trial, allow_split=True)
    print("autotune choice", choice)
    assert choice["rejected"] == len(choice["candidates_ms"]) >= 3, choice
    assert all(m == -1.0 for m in choice["candidates_ms"]) and choice["ms_per_step"] == -1.0
    assert choice["kernels"] == "fused" and choice["threads"] == 128


def test_kernel_timing_changes_no_bit():
    """per-launch timing (a start/stop event pair bound to each launch, read
    back when the pool of 256 pairs is full and by kernel_times) launches the
    same kernels: states bit-identical with it on and off; every launch is
    counted once, with a positive duration, and the durations of a call add up
    to less than its wall-clock time"""
    import time
    mesh, stim = uneven(16, lo=300, hi=900, seed=4)
    states = {}
    for prof in (0, 1):
        for kernels in ("split", "fused"):
            g = _batch(mesh, scheme="CN", kernels=kernels, thomas="partitioned")
            g.set_current_stimulation(**stim)
            g.profile_kernels(bool(prof))
            g.synchronize()
            t0 = time.perf_counter()
            g.advance(0.0, 1e-3, 300)   # > 256 timed launches on the split kernels
            g.synchronize()
            wall_ms = (time.perf_counter() - t0) * 1e3
            times = g.kernel_times(reset=True)
            g.advance(0.3, 1e-3, 7)
            times2 = g.kernel_times()
            states[(prof, kernels)] = g.get_states()
            g.close()
            print(prof, kernels, times, times2)
            if not prof:
                assert all(v == (0.0, 0) for v in times.values())
                continue
            assert sum(ms for ms, _ in times.values()) < wall_ms
            if kernels == "split":
                # paired halves: the call's first and last half alone, K - 1 pairs,
                # K solves, a stimulation-flag launch per reaction launch
                assert times["reaction"][1] == 2 and times["reaction_pair"][1] == 299
                assert times["diffusion"][1] == 300 and times["other"][1] == 301
                assert times2["reaction_pair"][1] == 6 and times2["diffusion"][1] == 7
            else:
                assert times["fused"][1] >= 1 and times2["fused"][1] >= 1
            for k, (ms, n) in list(times.items()) + list(times2.items()):
                assert (ms > 0) == (n > 0), (k, ms, n)
    for kernels in ("split", "fused"):
        np.testing.assert_array_equal(states[(0, kernels)], states[(1, kernels)], err_msg=kernels)
