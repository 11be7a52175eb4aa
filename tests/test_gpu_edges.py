"""Fibre lengths at the edges of the kernels' assumptions, in one batch, against
the oracle:

* 2 and 3 nodes (one chunk per fibre in the partitioned solve, P = n/2 threads);
* one node either side of 64, 128, 256, 512, 1024 and 2048 (the per-fibre
  reaction's 64-node passes and 32-node alignment shadows, the partitioned
  solve's chunk counts and its register arrays of 4, 8 and 16 rows);
* 4096 nodes, the longest fibre the fused kernel and the partitioned solve take.

Serial solve: bit-identical states and spike steps; partitioned solve: relative
L2 error <= 1e-10 and equal spike steps. The 56-state model runs the same
lengths through the per-fibre reaction with 64 and 256 threads.
"""
import os

import numpy as np
import pytest

from oracle_binding import OracleBatch, ORACLE_BUILD

pytestmark = pytest.mark.gpu

LENGTHS = [2, 3, 4, 5, 63, 64, 65, 127, 129, 255, 256, 257, 511, 513, 1023, 1025, 2047, 2048, 2049, 4096]
SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"


def _mesh(lengths, h=0.01):
    n = np.array(lengths, np.int32)
    F = n.size
    elem = np.concatenate([np.full(k - 1, h) for k in n])
    pf = np.random.default_rng(1).uniform(3.0, 4.5, F) / 500.0
    nmj = n // 2
    sb = np.maximum(nmj - 1, 0).astype(np.int32)
    se = np.minimum(nmj + 2, n).astype(np.int32)
    onset = np.random.default_rng(2).uniform(0.0, 0.3, F)
    stim = dict(stim_begin=sb, stim_end=se, onset=onset, duration=np.full(F, 0.1), amplitude=np.full(F, 1200.0),
                stim_param=0)
    return dict(n_nodes=n, element_lengths=elem, prefactor=pf), stim


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kernels,thomas", [("split", "serial"), ("split", "partitioned"), ("fused", "partitioned")])
def test_edge_lengths_hh(kernels, thomas):
    from opendihu_amd.fibers import FiberBatch
    mesh, stim = _mesh(LENGTHS)
    g = FiberBatch(**mesh, scheme="CN", kernels=kernels, thomas=thomas)
    o = OracleBatch(**mesh, scheme="CN")
    for b in (g, o):
        b.set_current_stimulation(**stim)
        b.enable_spike_detection(0.0)
    t = 0.0
    for k in (1, 300, 699):
        for b in (g, o):
            b.advance(t, 1e-3, k)
        t += k * 1e-3
    if thomas == "serial":
        np.testing.assert_array_equal(g.get_states(), o.get_states())
    else:
        err = _rel(g.vm(), o.vm())
        print(f"{kernels} partitioned: Vm rel L2 err {err:.3e}")
        assert err <= 1e-10
    for a, b in zip(g.spikes(), o.spikes()):
        np.testing.assert_array_equal(a, b)
    first = o.spikes()[0]
    off = np.concatenate([[0], np.cumsum(mesh["n_nodes"])])
    fired = [f for f in range(len(LENGTHS)) if (first[off[f]:off[f + 1]] >= 0).any()]
    assert len(fired) >= len(LENGTHS) - 4, f"most fibres must fire: {fired}"


@pytest.mark.parametrize("threads", [64, 256])
def test_edge_lengths_large_model_per_fibre_reaction(threads):
    from opendihu_amd.build import LIB_DIR
    from opendihu_amd.fibers import FiberBatch
    lib = f"libfms_{SHORTEN}.so"
    if not os.path.exists(os.path.join(LIB_DIR, lib)) or \
            not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}.so")):
        pytest.skip("the 56-state library or oracle model is not built")
    lengths = [2, 3, 31, 33, 63, 65, 97, 129, 1481]
    mesh, _ = _mesh(lengths)
    n = mesh["n_nodes"]
    F = n.size
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    table = np.ones((4, 2), np.uint8)
    stim = dict(stim_begin=(n // 2).astype(np.int32), stim_end=np.minimum(n // 2 + 3, n).astype(np.int32),
                firing_events=table, motor_unit=np.arange(F, dtype=np.int32) % 2, call_frequency=1.0,
                repeat_after_first_call=0.1, call_enable_begin=0.0)
    g = FiberBatch(**mesh, scheme="IE", kernels="split", thomas="serial", library=lib)
    g.set_tuning(reaction_per_fibre=1, reaction_threads=threads, pair_halves=1)
    o = OracleBatch(**mesh, scheme="IE", model=SHORTEN, threads=8)
    assert np.any(off % 32 != 0)
    for b in (g, o):
        b.set_firing_table_stimulation(**stim)
        b.enable_spike_detection(0.0)
        b.advance(0.0, 2e-4, 400)
    for k in range(g.n_states):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k}")
    for a, b in zip(g.spikes(), o.spikes()):
        np.testing.assert_array_equal(a, b)
