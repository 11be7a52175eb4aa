"""Parity at the benchmark sizes. Fibres are independent within a split step, so
a whole batch on the GPU is checked fibre by fibre against the oracle run on a
few sampled fibres only (with the fibre-local stimulation ranges and motor
units they have in the full batch): the Shorten batch of 1024 fibres x 1481
nodes, the same batch with the 57-state TK2014 model, the 4096 uneven HH fibres
and the 16384-fibre mixed-precision batch.
"""
import numpy as np
import pytest

from cases import c2, c5, uneven
from oracle_binding import OracleBatch

pytestmark = pytest.mark.gpu

SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"
PER_FIBRE = ("stim_begin", "stim_end", "motor_unit", "onset", "duration", "amplitude")


def _subset(mesh, stim, fibres):
    n = mesh["n_nodes"]
    eoff = np.concatenate([[0], np.cumsum(n - 1)])
    sub_mesh = dict(n_nodes=n[fibres].copy(),
                    element_lengths=np.concatenate([mesh["element_lengths"][eoff[f]:eoff[f + 1]] for f in fibres]),
                    prefactor=np.asarray(mesh["prefactor"])[fibres].copy())
    sub_stim = {k: (np.asarray(v)[fibres].copy() if k in PER_FIBRE else v) for k, v in stim.items()}
    return sub_mesh, sub_stim


def _slices(mesh, fibres):
    off = np.concatenate([[0], np.cumsum(mesh["n_nodes"])])
    return [slice(off[f], off[f + 1]) for f in fibres]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("thomas", ["serial", "partitioned"])
def test_shorten_1024x1481_sampled_fibres(thomas):
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import SHORTEN_LIBRARY
    dt, steps = 2.5e-5, 300
    mesh, stim = c2(rate_hz=4000.0)   # the benchmark's batch, motor units firing often
    fibres = [0, 389, 1023]
    g = FiberBatch(**mesh, scheme="IE", kernels="split", thomas=thomas, library=SHORTEN_LIBRARY)
    g.set_firing_table_stimulation(**stim)
    g.enable_spike_detection(-40.0)
    g.advance(0.0, dt, steps)
    vm, gate0, first = g.vm(), g.get_state(1), g.spikes()[0]
    g.close()
    sm, ss = _subset(mesh, stim, fibres)
    o = OracleBatch(**sm, scheme="IE", model=SHORTEN, threads=8)
    o.set_firing_table_stimulation(**ss)
    o.enable_spike_detection(-40.0)
    o.advance(0.0, dt, steps)
    ovm, og0, ofirst = o.vm(), o.get_state(1), o.spikes()[0]
    osl = _slices(sm, range(len(fibres)))
    for s, so in zip(_slices(mesh, fibres), osl):
        if thomas == "serial":
            np.testing.assert_array_equal(vm[s], ovm[so])
            np.testing.assert_array_equal(gate0[s], og0[so])
        else:
            assert _rel(vm[s], ovm[so]) <= 1e-10
        np.testing.assert_array_equal(first[s], ofirst[so])
    assert (ofirst >= 0).sum() > 0, "the sampled fibres must fire"


def test_tk2014_1024x1481_sampled_fibres():
    import os
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import LIB_DIR, TK2014_LIBRARY
    from oracle_binding import ORACLE_BUILD
    if not os.path.exists(os.path.join(LIB_DIR, TK2014_LIBRARY)) or \
            not os.path.exists(os.path.join(ORACLE_BUILD, "libfmso_tk2014.so")):
        pytest.skip("TK2014 library or oracle model not built (needs the reference checkout at build time)")
    dt, steps = 2.5e-5, 300
    mesh, stim = c2(rate_hz=4000.0)
    fibres = [0, 511, 1023]
    g = FiberBatch(**mesh, scheme="IE", kernels="split", thomas="serial", library=TK2014_LIBRARY)
    g.set_firing_table_stimulation(**stim)
    g.enable_spike_detection(-40.0)
    g.advance(0.0, dt, steps)
    vm, gate, first = g.vm(), g.get_state(18), g.spikes()[0]
    g.close()
    sm, ss = _subset(mesh, stim, fibres)
    o = OracleBatch(**sm, scheme="IE", model="tk2014", threads=8)
    o.set_firing_table_stimulation(**ss)
    o.enable_spike_detection(-40.0)
    o.advance(0.0, dt, steps)
    ovm, og, ofirst = o.vm(), o.get_state(18), o.spikes()[0]
    for s, so in zip(_slices(mesh, fibres), _slices(sm, range(len(fibres)))):
        np.testing.assert_array_equal(vm[s], ovm[so])
        np.testing.assert_array_equal(gate[s], og[so])
        np.testing.assert_array_equal(first[s], ofirst[so])
    assert (ofirst >= 0).sum() > 0, "the sampled fibres must fire"


def test_hh_4096_uneven_sampled_fibres():
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import HH_LIBRARY
    dt, steps = 1e-3, 2500
    mesh, stim = uneven(4096, seed=3)
    fibres = [0, 17, 2048, 4095]
    g = FiberBatch(**mesh, scheme="CN", kernels="fused", thomas="partitioned", library=HH_LIBRARY)
    g.set_current_stimulation(**stim)
    g.enable_spike_detection(0.0)
    g.advance(0.0, dt, steps)
    vm, first = g.vm(), g.spikes()[0]
    g.close()
    sm, ss = _subset(mesh, stim, fibres)
    o = OracleBatch(**sm, scheme="CN", threads=8)
    o.set_current_stimulation(**ss)
    o.enable_spike_detection(0.0)
    o.advance(0.0, dt, steps)
    ovm, ofirst = o.vm(), o.spikes()[0]
    for s, so in zip(_slices(mesh, fibres), _slices(sm, range(len(fibres)))):
        assert _rel(vm[s], ovm[so]) <= 1e-10
        np.testing.assert_array_equal(first[s], ofirst[so])
    assert (ofirst >= 0).sum() > 100, "the sampled fibres must conduct action potentials"


def test_mixed_16384_sampled_fibres_vs_fp64():
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import SHORTEN_LIBRARY
    dt, steps = 2.5e-5, 200
    mesh, stim = c5()
    fibres = [0, 5000, 16383]
    g = FiberBatch(**mesh, scheme="IE", kernels="split", thomas="partitioned", precision="mixed",
                   library=SHORTEN_LIBRARY)
    g.set_firing_table_stimulation(**stim)
    g.enable_spike_detection(-40.0)
    g.advance(0.0, dt, steps)
    vm, first = g.vm(), g.spikes()[0]
    g.close()
    sm, ss = _subset(mesh, stim, fibres)
    o = OracleBatch(**sm, scheme="IE", model=SHORTEN, threads=8)
    o.set_firing_table_stimulation(**ss)
    o.enable_spike_detection(-40.0)
    o.advance(0.0, dt, steps)
    ovm, ofirst = o.vm(), o.spikes()[0]
    for s, so in zip(_slices(mesh, fibres), _slices(sm, range(len(fibres)))):
        err = float(np.max(np.abs(vm[s] - ovm[so])))
        print(f"mixed vs FP64 oracle: Vm max abs err {err:.3e} mV")
        assert err <= 1e-2
        a, b = first[s], ofirst[so]
        assert np.array_equal(a >= 0, b >= 0)
        both = (a >= 0) & (b >= 0)
        assert np.all(np.abs(a[both] - b[both]) <= 1)
    assert (ofirst >= 0).sum() > 0, "the sampled fibres must fire"


def test_shorten_1024x1481_bench_schedule_1ms():
    """1 ms (40000 split steps) of the 1024 x 1481 Shorten batch under the
    benchmark's schedule (split kernels, per-fibre reaction with 64-thread
    workgroups, paired reaction halves, graphs of 8 steps, partitioned solve),
    every motor unit firing at t = 0 and at Poisson times after it: sampled
    fibres against the oracle, Vm within 1e-10 (relative L2) and spike steps equal"""
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import SHORTEN_LIBRARY
    import cases
    dt, steps = 2.5e-5, 40000
    mesh, stim = c2(rate_hz=400.0)
    stim["firing_events"] = cases.poisson_firing_table(200000, 10, 0.4, dt, 0.1, 2, first_at_zero=True)
    fibres = [0, 389, 700, 1023]
    g = FiberBatch(**mesh, scheme="IE", kernels="split", thomas="partitioned", library=SHORTEN_LIBRARY)
    g.set_tuning(reaction_per_fibre=1, reaction_threads=64, pair_halves=1, graph_steps=8)
    g.set_firing_table_stimulation(**stim)
    g.enable_spike_detection(-40.0)
    for k in range(4):
        g.advance(k * 10000 * dt, dt, 10000)
    vm, first, count = g.vm(), g.spikes()[0], g.spikes()[1]
    replays = g.graph_counts()["replays"]
    g.close()
    assert replays == 4 * (9999 // 8)
    sm, ss = _subset(mesh, stim, fibres)
    o = OracleBatch(**sm, scheme="IE", model=SHORTEN, threads=len(fibres))
    o.set_firing_table_stimulation(**ss)
    o.enable_spike_detection(-40.0)
    for k in range(4):
        o.advance(k * 10000 * dt, dt, 10000)
    ovm, ofirst, ocount = o.vm(), o.spikes()[0], o.spikes()[1]
    for s, so in zip(_slices(mesh, fibres), _slices(sm, range(len(fibres)))):
        err = _rel(vm[s], ovm[so])
        print(f"1 ms, bench schedule: Vm rel L2 {err:.3e}, max abs {np.max(np.abs(vm[s] - ovm[so])):.3e} mV")
        assert err <= 1e-10
        np.testing.assert_array_equal(first[s], ofirst[so])
        np.testing.assert_array_equal(count[s], ocount[so])
    assert (ofirst >= 0).sum() > 100, "action potentials must propagate along the sampled fibres"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hh_4096_uneven_full_10ms(precision):
    """the whole 10 ms (10000 steps) of the 4096 uneven HH fibres under the
    fused schedule: FP64 within 1e-10 of the oracle with equal spike steps; the
    mixed form (FP32 gates) with spike-time error <= 1 step against the FP64
    oracle"""
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import HH_LIBRARY
    dt, steps = 1e-3, 10000
    mesh, stim = uneven(4096, seed=3)
    fibres = [0, 17, 2048, 3001, 4095]
    g = FiberBatch(**mesh, scheme="CN", kernels="fused", thomas="partitioned", precision=precision,
                   library=HH_LIBRARY)
    g.set_current_stimulation(**stim)
    g.enable_spike_detection(0.0)
    g.advance(0.0, dt, steps)
    vm, first = g.vm(), g.spikes()[0]
    g.close()
    sm, ss = _subset(mesh, stim, fibres)
    o = OracleBatch(**sm, scheme="CN", threads=len(fibres))
    o.set_current_stimulation(**ss)
    o.enable_spike_detection(0.0)
    o.advance(0.0, dt, steps)
    ovm, ofirst = o.vm(), o.spikes()[0]
    for s, so in zip(_slices(mesh, fibres), _slices(sm, range(len(fibres)))):
        a, b = first[s], ofirst[so]
        if precision == "fp64":
            assert _rel(vm[s], ovm[so]) <= 1e-10
            np.testing.assert_array_equal(a, b)
        else:
            print(f"mixed, 10 ms: Vm max abs err {np.max(np.abs(vm[s] - ovm[so])):.3e} mV")
            assert np.array_equal(a >= 0, b >= 0)
            both = (a >= 0) & (b >= 0)
            assert np.max(np.abs(a[both] - b[both])) <= 1
    assert (ofirst >= 0).sum() > 1000, "the sampled fibres must conduct action potentials"


def test_adjoint_256x1000_full_size():
    """the adjoint at its benchmark size: HH, 256 fibres x 1000 nodes, 5000 split
    steps (5 ms), 32 sensors per fibre, checkpoint every 100 steps; the
    directional derivative of the trace misfit against central differences
    (h = 1e-6) within 1e-5"""
    import torch
    from cases import AM, CM
    from opendihu_amd.adjoint import AdjointProblem
    from opendihu_amd.fibers import FiberBatch
    F, n, K, dt = 256, 1000, 5000, 1e-3
    mesh, stim = uneven(F, lo=n, hi=n, seed=41)
    b = FiberBatch(**mesh, scheme="CN", kernels="fused", thomas="partitioned")
    b.set_current_stimulation(**stim)
    steps = np.stack([np.arange(K) * dt, np.full(K, dt), np.full(K, dt / 2), np.full(K, dt)], 1)
    sensors = np.tile(np.round(np.linspace(0, n - 1, 32)).astype(np.int32), (F, 1))
    prob = AdjointProblem(b, steps, sensors, checkpoint_interval=100, am=AM, cm=CM)
    rng = np.random.default_rng(42)
    sigma = mesh["prefactor"] * AM * CM
    theta_true = np.stack([rng.uniform(100, 140, F), rng.uniform(30, 42, F), rng.uniform(0.2, 0.4, F), sigma], 1)
    theta0 = theta_true * rng.uniform(0.95, 1.05, theta_true.shape)
    dev = torch.device("cuda")
    target = prob.forward_tensor(torch.tensor(theta_true, device=dev))
    tr = prob.forward_tensor(torch.tensor(theta0, device=dev))
    grad = prob.backward_tensor(tr - target).cpu().numpy()

    def loss(th):
        return 0.5 * float(((prob.forward_tensor(torch.tensor(th, device=dev)) - target) ** 2).sum())
    v = np.random.default_rng(43).standard_normal(theta0.shape) * theta0
    h = 1e-6
    fd = (loss(theta0 + h * v) - loss(theta0 - h * v)) / (2 * h)
    ad = float(np.sum(grad * v))
    print(f"C4 full size: adjoint {ad:.12e} central FD {fd:.12e} rel {abs(ad - fd) / abs(fd):.2e}")
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    b.close()
