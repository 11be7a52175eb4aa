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
