"""Generated CellML models on the GPU (through the C ABI) against the oracle.

* the Hodgkin-Huxley library generated from the reference's model file must be
  bit-identical to the hand-written HH library (same expression order);
* the 56-state Shorten library must be bit-identical to the oracle's Shorten
  model with the serial Thomas solve, and within 1e-10 (relative L2) with the
  partitioned solve.
The generated libraries are built from the reference's CellML inputs where that
checkout exists (build()); the tests skip when the library was not built.
"""
import os

import numpy as np
import pytest

from cases import c1, c2, uneven
from oracle_binding import OracleBatch, ORACLE_BUILD

pytestmark = pytest.mark.gpu

SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"


def _lib_or_skip(name):
    from opendihu_amd.build import LIB_DIR
    if not os.path.exists(os.path.join(LIB_DIR, name)):
        pytest.skip(f"{name} not built (needs the reference's CellML input at build time)")
    return name


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_generated_hh_equals_handwritten():
    from opendihu_amd.fibers import FiberBatch
    lib = _lib_or_skip("libfms_gen_hodgkin_huxley_1952.so")
    mesh, stim = uneven(8, lo=300, hi=900, seed=4)
    runs = []
    for name in ("libfms_hodgkin_huxley_1952.so", lib):
        b = FiberBatch(**mesh, scheme="CN", kernels="fused", library=name)
        b.set_current_stimulation(**stim)
        b.advance(0.0, 1e-3, 2500)
        runs.append(b.get_states())
    np.testing.assert_array_equal(runs[0], runs[1])


def _shorten_pair(kernels, thomas, F=6, n=301, steps=1500, precision="fp64", dt=2e-4):
    from opendihu_amd.fibers import FiberBatch
    lib = _lib_or_skip(f"libfms_{SHORTEN}.so")
    if not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}.so")):
        pytest.skip("oracle Shorten model not built")
    mesh, stim = c2(n_fibers=F, n_nodes=n, length_cm=(n - 1) * 0.01, dt=dt, end_time=steps * dt, rate_hz=4000.0)
    g = FiberBatch(**mesh, scheme="IE", kernels=kernels, thomas=thomas, precision=precision, library=lib)
    o = OracleBatch(**mesh, scheme="IE", model=SHORTEN, threads=8)
    for b in (g, o):
        b.set_firing_table_stimulation(**stim)
        b.enable_spike_detection(0.0)
        b.advance(0.0, dt, steps)
    return g, o


@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_shorten_bit_exact(kernels):
    g, o = _shorten_pair(kernels, "serial")
    for k in range(g.n_states):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k}")
    assert (o.spikes()[1] > 0).sum() > 0, "stimulus must trigger action potentials"


def test_shorten_partitioned():
    g, o = _shorten_pair("fused", "partitioned")
    err = _rel(g.vm(), o.vm())
    print(f"shorten fused partitioned: Vm rel L2 err {err:.3e}")
    assert err <= 1e-10
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])


def test_shorten_mixed_precision():
    g, o = _shorten_pair("fused", "partitioned", precision="mixed")
    gf, of = g.spikes()[0], o.spikes()[0]
    a = np.where(gf > 0, gf, 10**9)
    b = np.where(of > 0, of, 10**9)
    fired = (gf > 0) | (of > 0)
    shift = int(np.max(np.abs(a[fired] - b[fired]))) if fired.any() else 0
    print(f"shorten mixed: spike-time error {shift} steps, Vm max abs err {np.max(np.abs(g.vm() - o.vm())):.3e} mV")
    assert shift <= 1
