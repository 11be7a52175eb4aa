"""Generated CellML models on the GPU (through the C ABI) against the oracle.

* the Hodgkin-Huxley library generated from the reference's model file must be
  bit-identical to the hand-written HH library (same expression order);
* the 56-state Shorten library must be bit-identical to the oracle's Shorten
  model with the serial Thomas solve, and within 1e-10 (relative L2) with the
  partitioned solve;
* the 57-state TK2014 library (converted from the reference's hand-written Vc
  kernel) must be bit-identical to the oracle's TK2014 model, which takes the
  kernel's lines by another route.
The generated libraries are built from the reference's CellML inputs where that
checkout exists (build()); the tests skip when the library was not built.
"""
import os

import numpy as np
import pytest

from cases import c2, uneven
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


def _shorten_pair(kernels, thomas, F=6, n=301, steps=1500, precision="fp64", dt=2e-4, model=SHORTEN):
    from opendihu_amd.fibers import FiberBatch
    lib = _lib_or_skip(f"libfms_{model}.so")
    if not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{model}.so")):
        pytest.skip(f"oracle model {model} not built")
    mesh, stim = c2(n_fibers=F, n_nodes=n, length_cm=(n - 1) * 0.01, dt=dt, end_time=steps * dt, rate_hz=4000.0)
    per_node = kernels.startswith("split-per-node")
    unpaired = kernels.endswith("-unpaired")
    threads = 64 if "-64" in kernels else 256
    kernels = "split" if kernels.startswith("split") else kernels
    g = FiberBatch(**mesh, scheme="IE", kernels=kernels, thomas=thomas, precision=precision, library=lib)
    g.set_tuning(reaction_per_fibre=0 if per_node else 1, pair_halves=0 if unpaired else 1, reaction_threads=threads)
    o = OracleBatch(**mesh, scheme="IE", model=model, threads=8)
    for b in (g, o):
        b.set_firing_table_stimulation(**stim)
        b.enable_spike_detection(0.0)
        b.advance(0.0, dt, steps)
    return g, o


@pytest.mark.parametrize("threads", [64, 256])
def test_shorten_only_if_stimulated_idle_fibres(threads):
    """onlyComputeIfHasBeenStimulated with the per-fibre reaction kernel: the
    workgroup of a fibre none of whose Heun steps computes leaves its states
    alone (the spike detector still reads Vm); half the motor units never fire,
    so their fibres stay at their initial states. Bit-exact vs the oracle."""
    from opendihu_amd.fibers import FiberBatch
    lib = _lib_or_skip(f"libfms_{SHORTEN}.so")
    if not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}.so")):
        pytest.skip("oracle model not built")
    F, n, steps, dt = 6, 301, 600, 2e-4
    mesh, stim = c2(n_fibers=F, n_nodes=n, length_cm=(n - 1) * 0.01, dt=dt, end_time=steps * dt, rate_hz=4000.0)
    mu = stim["motor_unit"]
    silent = sorted(set(mu.tolist()))[::2]
    for m in silent:
        stim["firing_events"][:, m % stim["firing_events"].shape[1]] = 0
    g = FiberBatch(**mesh, scheme="IE", kernels="split", thomas="serial", library=lib)
    g.set_tuning(reaction_per_fibre=1, reaction_threads=threads, pair_halves=1)
    o = OracleBatch(**mesh, scheme="IE", model=SHORTEN, threads=8)
    for b in (g, o):
        b.set_firing_table_stimulation(**stim)
        b.set_only_compute_if_stimulated(True)
        b.enable_spike_detection(0.0)
        b.advance(0.0, dt, steps)
    for k in range(g.n_states):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k}")
    for a, b in zip(g.spikes(), o.spikes()):
        np.testing.assert_array_equal(a, b)
    idle = [f for f in range(F) if mu[f] in silent]
    busy = [f for f in range(F) if mu[f] not in silent]
    assert idle and busy, (mu, silent)
    y0 = o.get_state(2)
    for f in idle:
        assert np.all(y0[f * n:(f + 1) * n] == y0[idle[0] * n]), "idle fibres keep their initial states"
    assert (o.spikes()[1] > 0).sum() > 0


@pytest.mark.parametrize("kernels", ["split", "split-unpaired", "split-64", "split-64-unpaired", "split-per-node",
                                     "split-per-node-unpaired", "fused"])
def test_shorten_bit_exact(kernels):
    g, o = _shorten_pair(kernels, "serial")
    for k in range(g.n_states):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k}")
    assert (o.spikes()[1] > 0).sum() > 0, "stimulus must trigger action potentials"
    for a, b in zip(g.spikes(), o.spikes()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kernels", ["split", "split-64-unpaired", "split-per-node", "fused"])
def test_tk2014_bit_exact(kernels):
    g, o = _shorten_pair(kernels, "serial", F=4, n=201, steps=1200, dt=1e-4, model="tk2014")
    for k in range(g.n_states):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k}")
    assert (o.spikes()[1] > 0).sum() > 0, "stimulus must trigger action potentials"
    for a, b in zip(g.spikes(), o.spikes()):
        np.testing.assert_array_equal(a, b)


def test_tk2014_partitioned():
    g, o = _shorten_pair("split", "partitioned", F=4, n=201, steps=1200, dt=1e-4, model="tk2014")
    err = _rel(g.vm(), o.vm())
    print(f"tk2014 split partitioned: Vm rel L2 err {err:.3e}")
    assert err <= 1e-10
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])


def test_shorten_partitioned():
    g, o = _shorten_pair("fused", "partitioned")
    err = _rel(g.vm(), o.vm())
    print(f"shorten fused partitioned: Vm rel L2 err {err:.3e}")
    assert err <= 1e-10
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])


def test_tk2014_mixed_precision():
    g, o = _shorten_pair("fused", "partitioned", F=4, n=201, steps=1200, dt=1e-4, precision="mixed", model="tk2014")
    gf, of = g.spikes()[0], o.spikes()[0]
    fired = (gf > 0) | (of > 0)
    assert fired.any(), "stimulus must trigger action potentials"
    a = np.where(gf > 0, gf, 10**9)
    b = np.where(of > 0, of, 10**9)
    shift = int(np.max(np.abs(a[fired] - b[fired])))
    print(f"tk2014 mixed: spike-time error {shift} steps, Vm max abs err {np.max(np.abs(g.vm() - o.vm())):.3e} mV")
    assert shift <= 1


def test_shorten_mixed_precision():
    g, o = _shorten_pair("fused", "partitioned", precision="mixed")
    gf, of = g.spikes()[0], o.spikes()[0]
    a = np.where(gf > 0, gf, 10**9)
    b = np.where(of > 0, of, 10**9)
    fired = (gf > 0) | (of > 0)
    shift = int(np.max(np.abs(a[fired] - b[fired]))) if fired.any() else 0
    print(f"shorten mixed: spike-time error {shift} steps, Vm max abs err {np.max(np.abs(g.vm() - o.vm())):.3e} mV")
    assert shift <= 1


def _random_points(lib_name, oracle_model, n=4000, seed=0):
    from opendihu_amd import _lib as L
    from oracle_binding import oracle_library, P, _dp
    g = L.load_library(lib_name)
    o = oracle_library(oracle_model)
    ns, na, npar, _ = L.model_info(g)
    y0 = np.zeros(ns)
    o.fmso_initial_states(P(y0, _dp))
    p0 = np.zeros(max(npar, 1))
    g.fms_model_defaults(None, P(p0, _dp))
    rng = np.random.default_rng(seed)
    Y = y0[:, None] * (1 + 0.02 * rng.standard_normal((ns, n)))
    Y[0] = rng.uniform(-90, 40, n)
    Pm = np.repeat(p0[:npar, None], n, axis=1) if npar else np.zeros((0, n))
    return L, g, o, P, _dp, Y, Pm


@pytest.mark.parametrize("which", ["hh", SHORTEN, "tk2014"])
@pytest.mark.parametrize("exp_mode", ["approx", "libm"])
def test_device_rhs_and_heun_match_oracle(which, exp_mode):
    """the compiled model evaluated on the device at random states: rates and ten
    Heun steps bit-identical to the oracle with the approximate exponential; with
    the math library exp/log (device libm vs glibc) within 1e-10 relative"""
    lib = "libfms_hodgkin_huxley_1952.so" if which == "hh" else _lib_or_skip(f"libfms_{which}.so")
    if which != "hh" and not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{which}.so")):
        pytest.skip(f"oracle model {which} not built")
    L, g, o, P, _dp, Y, Pm = _random_points(lib, which)
    mode = 0 if exp_mode == "approx" else 1
    n = Y.shape[1]
    r = L.eval_rhs(g, Y, Pm, 0.0, exp_mode)
    yh = L.eval_heun(g, Y, Pm, 0.0, 2e-4, 10, exp_mode)
    ro, yo = np.empty_like(r), np.empty_like(yh)
    rr = np.zeros(Y.shape[0])
    aa = np.zeros(256)
    for j in range(n):
        yy = np.ascontiguousarray(Y[:, j])
        pp = np.ascontiguousarray(Pm[:, j]) if Pm.shape[0] else np.zeros(1)
        o.fmso_rhs(P(yy, _dp), P(pp, _dp), 0.0, P(rr, _dp), P(aa, _dp), mode)
        ro[:, j] = rr
        o.fmso_heun_cell(P(yy, _dp), P(pp, _dp), 0.0, 2e-4, 10, mode)
        yo[:, j] = yy
    if exp_mode == "approx":
        np.testing.assert_array_equal(r, ro)
        np.testing.assert_array_equal(yh, yo)
    else:
        np.testing.assert_allclose(r, ro, rtol=1e-10, atol=1e-300)
        np.testing.assert_allclose(yh, yo, rtol=1e-10, atol=1e-12)
    # repeated device evaluation gives the same bits
    np.testing.assert_array_equal(L.eval_rhs(g, Y, Pm, 0.0, exp_mode), r)


@pytest.mark.parametrize("model,kernels", [("hh", "split"), ("hh", "fused"), (SHORTEN, "split"), (SHORTEN, "fused"),
                                           ("tk2014", "split")])
def test_algebraics_for_transfer_match_oracle(model, kernels):
    """algebraicsForTransfer: the algebraics of the last Heun step's second
    evaluation (GEN/03_generator_vc.cpp:1131-1162), bit-identical to the
    oracle's; an algebraic that is a parameter (parametersUsedAsAlgebraic, the
    Shorten model's I_Stim = algebraic 32) is not stored (:1143-1146), so its
    transfer values keep their initial zeros"""
    from opendihu_amd.fibers import FiberBatch
    if model == "hh":
        mesh, stim = uneven(6, lo=300, hi=600, seed=12)
        g = FiberBatch(**mesh, scheme="IE", thomas="serial", kernels=kernels)
        o = OracleBatch(**mesh, scheme="IE")
        idx = [0, 3, 4, 8]
        for b in (g, o):
            b.set_current_stimulation(**stim)
        dt, steps, calls = 1e-3, 350, 2
    else:
        lib = _lib_or_skip(f"libfms_{model}.so")
        mesh, stim = c2(n_fibers=4, n_nodes=301, length_cm=3.0, dt=2e-4, end_time=0.2, rate_hz=4000.0)
        g = FiberBatch(**mesh, scheme="IE", thomas="serial", kernels=kernels, library=lib)
        o = OracleBatch(**mesh, scheme="IE", model=model, threads=4)
        idx = [0, 5, 32, 70] if model == SHORTEN else [0, 9, 44, 70]
        for b in (g, o):
            b.set_firing_table_stimulation(**stim)
        dt, steps, calls = 2e-4, 250, 2
    for b in (g, o):
        b.set_algebraics_for_transfer(idx)
    for c in range(calls):
        for b in (g, o):
            b.advance(c * steps * dt, dt, steps)
        ga, oa = g.get_algebraics(), o.algebraics()
        for q, k in enumerate(idx):
            np.testing.assert_array_equal(ga[q], oa[q], err_msg=f"{model} algebraic {k} after call {c}")
    if model == SHORTEN:
        assert np.all(ga[2] == 0.0), "a parameter algebraic is not stored"
    assert np.all(np.isfinite(ga)) and np.count_nonzero(ga) > 0
    for k in range(g.n_states):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k}")
