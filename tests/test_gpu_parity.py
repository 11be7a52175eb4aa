"""Parity of the HIP path (through the C ABI) with the CPU oracle, on MI355X.

Bars (FP64):
  * thomas="serial": bit-exact in every state (same operations in the same order)
  * thomas="partitioned": relative L2 error ||dVm||_2 / ||Vm||_2 <= 1e-10 after N
    split steps (the max-norm error is printed beside it: in the action-potential
    upstroke dV/dt ~ 400 mV/ms turns a 1e-16 rounding difference of the solver
    into ~1e-8 mV at single nodes)
  * spike times (first upward crossing of 0 mV, per node) equal at step resolution
Mixed precision (FP32 gates, FP64 Vm/Thomas) is compared with the FP64 oracle
with the error stated in the assertion (spike-time error <= 1 step).
"""
import numpy as np
import pytest

from cases import c1, uneven, vc_lane_stim_range, firing_table
from oracle_binding import OracleBatch

pytestmark = pytest.mark.gpu

REL_TOL = 1e-10


def _gpu_batch(mesh, **kw):
    from opendihu_amd.fibers import FiberBatch
    return FiberBatch(**mesh, **kw)


def _rel_err(a, b):
    """relative L2 error of a against b"""
    return float(np.linalg.norm(a - b) / max(1e-300, np.linalg.norm(b)))


def _max_rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def _run_pair(mesh, stim, n_steps, dt=1e-3, scheme="CN", gpu_kw=None, oracle_exp="approx", spikes=True,
              params=None):
    gpu_kw = dict(gpu_kw or {})
    gpu_kw.setdefault("exp_mode", oracle_exp)
    g = _gpu_batch(mesh, scheme=scheme, parameters=params, **gpu_kw)
    o = OracleBatch(**mesh, scheme=scheme, exp_mode=oracle_exp, parameters=params, threads=8)
    for b in (g, o):
        b.set_current_stimulation(**stim)
        if spikes:
            b.enable_spike_detection(0.0)
        b.advance(0.0, dt, n_steps)
    return g, o


@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_c1_serial_thomas_bit_exact(kernels):
    mesh, stim = c1()
    g, o = _run_pair(mesh, stim, 3000, gpu_kw=dict(thomas="serial", kernels=kernels))
    for k in range(4):
        np.testing.assert_array_equal(g.get_state(k), o.get_state(k), err_msg=f"state {k} ({kernels})")
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])


@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_c1_partitioned_thomas(kernels):
    mesh, stim = c1()
    g, o = _run_pair(mesh, stim, 10000, gpu_kw=dict(thomas="partitioned", kernels=kernels))
    err = _rel_err(g.vm(), o.vm())
    print(f"c1 {kernels}: Vm rel L2 err {err:.3e}, max-norm {_max_rel_err(g.vm(), o.vm()):.3e}")
    assert err <= REL_TOL, f"Vm rel err {err:.3e} ({kernels})"
    gf, gc = g.spikes()
    of, oc = o.spikes()
    assert (of > 0).sum() > 300, "the action potential must propagate"
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gc, oc)


def test_c1_libm_exp():
    mesh, stim = c1()
    g, o = _run_pair(mesh, stim, 4000, oracle_exp="libm", gpu_kw=dict(thomas="serial", kernels="fused"))
    err = _rel_err(g.vm(), o.vm())
    assert err <= REL_TOL, f"Vm rel err {err:.3e} with device exp vs libm exp"
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])


@pytest.mark.parametrize("kernels,scheme", [("fused", "CN"), ("split", "IE"), ("fused", "IE")])
def test_uneven_batch(kernels, scheme):
    mesh, stim = uneven(48, seed=11)
    g, o = _run_pair(mesh, stim, 2500, scheme=scheme, gpu_kw=dict(kernels=kernels))
    err = _rel_err(g.vm(), o.vm())
    print(f"uneven {kernels} {scheme}: Vm rel L2 err {err:.3e}, max-norm {_max_rel_err(g.vm(), o.vm()):.3e}")
    assert err <= REL_TOL, f"Vm rel err {err:.3e}"
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])
    for k in range(1, 4):
        assert _rel_err(g.get_state(k), o.get_state(k)) <= REL_TOL


def test_per_fibre_parameters():
    mesh, stim = uneven(16, lo=300, hi=700, seed=5)
    rng = np.random.default_rng(6)
    params = np.stack([np.zeros(16), rng.uniform(100, 140, 16), rng.uniform(30, 42, 16), rng.uniform(0.2, 0.4, 16)], 1)
    g, o = _run_pair(mesh, stim, 2000, scheme="CN", gpu_kw=dict(kernels="fused", thomas="serial"), params=params)
    np.testing.assert_array_equal(g.vm(), o.vm())


def _firing_setup(batch, mesh, width):
    n = mesh["n_nodes"]
    F = n.size
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    nmj = n // 2
    sb, se = vc_lane_stim_range(off, nmj, n, width)
    fe = firing_table(40, 10, 4.0, 0.1, seed=2)
    mu = (np.arange(F) * 7) % 10 + 1
    jit = [[0.1, -0.1, 0.05][: (f % 4)] for f in range(F)]
    batch.set_firing_table_stimulation(sb, se, fe, mu, call_frequency=1.0, repeat_after_first_call=0.1,
                                       call_enable_begin=0.2 * (np.arange(F) % 3), jitter_lists=jit,
                                       log_capacity=4096)


@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_firing_table_stimulation(kernels):
    mesh, _ = uneven(12, lo=200, hi=600, seed=9)
    g = _gpu_batch(mesh, scheme="IE", kernels=kernels, thomas="serial")
    o = OracleBatch(**mesh, scheme="IE")
    for b in (g, o):
        _firing_setup(b, mesh, width=4)
        b.enable_spike_detection(0.0)
        b.advance(0.0, 2e-3, 1500)
    np.testing.assert_array_equal(g.vm(), o.vm())
    np.testing.assert_array_equal(g.spikes()[1], o.spikes()[1])
    gt, gf = g.stimulation_log()
    ot, of = o.stimulation_log()
    assert len(ot) > 0
    ga = sorted(zip(gf.tolist(), gt.tolist()))
    oa = sorted(zip(of.tolist(), ot.tolist()))
    assert ga == oa


def spike_time_error(first_a, first_b, n_steps):
    """max |step difference| of first spikes over nodes that fired in either run;
    a node that did not fire counts as firing after the last step"""
    a = np.where(first_a > 0, first_a, n_steps + 1)
    b = np.where(first_b > 0, first_b, n_steps + 1)
    fired = (first_a > 0) | (first_b > 0)
    return int(np.max(np.abs(a[fired] - b[fired]))) if fired.any() else 0


@pytest.mark.parametrize("exp_mode", ["approx", "libm"])
def test_mixed_precision_error_vs_fp64(exp_mode):
    mesh, stim = uneven(32, seed=21)
    n_steps = 4000
    g, o = _run_pair(mesh, stim, n_steps, scheme="CN", oracle_exp=exp_mode,
                     gpu_kw=dict(kernels="fused", precision="mixed"))
    vg, vo = g.vm(), o.vm()
    abs_err = float(np.max(np.abs(vg - vo)))
    shift = spike_time_error(g.spikes()[0], o.spikes()[0], n_steps)
    print(f"mixed {exp_mode}: spike-time error {shift} steps, Vm max abs err {abs_err:.3e} mV, "
          f"rel L2 {_rel_err(vg, vo):.3e}")
    assert shift <= 1, f"spike-time error {shift} steps (Vm max abs err {abs_err:.3e} mV)"


def test_schedule_matches_oracle():
    from opendihu_amd import _lib as L
    from oracle_binding import oracle_library, schedule
    lib = L.load_library()
    for args in [(10.0, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3), (5.0, 2.5e-5, 2.5e-5, 2.5e-5, 2.5e-5, 2.5e-5, 2.5e-5, 2.5e-5),
                 (100.0, 2e-3, 2e-3, 2e-3, 2e-4, 2e-4, 2e-3, 2e-3)]:
        a = L.schedule(lib, *args)
        b = schedule(oracle_library(), *args)
        np.testing.assert_array_equal(a, b)


def test_schedule_driven_run_bit_exact():
    from opendihu_amd import _lib as L
    mesh, stim = c1(200, 2.0)
    lib = L.load_library()
    sched = L.schedule(lib, 1.0, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3)
    steps = sched[:, [0, 1, 3, 5]]
    assert np.all(sched[:, 2] == 1) and np.all(sched[:, 4] == 1)
    g = _gpu_batch(mesh, scheme="CN", thomas="serial", kernels="fused")
    o = OracleBatch(**mesh, scheme="CN")
    for b in (g, o):
        b.set_current_stimulation(**stim)
        b.advance_schedule(steps)
    np.testing.assert_array_equal(g.vm(), o.vm())


def test_reference_abi_matches_native():
    """initializeArrays/computeMonodomain (the reference's dlopen boundary)
    against the oracle: states and algebraics for transfer after every phase;
    parameters and prefactor are taken from every call (gpu.tpp:481,490), a
    call whose instances of one fibre carry different parameters is refused
    and changes nothing"""
    import ctypes
    from opendihu_amd import _lib as L
    lib = L.load_library()
    F, n = 3, 301
    N = F * n
    d = L.FmsReferenceDims(n_fibers_to_compute=F, n_instances_per_fiber=n, n_parameters_per_instance=2,
                           firing_events_rows=2, firing_events_cols=100, frequency_jitter_cols=2,
                           n_algebraics_for_transfer=2, n_states_for_transfer=2, scheme=0, precision=0, exp_mode=0)
    L.check(lib, lib.fms_reference_set_dims(ctypes.byref(d)), "set dims")
    y0 = np.array([-75, 0.05, 0.6, 0.325])
    alg_idx = np.array([0, 6], np.int32)
    st_idx = np.array([0, 3], np.int32)
    fe = np.ones(200, np.uint8)
    jitter = np.array([0.1, -0.1] * F)
    mu = np.array([1, 2, 3], np.int32)
    nmj = np.array([n // 2] * F, np.int32)
    freq = np.full(F, 0.1)
    rep = np.full(F, 0.1)
    beg = np.array([0.0, 1.0, 2.0])
    last = beg - 1e-13 - 1.0 / freq
    dp, ip = L._dp, L._ip
    lib.initializeArrays(y0.ctypes.data_as(dp), alg_idx.ctypes.data_as(ip), st_idx.ctypes.data_as(ip),
                         fe.tobytes(), jitter.ctypes.data_as(dp), mu.ctypes.data_as(ip), nmj.ctypes.data_as(ip),
                         last.ctypes.data_as(dp), freq.ctypes.data_as(dp), rep.ctypes.data_as(dp),
                         beg.ctypes.data_as(dp))
    # parameters[p*N + instance]: i_Stim = 0 and g_Na = 120 on every node
    params = np.zeros(2 * N, np.float32)
    params[N:] = 120.0
    lengths = np.full(F * (n - 1), 0.01, np.float32)
    states_out = np.zeros(2 * N)
    alg_out = np.zeros(2 * N)
    fp = ctypes.POINTER(ctypes.c_float)
    o = OracleBatch(n_nodes=[n] * F, element_lengths=lengths.astype(np.float64), prefactor=0.007656, scheme="IE")
    o.set_firing_table_stimulation(nmj - 1, nmj + 2, fe.reshape(2, 100), mu, freq, rep, beg,
                                   jitter_lists=[[0.1, -0.1]] * F, last_check_time=last)
    o.set_algebraics_for_transfer(alg_idx)

    def run(k0, k1, pf):
        for k in range(k0, k1):
            lib.computeMonodomain(params.ctypes.data_as(fp), alg_out.ctypes.data_as(dp),
                                  states_out.ctypes.data_as(dp), lengths.ctypes.data_as(fp), k * 2e-3, 2e-3, 1, 1e-3,
                                  1, 2e-3, 1, pf, 20.0)
            assert L.last_error(lib) == "", L.last_error(lib)
            o.advance(k * 2e-3, 2e-3, 1, 1e-3, 1, 2e-3, 1)

    def check(phase):
        assert _rel_err(states_out[:N], o.vm()) <= REL_TOL, phase
        assert _rel_err(states_out[N:], o.get_state(3)) <= REL_TOL, phase
        alg = o.algebraics()
        for q in range(2):
            assert _rel_err(alg_out[q * N:(q + 1) * N], alg[q]) <= 1e-9, (phase, q)

    run(0, 600, 0.007656)
    check("phase 1")
    assert np.max(o.vm()) > 0.0, "the stimulated fibres fire"
    # a coupled solver changes the prefactor and fibre 1's parameters
    params[n:2 * n] = 2.5
    params[N + n:N + 2 * n] = 100.0
    o.pf[:] = 0.009
    o.params[1, 0], o.params[1, 1] = 2.5, 100.0
    run(600, 1000, 0.009)
    check("phase 2")
    # per-node parameters inside one fibre are refused; the states stay
    before = states_out.copy()
    params[N + 2 * n + 7] = 110.0
    lib.computeMonodomain(params.ctypes.data_as(fp), alg_out.ctypes.data_as(dp), states_out.ctypes.data_as(dp),
                          lengths.ctypes.data_as(fp), 1000 * 2e-3, 2e-3, 1, 1e-3, 1, 2e-3, 1, 0.009, 20.0)
    assert "differs between nodes of fibre 2" in L.last_error(lib)
    np.testing.assert_array_equal(states_out, before)
    params[N + 2 * n + 7] = 120.0
    run(1000, 1100, 0.009)
    check("phase 3")


@pytest.mark.parametrize("resident", [0, 1])
def test_fused_gate_residency_bit_exact(resident):
    """gates kept in registers or streamed from HBM by the fused kernel give the
    same bits as the split kernels with the serial solve"""
    mesh, stim = uneven(10, lo=300, hi=1500, seed=8)
    runs = []
    for kernels in ("split", "fused"):
        g = _gpu_batch(mesh, scheme="CN", thomas="serial", kernels=kernels)
        if kernels == "fused":
            g.set_tuning(steps_per_launch=97, resident_gates=resident)
        g.set_current_stimulation(**stim)
        g.advance(0.0, 1e-3, 600)
        runs.append(g.get_states())
    np.testing.assert_array_equal(runs[0], runs[1])


def test_algebraics_for_transfer_fused_equals_split():
    mesh, stim = uneven(6, lo=300, hi=600, seed=12)
    out = []
    for kernels in ("split", "fused"):
        g = _gpu_batch(mesh, scheme="IE", thomas="serial", kernels=kernels)
        g.set_current_stimulation(**stim)
        g.set_algebraics_for_transfer([0, 4, 8])
        g.advance(0.0, 1e-3, 700)
        out.append((g.get_states(), g.get_algebraics()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.all(np.isfinite(out[0][1])) and np.any(out[0][1] != 0)


@pytest.mark.parametrize("kernels", ["split", "fused"])
def test_only_compute_if_stimulated(kernels):
    """onlyComputeIfHasBeenStimulated: fibres keep their initial state until their
    first stimulation (firing table), then evolve; bit-exact vs the oracle"""
    mesh, _ = uneven(12, lo=200, hi=600, seed=9)
    g = _gpu_batch(mesh, scheme="IE", kernels=kernels, thomas="serial")
    o = OracleBatch(**mesh, scheme="IE")
    for b in (g, o):
        _firing_setup(b, mesh, width=4)
        b.set_only_compute_if_stimulated(True)
        b.advance(0.0, 2e-3, 400)
    np.testing.assert_array_equal(g.get_states(), o.get_states())
    gates = o.get_state(1)
    off = np.concatenate([[0], np.cumsum(mesh["n_nodes"])])
    untouched = [f for f in range(12) if np.all(gates[off[f]:off[f + 1]] == 0.05)]
    assert 0 < len(untouched) < 12, "some fibres must still be unstimulated at t = 0.8 ms"


@pytest.mark.parametrize("kernels,pair", [("split", 1), ("split", 0), ("fused", 1)])
def test_only_compute_if_stimulated_idle_fibres_spikes_and_algebraics(kernels, pair):
    """onlyComputeIfHasBeenStimulated with fibres that stay unstimulated: waves
    and workgroups none of whose Heun steps computes skip their states, the
    spike detector still follows Vm (changed by diffusion only), and a node whose
    last Heun step did not compute keeps the algebraics it had (zero before its
    first computed step), as the reference's skipped point buffers do; fibres of
    200-600 nodes put stimulated and idle fibres in one wave. Bit-exact vs the
    oracle over several calls."""
    mesh, _ = uneven(12, lo=200, hi=600, seed=9)
    g = _gpu_batch(mesh, scheme="IE", kernels=kernels, thomas="serial")
    g.set_tuning(pair_halves=pair)
    o = OracleBatch(**mesh, scheme="IE")
    for b in (g, o):
        _firing_setup(b, mesh, width=4)
        b.set_only_compute_if_stimulated(True)
        b.enable_spike_detection(-70.0)
        b.set_algebraics_for_transfer([0, 3, 5])
    t = 0.0
    for k in (150, 1, 249):
        for b in (g, o):
            b.advance(t, 2e-3, k)
        t += k * 2e-3
        np.testing.assert_array_equal(g.get_states(), o.get_states(), err_msg=f"after t = {t}")
        np.testing.assert_array_equal(g.get_algebraics(), o.algebraics(), err_msg=f"algebraics after t = {t}")
    for a, b in zip(g.spikes(), o.spikes()):
        np.testing.assert_array_equal(a, b)
    off = np.concatenate([[0], np.cumsum(mesh["n_nodes"])])
    idle = [f for f in range(12) if np.all(o.get_state(1)[off[f]:off[f + 1]] == 0.05)]
    assert 0 < len(idle) < 12, "some fibres must stay unstimulated"
    alg = o.algebraics()
    assert all(np.all(alg[:, off[f]:off[f + 1]] == 0) for f in idle), "idle fibres keep zero algebraics"
    assert (o.spikes()[1] > 0).sum() > 0, "stimulated fibres must fire"


@pytest.mark.parametrize("kernels,n0D", [("split", 1), ("split", 3), ("fused", 1)])
def test_only_compute_if_stimulated_point_buffer_order(kernels, n0D):
    """onlyComputeIfHasBeenStimulated in the reference's point-buffer order
    (compute.tpp:131-216) against the oracle's sweep over buffers of four nodes:
    equal fibres of 301 nodes (buffers straddle fibre ends, each taking its
    first node's fibre), the stimulation checked in the centre buffer only, and
    in the half step of a fibre's first stimulation the buffers before the
    centre skip the half while those after it compute all of it. Bit-exact;
    the per-fibre rule gives other states on the same inputs."""
    F, n = 10, 301
    mesh = dict(n_nodes=np.full(F, n, np.int32), element_lengths=np.full(F * (n - 1), 0.01),
                prefactor=np.full(F, 0.007656))
    nmj = np.full(F, n // 2, np.int32)
    dt = 2e-3
    g = _gpu_batch(mesh, scheme="IE", kernels=kernels, thomas="serial")
    o = OracleBatch(**mesh, scheme="IE")
    o_fibre = OracleBatch(**mesh, scheme="IE")
    for b in (g, o, o_fibre):
        _firing_setup(b, mesh, width=4)
        b.set_only_compute_if_stimulated(True)
        b.enable_spike_detection(0.0)
    g.set_point_buffer_order(4, nmj)
    o.enable_vc_sweep(nmj, width=4, equilibrium_skip=False)
    for k in range(3):
        for b in (g, o, o_fibre):
            b.advance(k * 200 * dt, dt, 200, dt0D=dt / 2 / n0D, n0D=n0D)
        np.testing.assert_array_equal(g.get_states(), o.get_states(), err_msg=f"call {k}")
    np.testing.assert_array_equal(g.spikes()[0], o.spikes()[0])
    assert (o.spikes()[1] > 0).sum() > 0, "fibres must fire"
    assert not np.array_equal(o.get_states(), o_fibre.get_states()), "the sweep order must matter here"
    cnt = o.vc_counters()
    assert cnt["skipped_not_stimulated"] > 0 and cnt["computed"] > 0
