"""CPU checks that pin the oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something other than a retyping of
its own formula:
  * the reference's own known-answer values (tests/golden/reference_kats.json,
    written by tests/golden/extract_reference_kats.py from the reference's tests)
  * closed forms / invariants of the discretisation (dense solve, conservation
    of 1^T M u, discrete Fourier modes, fixed points)
  * the reference's documented numbers (exp approximation error table,
    W-lane stimulation nodes)
  * the reference's model files compiled as they are (oracle/_ref), where built
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from oracle_binding import OracleBatch, oracle_library, ORACLE_DIR, ORACLE_BUILD, P, _dp, schedule
from cases import c1, vc_lane_stim_range

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")
KATS = json.load(open(GOLDEN))
SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"


def _euler(lib, y, p, dt, n, mode):
    y = np.ascontiguousarray(y, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    lib.fmso_euler_cell(P(y, _dp), P(p, _dp), ctypes.c_double(0.0), ctypes.c_double(dt), ctypes.c_long(n), mode)
    return y


def test_hh_rhs_reference_kat_bit_exact():
    """HodgkinHuxley* KAT (cellml.cpp:64): 90001 explicit-Euler steps with libm exp."""
    k = KATS["hodgkin_huxley_explicit_euler"]
    y = _euler(oracle_library("hh"), k["initial_states"], [k["i_stim"], 120, 36, 0.3], k["dt"], k["n_steps"], 1)
    np.testing.assert_array_equal(y, np.array(k["states"]))


def test_hh_kat_distinguishes_approximate_exp():
    """the same run with the (1+x/1024)^1024 exponential is off by > 1e-4 relative:
    the KAT pins the exp mode, not just the rate functions"""
    k = KATS["hodgkin_huxley_explicit_euler"]
    y = _euler(oracle_library("hh"), k["initial_states"], [k["i_stim"], 120, 36, 0.3], k["dt"], k["n_steps"], 0)
    assert np.max(np.abs(y - k["states"]) / np.abs(k["states"])) > 1e-4


def test_shorten_rhs_reference_kat():
    """ShortenOpenCOR KAT (cellml.cpp:215): 56 states after 90001 explicit-Euler steps.
    Agreement to 1e-14 relative (the reference run evaluates pow(x,n) with libm,
    the FastMonodomainSolver form with recursive products)."""
    path = os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}.so")
    if not os.path.exists(path):
        pytest.skip("oracle Shorten model not built (needs the reference's CellML input)")
    k = KATS["shorten_explicit_euler"]
    lib = oracle_library(SHORTEN)
    y0 = np.zeros(56)
    lib.fmso_initial_states(P(y0, _dp))
    y = _euler(lib, y0, k["parameters"], k["dt"], k["n_steps"], 1)
    ref = np.array(k["states"])
    assert np.max(np.abs(y - ref) / np.abs(ref)) < 1e-14


def test_diffusion_implicit_euler_lumped_kat():
    """ImplicitEuler1D KAT (diffusion.cpp:150), lumped mass, prefactor 1, 5 steps of
    dt 0.02. The reference solved with PETSc GMRES at its default rtol 1e-5."""
    k = KATS["diffusion_implicit_euler_lumped"]
    lib = oracle_library("hh")
    h = np.diff(np.array(k["x"]))
    u = np.array(k["u0"], np.float64)
    for _ in range(k["n_steps"]):
        lib.fmso_thomas(P(u, _dp), 6, P(h, _dp), 1.0, k["dt"], 0, 1)
    np.testing.assert_allclose(u, k["u"], rtol=0, atol=1e-5)
    assert np.max(np.abs(u - k["u"])) > 1e-8   # a direct solve, not the iterative one


def _dense(n, h, pf, dt, scheme, mass="consistent"):
    K = np.zeros((n, n))
    M = np.zeros((n, n))
    for e in range(n - 1):
        k = pf / h[e]
        K[e:e + 2, e:e + 2] += np.array([[-k, k], [k, -k]])
        if mass == "consistent":
            M[e:e + 2, e:e + 2] += h[e] * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
        else:
            M[e:e + 2, e:e + 2] += h[e] * np.array([[0.5, 0], [0, 0.5]])
    if scheme == 0:
        return K - M / dt, -M / dt, K, M
    return K / 2 - M / dt, -K / 2 - M / dt, K, M


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("n", [2, 3, 5, 17, 64])
def test_thomas_matches_dense_solve(scheme, n):
    rng = np.random.default_rng(n + 10 * scheme)
    h = rng.uniform(0.005, 0.02, n - 1)
    u = rng.normal(-60, 20, n)
    pf, dt = 0.007656, 1e-3
    A, B, _, _ = _dense(n, h, pf, dt, scheme)
    expect = np.linalg.solve(A, B @ u)
    got = u.copy()
    oracle_library("hh").fmso_thomas(P(got, _dp), n, P(h, _dp), pf, dt, scheme, 0)
    np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("scheme", [0, 1])
def test_thomas_conserves_mass_weighted_sum(scheme):
    """natural Neumann ends: 1^T K = 0, so 1^T M u is invariant for IE and CN"""
    rng = np.random.default_rng(3)
    n = 200
    h = rng.uniform(0.005, 0.02, n - 1)
    u = rng.normal(-60, 20, n)
    _, _, _, M = _dense(n, h, 0.0077, 1e-2, scheme)
    before = np.ones(n) @ M @ u
    lib = oracle_library("hh")
    for _ in range(50):
        lib.fmso_thomas(P(u, _dp), n, P(h, _dp), 0.0077, 1e-2, scheme, 0)
    after = np.ones(n) @ M @ u
    assert abs(after - before) <= 1e-12 * abs(before)


@pytest.mark.parametrize("scheme", [0, 1])
def test_thomas_fourier_mode_decay(scheme):
    """on a uniform mesh cos(j*pi*i/(n-1)) is an eigenvector of K and M with
    eigenvalues -(pf/h)(2-2cos t) and (h/6)(4+2cos t); IE multiplies it by
    (m/dt)/(m/dt+k), CN by (m/dt-k/2)/(m/dt+k/2)"""
    n, h, pf, dt = 101, 0.01, 0.007656, 1e-3
    for j in (1, 7, 50, 100):
        th = j * np.pi / (n - 1)
        u = np.cos(th * np.arange(n))
        k = pf / h * (2 - 2 * np.cos(th))
        m = h / 6 * (4 + 2 * np.cos(th))
        g = (m / dt) / (m / dt + k) if scheme == 0 else (m / dt - k / 2) / (m / dt + k / 2)
        got = u.copy()
        oracle_library("hh").fmso_thomas(P(got, _dp), n, P(np.full(n - 1, h), _dp), pf, dt, scheme, 0)
        np.testing.assert_allclose(got, g * u, rtol=0, atol=1e-12)


def test_thomas_uniform_fixed_point():
    n = 50
    u = np.full(n, -75.0)
    oracle_library("hh").fmso_thomas(P(u, _dp), n, P(np.full(n - 1, 0.01), _dp), 0.01, 1e-3, 1, 0)
    np.testing.assert_allclose(u, -75.0, rtol=1e-15)


def test_heun_second_order():
    """the Heun step converges with order 2 on a single HH cell (no stimulus
    current clamp), error of dt against dt/2 vs dt/2 against dt/4"""
    lib = oracle_library("hh")
    y0 = np.array([-60.0, 0.05, 0.6, 0.325])
    p = np.array([10.0, 120, 36, 0.3])

    def run(dt):
        y = y0.copy()
        lib.fmso_heun_cell(P(y, _dp), P(p, _dp), 0.0, dt, int(round(1.0 / dt)), 1)
        return y

    a, b, c = run(2e-3), run(1e-3), run(5e-4)
    ratio = np.linalg.norm(a - b) / np.linalg.norm(b - c)
    assert 3.5 < ratio < 4.5, ratio


def _exp_approx_py(x):
    x = 1.0 + x / 1024.0
    for _ in range(10):
        x *= x
    return x


def test_exp_approximation_error_table():
    """the relative errors listed with the approximation (03_generator_vc.cpp:398-405)"""
    lib = oracle_library("hh")
    table = {1: 0.00048784455634225593, 3: 0.0043763626896140342, 5: 0.012093715791500804,
             9: 0.038557535762274039, 12: 0.067389808619653505}
    for x, err in table.items():
        # the oracle's exponential through the HH rate a[5] = 4 exp(-(V+75)/18) with V = -75 - 18 x
        y = np.array([-75.0 - 18.0 * x, 0.05, 0.6, 0.325])
        r = np.zeros(4)
        a = np.zeros(9)
        lib.fmso_rhs(P(y, _dp), P(np.array([0, 120, 36, 0.3]), _dp), 0.0, P(r, _dp), P(a, _dp), 0)
        approx = a[5] / 4.0
        assert approx == _exp_approx_py(x)
        assert abs(abs(approx - math.exp(x)) / math.exp(x) - err) < 1e-6 * err + 1e-12


def test_schedule_step_count_rule():
    """RepeatedCall + TimeSteppingScheme bookkeeping: every call of the HH
    example is one Strang step with 5 Heun steps of 2e-4 (settings_fast_fiber.py:15-21);
    for C1 10000 calls of one step with dt0D = dt/2"""
    lib = oracle_library("hh")
    s = schedule(lib, 100.0, 2e-3, 2e-3, 2e-3, 2e-4, 2e-4, 2e-3, 2e-3)
    assert s.shape[0] == 50000
    assert np.all(s[:, 2] == 1) and np.all(s[:, 4] == 5) and np.all(s[:, 6] == 1)
    np.testing.assert_allclose(s[:, 3], 2e-4, rtol=1e-9)
    s = schedule(lib, 10.0, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3)
    assert s.shape[0] == 10000 and np.all(s[:, 4] == 1)
    k = np.arange(10000)
    np.testing.assert_array_equal(s[:, 0], 0.0 + k.astype(float) / 10000 * 10.0)
    np.testing.assert_array_equal(s[:, 1], (s[:, 0] + 1e-3) - s[:, 0])
    # a span slightly longer than the step keeps one step (epsilon 0.1) ...
    s = schedule(lib, 1.09, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
    assert s.shape[0] == 1 and s[0, 1] == 1.09
    # ... a longer one is split in two
    s = schedule(lib, 1.11, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
    assert s.shape[0] == 2


def test_vc_lane_stimulation_nodes():
    """W-aligned stimulated nodes of the Vc path: for a 1001-node fibre at offset
    0 (NMJ 500) W=4 clamps 500..502 and W=8 clamps 496..498"""
    b, e = vc_lane_stim_range([0], [500], np.array([1001]), 4)
    assert (b[0], e[0]) == (500, 503)
    b, e = vc_lane_stim_range([0], [500], np.array([1001]), 8)
    assert (b[0], e[0]) == (496, 499)
    b, e = vc_lane_stim_range([1001], [500], np.array([1001]), 4)
    assert (b[0], e[0]) == (499, 502)


def test_stimulation_windows_and_jitter():
    """firing-table state machine: with frequency f and RepeatAfterFirstCall r a
    window opens every 1/(f + jitter*f) and stays open for r; the log holds
    the window onsets (compute.tpp:538-578)"""
    mesh = dict(n_nodes=np.array([20]), element_lengths=np.full(19, 0.01), prefactor=np.array([0.0077]))
    o = OracleBatch(**mesh, scheme="IE")
    table = np.ones((1, 1), np.uint8)
    o.set_firing_table_stimulation(9, 12, table, [0], call_frequency=0.5, repeat_after_first_call=0.1,
                                   call_enable_begin=1.0, jitter_lists=[[0.25, -0.2]], log_capacity=64)
    o.advance(0.0, 1e-2, 800)
    t, f = o.stimulation_log()
    # onsets: 1.0 (= enableBegin), then periods 1/(0.5+0.125) = 1.6, 1/(0.5-0.1) = 2.5, 1.6, ...
    expect = [1.0, 2.6, 5.1, 6.7]
    np.testing.assert_allclose(t[:4], expect, atol=1e-9)


def test_oracle_rhs_against_reference_model_files():
    """oracle models vs the reference's model files compiled as they are
    (oracle/_ref, libm exp): rates agree to 1e-13 relative at random states"""
    ref_dir = os.path.join(ORACLE_DIR, "_ref")
    cases = [("hh", "hodgkin_huxley_1952", 4, 9, 9, lambda C: [C[2], C[3], C[4], C[5]])]
    if os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}.so")):
        cases.append((SHORTEN, SHORTEN, 56, 71, 105, lambda C: [0.0, C[65]]))
    checked = 0
    for oname, rname, ns, na, nc, params in cases:
        path = os.path.join(ref_dir, f"libref_{rname}.so")
        if not os.path.exists(path):
            continue
        ref = ctypes.CDLL(path)
        ol = oracle_library(oname)
        C = np.zeros(nc)
        R = np.zeros(ns)
        S = np.zeros(ns)
        A = np.zeros(na)
        ref.initConsts(P(C, _dp), P(R, _dp), P(S, _dp))
        rng = np.random.default_rng(1)
        for _ in range(20):
            y = S * (1 + 0.05 * rng.standard_normal(ns))
            y[0] = rng.uniform(-90, 30)
            ref.computeRates(ctypes.c_double(1.0), P(C, _dp), P(R, _dp), P(y, _dp), P(A, _dp))
            r = np.zeros(ns)
            a = np.zeros(na)
            p = np.array(params(C), np.float64)
            ol.fmso_rhs(P(y, _dp), P(p, _dp), 1.0, P(r, _dp), P(a, _dp), 1)
            np.testing.assert_allclose(r, R, rtol=1e-13, atol=1e-13 * np.max(np.abs(R)))
        checked += 1
    if not checked:
        pytest.skip("oracle/_ref not built (needs the reference checkout)")


def test_c1_action_potential_propagates():
    """C1 in the oracle: the stimulus starts an action potential that travels
    along the fibre at a speed in the physiological range"""
    mesh, stim = c1()
    o = OracleBatch(**mesh, scheme="CN", threads=4)
    o.set_current_stimulation(**stim)
    o.enable_spike_detection(0.0)
    o.advance(0.0, 1e-3, 10000)
    first = o.spikes()[0]
    assert first[500] > 0 and first[650] > first[600] > first[500]
    assert first[400] == first[600]   # symmetric about the centre node
    v = 0.01 * 50 / ((first[650] - first[600]) * 1e-3)   # cm/ms
    assert 0.05 < v < 1.0, v


def avx512() -> bool:
    """the host runs AVX-512 code (the eight-lane oracle builds)"""
    try:
        with open("/proc/cpuinfo") as f:
            return any(line.startswith("flags") and " avx512f" in line for line in f)
    except OSError:
        return False


@pytest.mark.parametrize("stim", ["current", "firing_table"])
def test_four_lane_reaction_matches_one_node_at_a_time(stim):
    """the HH oracle evaluates four nodes per vector operation (approximate
    exponential), eight in its AVX-512 build (where the host has AVX-512); the
    same file built one node at a time gives bit-identical
    states and spike steps, with stimulated ranges that straddle lane groups,
    fibre lengths that leave a scalar tail, and clamped lanes"""
    n = np.array([23, 37, 41, 8], np.int32)
    mesh = dict(n_nodes=n, element_lengths=np.full(int((n - 1).sum()), 0.01),
                prefactor=np.array([0.0077, 0.0061, 0.0090, 0.0077]))
    out = []
    for model in ("hh", "hh_scalar") + (("hh_w8",) if avx512() else ()):
        o = OracleBatch(**mesh, scheme="CN", model=model)
        if stim == "current":
            o.set_current_stimulation(stim_begin=np.array([5, 0, 38, 3]), stim_end=np.array([9, 3, 41, 8]),
                                      onset=np.array([0.0, 0.2, 0.1, 0.0]), duration=np.full(4, 0.1),
                                      amplitude=np.full(4, 1200.0), stim_param=0)
        else:
            o.set_firing_table_stimulation(np.array([5, 0, 38, 3]), np.array([9, 3, 41, 8]),
                                           np.ones((1, 1), np.uint8), [0, 0, 0, 0], call_frequency=2.0,
                                           repeat_after_first_call=0.05, call_enable_begin=0.0)
        o.enable_spike_detection(0.0)
        o.advance(0.0, 1e-3, 1500)
        out.append((o.states.copy(), o.spikes()[0].copy()))
    for other in out[:1] + out[2:]:
        np.testing.assert_array_equal(other[0], out[1][0])
        np.testing.assert_array_equal(other[1], out[1][1])
    assert (out[1][1] >= 0).sum() > 20, "the fibres must fire"


def test_shorten_four_lane_reaction_matches_one_node_at_a_time():
    """the generated Shorten oracle (approximate exp/log, integer powers, lane
    selects for its two conditionals) gives bit-identical states and spike
    steps four nodes at a time and one node at a time, through the upstroke of
    propagating action potentials"""
    from cases import c2
    variants = ("", "_scalar") + (("_w8",) if avx512() else ())
    paths = [os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}{s}.so") for s in variants]
    if not all(os.path.exists(p) for p in paths):
        pytest.skip("generated oracle models not built (needs the reference checkout)")
    mesh, stim = c2(n_fibers=2, n_nodes=203, length_cm=2.02, rate_hz=4000.0)
    out = []
    for model in [SHORTEN + v for v in variants]:
        o = OracleBatch(**mesh, scheme="IE", model=model, threads=2)
        o.set_firing_table_stimulation(**stim)
        o.enable_spike_detection(-40.0)
        o.advance(0.0, 2.5e-5, 12000)
        out.append((o.states.copy(), o.spikes()[0].copy()))
    for other in out[:1] + out[2:]:
        np.testing.assert_array_equal(other[0], out[1][0])
        np.testing.assert_array_equal(other[1], out[1][1])
    assert (out[1][1] >= 0).sum() > 10, "the fibres must fire"
