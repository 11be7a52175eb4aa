"""Pins of the oracle functions the reference's own tests do not fix.

Each check derives the expected value from something other than the oracle's
own formula:
  * Strang composition: second-order convergence in the split step (a first
    order, Lie-type composition is shown to fail the same check), and one split
    step recomposed from the single-purpose primitives (Heun step of one cell,
    tridiagonal solve) with a stimulus in the second half only, which fixes the
    order of the halves and the second half's start time
  * approximate logarithm: the Taylor polynomials of log about 1, 3 and 9 built
    from the series definition (degrees and branch points as stated at
    GEN/03_generator_vc.cpp:431-490) followed by the stated Newton step against
    the approximate exponential; and the stated error bound on [0.2, 19]
  * algebraicsForTransfer: the second evaluation of the last Heun step,
    recomputed from the model's right-hand side
  * the reference's point-buffer sweep: equal to the fibre-order sweep without
    skips, and the skip pattern of onlyComputeIfHasBeenStimulated in buffer
    order as compute.tpp:172-206 states it
  * initializeStatesToEquilibrium: the RK4 search ends at the model's resting
    state, solved independently from the Hodgkin-Huxley rate functions
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle_binding import OracleBatch, oracle_library, ORACLE_BUILD, P, _dp, equilibrium
from cases import vc_lane_stim_range, poisson_firing_table, hh_steady_state_gates, hh_rate_constants

SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"
HH_PARAMS = np.array([0.0, 120.0, 36.0, 0.3])


def _arr(lib, fn, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    getattr(lib, fn)(P(x, _dp), P(y, _dp), ctypes.c_long(x.size))
    return y


# ---------------------------------------------------------------------------
# approximate logarithm
# ---------------------------------------------------------------------------

def _log_taylor(x, a, degree):
    """Taylor polynomial of log about a: log a + sum_k (-1)^(k+1) ((x-a)/a)^k / k"""
    t = (x - a) / a
    s = np.full_like(x, math.log(a))
    for k in range(1, degree + 1):
        s = s + (-1) ** (k + 1) * t ** k / k
    return s


def test_log_approximation_is_taylor_plus_newton():
    """logarithmic(x): Taylor about 1 (degree 6) below 2, about 3 (degree 4) on
    [2, 6), about 9 (degree 4) from 6 on, then y -= 1 - x/exponential(y).
    On [0.2, 19] (the stated range) the two agree to 2e-12 absolute (only the
    summation order differs); a dropped term, a wrong coefficient or a wrong
    branch point differs by > 1e-6."""
    lib = oracle_library("hh")
    x = np.concatenate([np.linspace(0.2, 19.0, 300001), [2.0, 6.0, np.nextafter(2.0, 0), np.nextafter(6.0, 0), 1.0]])
    y0 = np.where(x >= 6, _log_taylor(x, 9.0, 4), np.where(x >= 2, _log_taylor(x, 3.0, 4), _log_taylor(x, 1.0, 6)))
    y1 = y0 - (1 - x / _arr(lib, "fmso_exp_approx_n", y0))
    got = _arr(lib, "fmso_log_approx_n", x)
    assert np.max(np.abs(got - y1)) < 2e-12
    # the same check catches a perturbed polynomial
    y0_bad = np.where(x >= 6, _log_taylor(x, 9.0, 3), y0)
    y1_bad = y0_bad - (1 - x / _arr(lib, "fmso_exp_approx_n", y0_bad))
    assert np.max(np.abs(got - y1_bad)) > 1e-6


def test_log_approximation_stated_bound():
    """'The relative error of this implementation is below 0.04614465854334056
    for x in [0.2,19]' (GEN/03_generator_vc.cpp:480); the Newton step reduces the
    Taylor polynomials' error"""
    lib = oracle_library("hh")
    x = np.linspace(0.2, 19.0, 190001)
    x = x[np.abs(x - 1.0) > 1e-3]
    got = _arr(lib, "fmso_log_approx_n", x)
    rel = np.abs(got - np.log(x)) / np.abs(np.log(x))
    assert rel.max() < 0.04614465854334056
    y0 = np.where(x >= 6, _log_taylor(x, 9.0, 4), np.where(x >= 2, _log_taylor(x, 3.0, 4), _log_taylor(x, 1.0, 6)))
    assert np.max(np.abs(got - np.log(x))) < np.max(np.abs(y0 - np.log(x)))


# ---------------------------------------------------------------------------
# Strang composition
# ---------------------------------------------------------------------------

def _bump_batch(n=201, length=1.0, scheme="CN"):
    h = length / (n - 1)
    mesh = dict(n_nodes=np.array([n], np.int32), element_lengths=np.full(n - 1, h), prefactor=np.array([3.828 / 500]))
    o = OracleBatch(**mesh, scheme=scheme, initial_states=[-75.0, *hh_steady_state_gates(-75.0)])
    xs = np.arange(n) * h
    o.states[0][:] = -75.0 + 5.0 * np.exp(-((xs - 0.5 * length) / 0.1) ** 2)
    return o, mesh


def _strang_run(dt, T=0.5):
    o, _ = _bump_batch()
    o.advance(0.0, dt, int(round(T / dt)))
    return o.get_states()


def _lie_run(dt, T=0.5):
    """reaction over the whole step (two Heun steps of dt/2), then the solve:
    a first-order composition built from the oracle's primitives"""
    o, mesh = _bump_batch()
    lib = o.lib
    y = o.get_states()
    n = y.shape[1]
    for s in range(int(round(T / dt))):
        for i in range(n):
            c = np.ascontiguousarray(y[:, i])
            lib.fmso_heun_cell(P(c, _dp), P(HH_PARAMS, _dp), s * dt, dt / 2, 2, 0)
            y[:, i] = c
        v = np.ascontiguousarray(y[0])
        lib.fmso_thomas(P(v, _dp), n, P(mesh["element_lengths"], _dp), float(mesh["prefactor"][0]), dt, 1, 0)
        y[0] = v
    return y


def test_strang_second_order_in_dt_split():
    """sub-threshold, smooth Vm bump on one fibre (CN, one Heun step of dt/2 per
    half): the error against a fine run falls by 4x per halving of dt; the
    first-order Lie composition of the same primitives falls by 2x"""
    ref = _strang_run(0.05 / 64)
    err = [np.max(np.abs(_strang_run(dt) - ref)) for dt in (0.05, 0.025, 0.0125)]
    r = [err[0] / err[1], err[1] / err[2]]
    assert all(3.5 < q < 4.6 for q in r), (err, r)
    lie = [np.max(np.abs(_lie_run(dt) - ref)) for dt in (0.05, 0.025, 0.0125)]
    rl = [lie[0] / lie[1], lie[1] / lie[2]]
    assert all(q < 2.6 for q in rl), (lie, rl)


def test_strang_step_from_primitives():
    """one split step = Heun half at t, solve over dt, Heun half at t + dt/2:
    a current pulse that switches on at t + dt/2 acts in the second half only.
    Recomposed from fmso_heun_cell / fmso_thomas, bit for bit."""
    n, dt, t0 = 41, 0.02, 0.3
    mesh = dict(n_nodes=np.array([n], np.int32), element_lengths=np.full(n - 1, 0.01), prefactor=np.array([0.0077]))
    o = OracleBatch(**mesh, scheme="CN")
    o.set_current_stimulation(19, 22, onset=t0 + dt / 2, duration=1.0, amplitude=50.0, stim_param=0)
    y = o.get_states()
    o.advance(t0, dt, 1)
    lib = o.lib
    for i in range(n):   # first half, unstimulated
        c = np.ascontiguousarray(y[:, i])
        lib.fmso_heun_cell(P(c, _dp), P(HH_PARAMS, _dp), t0, dt / 2, 1, 0)
        y[:, i] = c
    v = np.ascontiguousarray(y[0])
    lib.fmso_thomas(P(v, _dp), n, P(mesh["element_lengths"], _dp), 0.0077, dt, 1, 0)
    y[0] = v
    for i in range(n):   # second half at t0 + dt/2: i_Stim on the stimulated nodes
        c = np.ascontiguousarray(y[:, i])
        prm = HH_PARAMS.copy()
        if 19 <= i < 22:
            prm[0] = 50.0
        lib.fmso_heun_cell(P(c, _dp), P(prm, _dp), t0 + dt / 2, dt / 2, 1, 0)
        y[:, i] = c
    np.testing.assert_array_equal(o.get_states(), y)


# ---------------------------------------------------------------------------
# algebraicsForTransfer
# ---------------------------------------------------------------------------

def _rhs(lib, y, prm, na, mode=0):
    r = np.zeros(y.size)
    a = np.zeros(na)
    lib.fmso_rhs(P(np.ascontiguousarray(y), _dp), P(np.ascontiguousarray(prm), _dp), 0.0, P(r, _dp), P(a, _dp), mode)
    return r, a


@pytest.mark.parametrize("n0D", [1, 2])
def test_algebraics_for_transfer_second_evaluation_of_last_heun_step(n0D):
    """after advance(), algebraics[q] = algebraic idx[q] of rhs(y*) of the last
    Heun step of the last reaction half, y* = y + dt0D rhs(y)"""
    n, dt, k = 30, 0.01, 3
    mesh = dict(n_nodes=np.array([n], np.int32), element_lengths=np.full(n - 1, 0.01), prefactor=np.array([0.0077]))
    idx = [8, 0, 5, 1]
    o = OracleBatch(**mesh, scheme="IE")
    o.set_current_stimulation(10, 13, onset=0.0, duration=1.0, amplitude=30.0, stim_param=0)
    o.set_algebraics_for_transfer(idx)
    o.advance(0.0, dt, k - 1, dt0D=dt / 2 / n0D, n0D=n0D)
    # the last step, by hand up to its last Heun step
    y = o.get_states()
    mid = OracleBatch(**mesh, scheme="IE")
    mid.states[:] = y
    mid.set_current_stimulation(10, 13, onset=0.0, duration=1.0, amplitude=30.0, stim_param=0)
    o.advance((k - 1) * dt, dt, 1, dt0D=dt / 2 / n0D, n0D=n0D)
    lib = o.lib
    ydt = dt / 2 / n0D
    for i in range(n):
        prm = HH_PARAMS.copy()
        if 10 <= i < 13:
            prm[0] = 30.0
        c = np.ascontiguousarray(y[:, i])
        lib.fmso_heun_cell(P(c, _dp), P(prm, _dp), 0.0, ydt, n0D, 0)
        y[:, i] = c
    v = np.ascontiguousarray(y[0])
    lib.fmso_thomas(P(v, _dp), n, P(mesh["element_lengths"], _dp), 0.0077, dt, 0, 0)
    y[0] = v
    want = np.zeros((len(idx), n))
    for i in range(n):
        prm = HH_PARAMS.copy()
        if 10 <= i < 13:
            prm[0] = 30.0
        c = np.ascontiguousarray(y[:, i])
        if n0D > 1:
            lib.fmso_heun_cell(P(c, _dp), P(prm, _dp), 0.0, ydt, n0D - 1, 0)
        r, _ = _rhs(lib, c, prm, 9)
        ys = c + ydt * r
        _, a = _rhs(lib, ys, prm, 9)
        want[:, i] = a[idx]
    np.testing.assert_array_equal(o.algebraics(), want)


def test_algebraics_for_transfer_parameter_algebraic_keeps_its_value():
    """an algebraic that is a parameter (Shorten: ALGEBRAIC[32] = I_Stim,
    parametersUsedAsAlgebraic=[32]) has no case in the generated copy
    (GEN/03_generator_vc.cpp:1147-1152): its transfer slot is not written"""
    if not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{SHORTEN}.so")):
        pytest.skip("oracle Shorten model not built (needs the reference's CellML input)")
    mesh = dict(n_nodes=np.array([8], np.int32), element_lengths=np.full(7, 0.01), prefactor=np.array([0.0077]))
    o = OracleBatch(**mesh, scheme="IE", model=SHORTEN)
    o.set_algebraics_for_transfer([32, 0])
    o._alg[0][:] = 123.0
    o.advance(0.0, 2.5e-5, 3)
    assert np.all(o.algebraics()[0] == 123.0)
    assert np.all(o.algebraics()[1] != 0.0)


# ---------------------------------------------------------------------------
# point-buffer sweep
# ---------------------------------------------------------------------------

def _vc_case(n_fibers=6, n=101, width=4, seed=2):
    dt = 1e-3
    mesh = dict(n_nodes=np.full(n_fibers, n, np.int32), element_lengths=np.full(n_fibers * (n - 1), 0.01),
                prefactor=np.full(n_fibers, 3.828 / 500))
    table = poisson_firing_table(400, 3, 0.05, dt, 0.1, seed, first_at_zero=True)
    nmj = np.full(n_fibers, n // 2, np.int32)
    off = np.arange(n_fibers) * n
    sb, se = vc_lane_stim_range(off, nmj, mesh["n_nodes"], width)
    stim = dict(stim_begin=sb, stim_end=se, firing_events=table, motor_unit=np.arange(n_fibers, dtype=np.int32) % 3,
                call_frequency=1 / dt, repeat_after_first_call=0.0, call_enable_begin=0.0)
    return mesh, stim, nmj, dt


@pytest.mark.parametrize("width", [1, 4, 8])
def test_point_buffer_sweep_equals_fibre_sweep_without_skips(width):
    """with both skip heuristics off the buffer order does not matter: the
    point-buffer sweep and the fibre sweep (stimulated nodes as the W-lane
    quirk places them) agree bit for bit, algebraics included"""
    mesh, stim, nmj, dt = _vc_case(width=width)
    a = OracleBatch(**mesh, scheme="CN")
    b = OracleBatch(**mesh, scheme="CN")
    b.enable_vc_sweep(nmj, width)
    for o in (a, b):
        o.set_firing_table_stimulation(**stim, log_capacity=64)
        o.set_algebraics_for_transfer([0, 4, 8])
        o.advance(0.0, dt, 150)
    np.testing.assert_array_equal(a.get_states(), b.get_states())
    np.testing.assert_array_equal(a.algebraics(), b.algebraics())
    np.testing.assert_array_equal(a.stimulation_log()[0], b.stimulation_log()[0])


def test_point_buffer_order_only_if_stimulated():
    """onlyComputeIfHasBeenStimulated in buffer order (compute.tpp:172-206): in
    the reaction half in which a fibre is first stimulated, its buffers before
    the centre buffer (the one holding the NMJ) skip the half and its later
    buffers compute it; fibres never stimulated stay at their initial states"""
    n, W = 101, 4
    mesh, stim, nmj, dt = _vc_case(n_fibers=2, n=n, width=W)
    stim["firing_events"] = np.zeros((400, 3), np.uint8)
    stim["firing_events"][:, 0] = 1   # motor unit 0 (fibre 0) fires from t = 0, fibre 1 (MU 1) never
    o = OracleBatch(**mesh, scheme="CN")
    o.set_firing_table_stimulation(**stim)
    o.set_only_compute_if_stimulated(True)
    o.enable_vc_sweep(nmj, W)
    y0 = o.get_states()
    o.advance(0.0, dt, 1)   # one split step, one Heun step per half
    c = o.vc_counters()
    n_buf = (2 * n + W - 1) // W
    centre = nmj[0] // W               # fibre 0 starts at node 0
    f1_first = -(-n // W)              # first buffer with int(b * W / n) == 1
    buffers_f0, buffers_f1 = f1_first, n_buf - f1_first
    # first half: fibre 0's buffers before the centre skip, the rest compute;
    # second half: all of fibre 0 computes; fibre 1 never computes
    assert c["skipped_not_stimulated"] == centre + 2 * buffers_f1
    assert c["computed"] == (buffers_f0 - centre) + buffers_f0
    # the gates (no diffusion) of fibre 1 stay at their initial values, except
    # its first nodes, which share fibre 0's last buffer and are computed with it
    straddle = f1_first * W - n
    np.testing.assert_array_equal(o.get_states()[1:, n + straddle:], y0[1:, n + straddle:])
    assert np.all(o.get_states()[1:, n:n + straddle] != y0[1:, n:n + straddle])


def test_equilibrium_skip_is_a_no_op_while_states_change():
    """disableComputationWhenStatesAreCloseToEquilibrium skips nothing while
    every buffer's states change by more than 1e-6 relative per half; with the
    reference's tolerance the results equal the sweep without it"""
    mesh, stim, nmj, dt = _vc_case(width=4)
    a = OracleBatch(**mesh, scheme="CN")
    b = OracleBatch(**mesh, scheme="CN")
    a.enable_vc_sweep(nmj, 4)
    b.enable_vc_sweep(nmj, 4, equilibrium_skip=True)
    for o in (a, b):
        o.set_firing_table_stimulation(**stim)
        o.advance(0.0, dt, 100)
    assert b.vc_counters()["skipped_equilibrium"] == 0
    np.testing.assert_array_equal(a.get_states(), b.get_states())


def test_equilibrium_skip_freezes_a_resting_fibre():
    """a fibre at the resting state of its model (libm exponential, the one the
    RK4 search uses): every buffer becomes inactive after
    the first half, skips its Heun steps from then on, and compute1D is skipped
    once the inactive counter reaches the number of buffers"""
    lib = oracle_library("hh")
    y_rest, it, _, _ = equilibrium(lib, [-75, 0.05, 0.6, 0.325], HH_PARAMS)
    assert it > 0
    n = 41
    mesh = dict(n_nodes=np.array([n], np.int32), element_lengths=np.full(n - 1, 0.01), prefactor=np.array([0.0077]))
    o = OracleBatch(**mesh, scheme="CN", initial_states=y_rest, exp_mode="libm")
    o.set_firing_table_stimulation(n // 2, n // 2 + 3, np.zeros((10, 1), np.uint8), [0], 1e3, 0.0, 0.0)
    o.enable_vc_sweep(np.array([n // 2]), 4, equilibrium_skip=True)
    o.advance(0.0, 1e-3, 20)
    c = o.vc_counters()
    n_buf = (n + 3) // 4
    # every buffer computes the first half (all start active) and turns inactive;
    # compute0D marks the first and last buffer active again in every half
    # (compute.tpp:155-156), so those two compute in each of the other 39 halves
    assert c["computed"] == n_buf + 2 * 39
    assert c["skipped_equilibrium"] == 39 * (n_buf - 2)
    eq = o.vc_equilibrium_states()
    assert np.all(eq == 0)


# ---------------------------------------------------------------------------
# initializeStatesToEquilibrium
# ---------------------------------------------------------------------------

def test_equilibrium_rk4_ends_at_the_resting_state():
    """the RK4 search (03_cellml_adapter.tpp:457-520) from the model file's
    initial values ends where the ionic current vanishes with the gates at
    their steady state alpha/(alpha+beta); V solved independently by bisection
    from the textbook rate functions (libm exp)"""
    lib = oracle_library("hh")
    y, it, inc, dt = equilibrium(lib, [-75, 0.05, 0.6, 0.325], HH_PARAMS)
    assert it > 0 and inc < 1e-5 and dt == 1e-4

    def current(v):
        m, h, nn = hh_steady_state_gates(v)
        return 120 * m ** 3 * h * (v - 40.0) + 36 * nn ** 4 * (v + 87.0) + 0.3 * (v + 64.387)

    lo, hi = -80.0, -70.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if current(lo) * current(mid) <= 0:
            hi = mid
        else:
            lo = mid
    v_rest = 0.5 * (lo + hi)
    assert abs(y[0] - v_rest) < 1e-3
    # the search stops at |dx/dt| < 1e-5 per ms: each gate is then within
    # 1e-5 tau of its steady state, tau = 1/(alpha+beta) (8.5 ms for h)
    tau = np.array([1.0 / (a + b) for a, b in hh_rate_constants(y[0])])
    assert np.all(np.abs(y[1:] - hh_steady_state_gates(y[0])) <= 1.05e-5 * tau)
    # started at the fixed point it stops at once
    y2, it2, _, _ = equilibrium(lib, y, HH_PARAMS)
    assert it2 == 0


def test_equilibrium_rk4_halves_a_divergent_step():
    """an RK4 step width at which the search diverges (increment > 1e4) is
    halved and the search restarts from the initial values"""
    lib = oracle_library("hh")
    y, it, inc, dt = equilibrium(lib, [-75, 0.05, 0.6, 0.325], HH_PARAMS, dt=1.0)
    assert dt == 0.5 and it > 0 and inc < 1e-5
    y_ref, _, _, _ = equilibrium(lib, [-75, 0.05, 0.6, 0.325], HH_PARAMS)
    assert abs(y[0] - y_ref[0]) < 1e-4
