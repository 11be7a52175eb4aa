"""Seeded workloads shared by the parity tests, smoke() and bench.py.

Sizes follow the fibre configurations the GPU path is measured on:
  c1     HH, 1 fibre x 1001 nodes (1000 elements, 10 cm), Am=500, Cm=1,
         sigma=3.828, CN, dt=1e-3 ms, 1200 uA/cm^2 on the central 3 nodes for
         0.1 ms; Vm0 = -75 mV with the gates at their steady state there
  uneven HH, fibres of U{lo..hi} nodes, per-fibre sigma ~ U(3.0,4.5),
         stimulus node ~ U{0..n-1} and onset ~ U(0,2) ms (current pulses)
"""
from __future__ import annotations

import numpy as np

AM, CM, SIGMA = 500.0, 1.0, 3.828


def hh_rate_constants(v):
    """(alpha, beta) of the gates m, h, n of the Hodgkin-Huxley model with the
    resting potential shifted to -75 mV (textbook forms, libm exp)"""
    v = np.asarray(v, np.float64)
    am = -0.1 * (v + 50.0) / (np.exp(-(v + 50.0) / 10.0) - 1.0)
    bm = 4.0 * np.exp(-(v + 75.0) / 18.0)
    ah = 0.07 * np.exp(-(v + 75.0) / 20.0)
    bh = 1.0 / (np.exp(-(v + 45.0) / 10.0) + 1.0)
    an = -0.01 * (v + 65.0) / (np.exp(-(v + 65.0) / 10.0) - 1.0)
    bn = 0.125 * np.exp((v + 75.0) / 80.0)
    return (am, bm), (ah, bh), (an, bn)


def hh_steady_state_gates(v):
    """m, h, n = alpha/(alpha+beta) at the membrane potential v"""
    return tuple(a / (a + b) for a, b in hh_rate_constants(v))


def c1(n_elements=1000, length_cm=10.0):
    n = n_elements + 1
    h = np.full(n_elements, length_cm / n_elements)
    pf = SIGMA / (AM * CM)
    mid = n // 2
    stim = dict(stim_begin=mid - 1, stim_end=mid + 2, onset=0.0, duration=0.1, amplitude=1200.0, stim_param=0)
    y0 = np.array([-75.0, *hh_steady_state_gates(-75.0)])
    return dict(n_nodes=np.array([n], np.int32), element_lengths=h, prefactor=np.array([pf]), initial_states=y0), stim


def uneven(n_fibers, lo=400, hi=2000, seed=3, h=0.01):
    rng = np.random.default_rng(seed)
    n = rng.integers(lo, hi + 1, size=n_fibers).astype(np.int32)
    sigma = np.random.default_rng(seed + 1).uniform(3.0, 4.5, size=n_fibers)
    r5 = np.random.default_rng(seed + 2)
    nmj = np.array([r5.integers(0, k) for k in n], np.int32)
    onset = r5.uniform(0.0, 2.0, size=n_fibers)
    elem = np.concatenate([np.full(k - 1, h) for k in n])
    pf = sigma / (AM * CM)
    sb = np.maximum(nmj - 1, 0).astype(np.int32)
    se = np.minimum(nmj + 2, n).astype(np.int32)
    stim = dict(stim_begin=sb, stim_end=se, onset=onset, duration=np.full(n_fibers, 0.1),
                amplitude=np.full(n_fibers, 1200.0), stim_param=0)
    return dict(n_nodes=n, element_lengths=elem, prefactor=pf), stim


def vc_lane_stim_range(offsets, nmj, n_nodes, width=4):
    """Nodes the reference's Vc path clamps: lanes 0..min(3,W)-1 of the W-aligned
    point buffer holding the NMJ (fast_monodomain_solver_compute.tpp:161-166,
    03_generator_vc.cpp:1032-1038); clipped to the fibre."""
    offsets = np.asarray(offsets)
    nmj = np.asarray(nmj)
    base = width * ((offsets + nmj) // width) - offsets
    begin = np.clip(base, 0, n_nodes)
    end = np.clip(base + min(3, width), 0, n_nodes)
    return begin.astype(np.int32), end.astype(np.int32)


def firing_table(n_rows, n_mu, rate_per_ms, dt_row, seed):
    """0/1 table, rows = sampling steps of width dt_row, columns = motor units (Poisson firing)."""
    rng = np.random.default_rng(seed)
    return (rng.random((n_rows, n_mu)) < rate_per_ms * dt_row).astype(np.uint8)


def poisson_firing_table(n_rows, n_mu, rate_per_ms, dt_row, duration_ms, seed, first_at_zero=False):
    """Firing table with one row per sampling step: each motor unit fires at
    Poisson times (rate_per_ms, a scalar or one rate per motor unit) and each
    firing marks duration_ms of rows; first_at_zero adds a firing of every unit
    at t=0."""
    rng = np.random.default_rng(seed)
    rates = np.broadcast_to(np.asarray(rate_per_ms, np.float64), (n_mu,))
    table = np.zeros((n_rows, n_mu), np.uint8)
    t_end = n_rows * dt_row
    width = max(1, int(round(duration_ms / dt_row)))
    for mu in range(n_mu):
        t = 0.0 if first_at_zero else rng.exponential(1.0 / rates[mu])
        while t < t_end:
            r0 = int(t / dt_row)
            table[r0:r0 + width, mu] = 1
            t += rng.exponential(1.0 / rates[mu])
    return table


def c2(n_fibers=1024, n_nodes=1481, length_cm=14.8, dt=2.5e-5, end_time=5.0, n_mu=10, rate_hz=20.0,
       seed_mu=1, seed_fire=2, lane_width=4):
    """57-state-class Shorten workload: equal fibres, IE, motor units with Poisson
    firing at the centre node; stimulation follows the reference's firing-table
    semantics with the table sampled once per split step."""
    n = np.full(n_fibers, n_nodes, np.int32)
    h = np.full(n_fibers * (n_nodes - 1), length_cm / (n_nodes - 1))
    pf = np.full(n_fibers, SIGMA / (AM * CM))
    mu = np.random.default_rng(seed_mu).integers(1, n_mu + 1, size=n_fibers).astype(np.int32)
    rows = int(round(end_time / dt))
    table = poisson_firing_table(rows, n_mu, rate_hz * 1e-3, dt, 0.1, seed_fire)
    off = np.arange(n_fibers, dtype=np.int64) * n_nodes
    sb, se = vc_lane_stim_range(off, np.full(n_fibers, n_nodes // 2), n, lane_width)
    stim = dict(stim_begin=sb, stim_end=se, firing_events=table, motor_unit=mu, call_frequency=1.0 / dt,
                repeat_after_first_call=0.0, call_enable_begin=0.0)
    return dict(n_nodes=n, element_lengths=h, prefactor=pf), stim


def c5(n_fibers=16384, lo=1000, hi=1481, n_mu=100, rate_hz=(10.0, 40.0), dt=2.5e-5, n_rows=4000, seed=5,
       lane_width=4, first_at_zero=True):
    """Large Shorten-type batch: fibres of U{lo..hi} nodes (h = 0.01 cm), motor
    units drawn uniformly, each firing at seeded Poisson times with its own rate
    ~ U(rate_hz) at the fibre's centre node; the firing table has one row per
    split step (n_rows rows, wrapping as the reference's table does)."""
    rng = np.random.default_rng(seed)
    n = rng.integers(lo, hi + 1, size=n_fibers).astype(np.int32)
    h = np.full(int(n.sum()) - n_fibers, 0.01)
    pf = np.full(n_fibers, SIGMA / (AM * CM))
    mu = np.random.default_rng(seed + 1).integers(1, n_mu + 1, size=n_fibers).astype(np.int32)
    rates = np.random.default_rng(seed + 2).uniform(rate_hz[0], rate_hz[1], size=n_mu) * 1e-3
    table = poisson_firing_table(n_rows, n_mu, rates, dt, 0.1, seed + 3, first_at_zero=first_at_zero)
    off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    sb, se = vc_lane_stim_range(off, n // 2, n, lane_width)
    stim = dict(stim_begin=sb, stim_end=se, firing_events=table, motor_unit=mu, call_frequency=1.0 / dt,
                repeat_after_first_call=0.0, call_enable_begin=0.0)
    return dict(n_nodes=n, element_lengths=h, prefactor=pf), stim
