"""Checks that pin the oracle's TK2014 model (57 states, the reference's
hand-written Vc kernel FMS/tk2014.cpp.disabled), which has no known answer in
the reference's tests. Each check is one the mathematics or a pinned model
fixes, so that a dropped term, a wrong sign or a wrong index in the
transcription fails it:

* shared equations: the rates of the states whose equations and constants the
  kernel shares with the Shorten 2007 model equal that model's rates bit for bit
  at random states; the Shorten model is pinned by the reference's known answer
  (testing/unit_testing/src/1_rank/cellml.cpp:215, tests/test_oracle.py);
* conservation: total calcium (free, buffered by troponin, ATP, calsequestrin,
  cross-bridge troponin sites and Ca-phosphate precipitate, over the four
  compartments), total magnesium, total ATP and the ryanodine-receptor
  occupancy are constants of the motion: sum_k w_k dy_k/dt = 0 with the
  compartment volumes (1 % / 99 % splits of the 95 % myoplasm and the 5 % SR)
  and the binding stoichiometry as weights; the same laws hold for the Shorten
  model;
* positivity: a probability or concentration at 0 has a non-negative rate, and
  a gate at 1 a non-positive one.

TK2014 state k is Shorten state k for k < 28 and k - 1 for k > 28 (state 28,
the activation integral, exists in TK2014 only). The rates of states 0-7 (the
membrane and t-tubule potentials and the ion concentrations) and 28 are pinned
by none of these checks.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle_binding import ORACLE_BUILD, P, _dp, model_info, oracle_library

SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"

# compartment volumes relative to the fibre: myoplasm near the triads / bulk
# myoplasm (1 % / 99 % of 95 %), terminal / longitudinal SR (1 % / 99 % of 5 %)
V_C1, V_C2, V_S1, V_S2 = 0.01 * 0.95, 0.99 * 0.95, 0.01 * 0.05, 0.99 * 0.05


def _libs():
    for m in (SHORTEN, "tk2014"):
        if not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{m}.so")):
            pytest.skip(f"oracle model {m} not built (needs the reference checkout at build time)")
    sh, tk = oracle_library(SHORTEN), oracle_library("tk2014")
    for lib in (sh, tk):
        lib.fmso_rhs.argtypes = [_dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_int]
    return sh, tk


def _rhs(lib, y, exp_mode=1):
    ns, na, npar = model_info(lib)
    p = np.zeros(max(npar, 1))
    lib.fmso_default_params(P(p, _dp))
    r, a = np.zeros(ns), np.zeros(na)
    lib.fmso_rhs(P(np.ascontiguousarray(y, np.float64), _dp), P(p, _dp), 0.0, P(r, _dp), P(a, _dp), exp_mode)
    return r


def _random_states(lib, n, seed):
    ns = model_info(lib)[0]
    y0 = np.zeros(ns)
    lib.fmso_initial_states(P(y0, _dp))
    rng = np.random.default_rng(seed)
    for _ in range(n):
        y = y0 * rng.uniform(0.5, 1.5, ns) + rng.uniform(0.0, 0.5, ns) * (y0 == 0)
        y[0], y[1] = rng.uniform(-100, 40, 2)
        yield y


def _to_shorten(k):
    return k if k < 28 else k - 1


def conservation_residuals(r, idx=lambda k: k):
    """relative residuals sum_k w_k r_k / sum_k |w_k r_k| of the four laws; r in
    the model's own indexing, idx maps a TK2014 state number to it"""
    def law(terms):
        v = np.array([w * r[idx(k)] for w, k in terms])
        return abs(v.sum()) / max(np.abs(v).sum(), 1e-300)
    calcium = ([(V_C1, k) for k in (29, 34, 40)]
               + [(V_C2, k) for k in (31, 35, 41, 33, 50)] + [(2 * V_C2, k) for k in (48, 51, 52, 53)]
               + [(V_S1, k) for k in (30, 38)] + [(V_S2, k) for k in (32, 39)] + [(1000 * V_S2, 56)])
    magnesium = [(V_C1, k) for k in (46, 42, 36)] + [(V_C2, k) for k in (47, 43, 37)]
    atp = [(V_C1, k) for k in (40, 42, 44)] + [(V_C2, k) for k in (41, 43, 45)]
    ryr = [(1.0, k) for k in range(18, 28)]
    return dict(calcium=law(calcium), magnesium=law(magnesium), atp=law(atp), ryanodine_receptor=law(ryr))


SHARED_WITH_SHORTEN = [8, 9, 10, 11, 13, 14, 15, 16, 18, 19, 20, 21, 23, 24, 25, 26, 38, 39, 56]


def test_tk2014_rates_shared_with_shorten_equal_bitwise():
    sh, tk = _libs()
    for mode in (0, 1):
        for y in _random_states(tk, 200, seed=2):
            rt = _rhs(tk, y, mode)
            rs = _rhs(sh, np.delete(y, 28), mode)
            for k in SHARED_WITH_SHORTEN:
                assert rt[k] == rs[_to_shorten(k)], (mode, k, rt[k], rs[_to_shorten(k)])


@pytest.mark.parametrize("model", [SHORTEN, "tk2014"])
def test_calcium_magnesium_atp_and_receptor_conservation(model):
    sh, tk = _libs()
    lib, idx = (tk, lambda k: k) if model == "tk2014" else (sh, _to_shorten)
    worst = {}
    for y in _random_states(lib, 200, seed=0):
        for name, res in conservation_residuals(_rhs(lib, y), idx).items():
            worst[name] = max(worst.get(name, 0.0), res)
    print(model, worst)
    assert max(worst.values()) <= 1e-14, worst


def test_conservation_detects_a_dropped_term():
    """the calcium law fails when one flux loses its counterpart: drop the
    myoplasmic diffusion term from the bulk compartment's rate (TK2014 line
    rate31: + constant64*((states[29] - states[31])/constant109))"""
    _, tk = _libs()
    y = next(_random_states(tk, 1, seed=5))
    r = _rhs(tk, y)
    assert conservation_residuals(r)["calcium"] <= 1e-14
    # constant64 = 0.75, constant109 = 0.99 * 0.95 * pi * 0.5^2 (the kernel's
    # constants 64, 66, 68, 106, 109)
    r_bad = r.copy()
    r_bad[31] -= 0.75 * (y[29] - y[31]) / (V_C2 * np.pi * 0.5 ** 2)
    assert conservation_residuals(r_bad)["calcium"] > 1e-11   # far above the 1e-14 bar


@pytest.mark.parametrize("model", [SHORTEN, "tk2014"])
def test_positivity_and_gate_bounds(model):
    sh, tk = _libs()
    lib, idx = (tk, lambda k: k) if model == "tk2014" else (sh, _to_shorten)
    states = [k for k in list(range(8, 28)) + list(range(29, 57))]
    for y in _random_states(lib, 100, seed=1):
        for k in states:
            for value, sign in ((0.0, 1.0), (1.0, -1.0)):
                if value == 1.0 and not 8 <= k <= 17:
                    continue
                yy = y.copy()
                yy[idx(k)] = value
                r = _rhs(lib, yy)
                assert sign * r[idx(k)] >= -1e-12 * (1.0 + np.abs(r).max()), (model, k, value, r[idx(k)])
