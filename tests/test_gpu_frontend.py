"""The settings-dict front end on the GPU against the oracle.

The settings follow examples/electrophysiology/monodomain/hodgkin_huxley_fast/
settings_fast_fiber.py (RepeatedCall at dt_splitting, Strang with one Heun of
5 sub-steps per half and one implicit-Euler step) with three fibres and the
reference's motor-unit files (tests/golden).
"""
import os

import numpy as np
import pytest

from oracle_binding import OracleBatch
from opendihu_amd.settings import vc_stimulation_range

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def settings(n_fibers=3, n_elements=100, end_time=4.0, dt_split=2e-3, dt_0D=2e-4, only_if=False, scheme="ImplicitEuler"):
    meshes = {f"MeshFiber_{i}": {"nElements": [n_elements + 10 * i], "physicalExtent": [(n_elements + 10 * i) / 100.],
                                 "physicalOffset": [0]} for i in range(n_fibers)}
    heuns, diffs = [], []
    for i in range(n_fibers):
        heuns.append({"ranks": [0], "Heun": {"timeStepWidth": dt_0D, "CellML": {
            "modelFilename": "hodgkin_huxley_1952.c", "optimizationType": "vc",
            "approximateExponentialFunction": True,
            "setSpecificStatesCallFrequency": 0.1, "setSpecificStatesFrequencyJitter": [0.1, -0.1],
            "setSpecificStatesRepeatAfterFirstCall": 0.1, "setSpecificStatesCallEnableBegin": 0.5 * i,
            "additionalArgument": i, "parametersUsedAsAlgebraic": [], "parametersUsedAsConstant": [2],
            "parametersInitialValues": [0.0], "meshName": f"MeshFiber_{i}", "stimulationLogFilename": ""}}})
        diffs.append({"ranks": [0], scheme: {"timeStepWidth": dt_split, "FiniteElementMethod": {
            "meshName": f"MeshFiber_{i}", "prefactor": 0.03}}})
    return {
        "Meshes": meshes,
        "RepeatedCall": {
            "timeStepWidth": dt_split, "endTime": end_time,
            "MultipleInstances": {"nInstances": 1, "instances": [{"ranks": [0], "StrangSplitting": {
                "timeStepWidth": dt_split, "endTime": dt_split,
                "Term1": {"MultipleInstances": {"nInstances": n_fibers, "instances": heuns}},
                "Term2": {"MultipleInstances": {"nInstances": n_fibers, "instances": diffs}}}}]},
            "fiberDistributionFile": os.path.join(GOLDEN, "MU_fibre_distribution_10MUs.txt"),
            "firingTimesFile": os.path.join(GOLDEN, "MU_firing_times_always.txt"),
            "onlyComputeIfHasBeenStimulated": only_if,
            "disableComputationWhenStatesAreCloseToEquilibrium": False,
            "valueForStimulatedPoint": 20,
            "thomas": "serial",
        },
    }


@pytest.mark.parametrize("only_if", [False, True])
def test_settings_front_end_matches_oracle(only_if):
    from opendihu_amd.fast_monodomain_solver import FastMonodomainSolver
    cfg = settings(only_if=only_if)
    s = FastMonodomainSolver(cfg)
    s.initialize()
    assert np.all(s.schedule[:, 4] == 5) and s.schedule.shape[0] == 2000
    # "kernels" defaults to "auto": the autotuner chose among bit-identical schedules
    assert s.tuned_schedule is not None and s.tuned_schedule["rejected"] == 0, s.tuned_schedule
    s.advance_time_span()
    n = s.n_nodes
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    sb, se = zip(*[vc_stimulation_range(int(o), int(k) // 2, int(k), 4) for o, k in zip(off, n)])
    o = OracleBatch(n_nodes=n, element_lengths=s.batch.element_lengths, prefactor=0.03, scheme="IE")
    o.set_firing_table_stimulation(np.array(sb), np.array(se), s.stim["firing_events"], s.stim["motor_unit"],
                                   call_frequency=0.1, repeat_after_first_call=0.1,
                                   call_enable_begin=0.5 * np.arange(3), jitter_lists=[[0.1, -0.1]] * 3)
    o.set_only_compute_if_stimulated(only_if)
    for t0, dts, ns, dt0, n0, dt1, n1 in s.schedule:
        o.advance(t0, dts, int(ns), dt0, int(n0), dt1, int(n1))
    np.testing.assert_array_equal(s.batch.vm(), o.vm())
    assert np.max(o.vm()) > 0, "the fibres fire"


@pytest.mark.parametrize("model", ["hh", "shorten_ocallaghan_davidson_soboleva_2007"])
def test_equilibrium_states_match_oracle(model):
    """fms_equilibrium_states (initializeStatesToEquilibrium) against the
    oracle's RK4 search from the model's initial values: same iterations, same
    bits. HH converges (436094 iterations at dt = 1e-4) to a rest state whose
    rates vanish to the stopping bar; the Shorten model's search does not
    converge within 200000 iterations, and both stop with the same last iterate
    and increment (the library reports FMS_ERR_STATE)"""
    import ctypes
    from opendihu_amd import _lib as L
    from opendihu_amd.build import LIB_DIR
    from oracle_binding import ORACLE_BUILD, equilibrium, oracle_library
    libname = "libfms_hodgkin_huxley_1952.so" if model == "hh" else f"libfms_{model}.so"
    if not os.path.exists(os.path.join(LIB_DIR, libname)) or \
            not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{model}.so")):
        pytest.skip(f"{model} not built")
    lib = L.load_library(libname)
    ns, _, npar, _ = L.model_info(lib)
    y0, p0 = np.zeros(ns), np.zeros(max(npar, 1))
    L.check(lib, lib.fms_model_defaults(L.as_ptr(y0, L._dp), L.as_ptr(p0, L._dp)), "defaults")
    max_it = 3_000_000 if model == "hh" else 200_000
    y = y0.copy()
    it, inc, dt_out = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
    rc = lib.fms_equilibrium_states(L.as_ptr(y, L._dp), L.as_ptr(p0, L._dp), 1e-4, max_it, ctypes.byref(it),
                                    ctypes.byref(inc), ctypes.byref(dt_out))
    ref_y, ref_it, ref_inc, ref_dt = equilibrium(oracle_library(model), y0, p0[:npar], dt=1e-4, max_iter=max_it)
    print(model, rc, it.value, ref_it, inc.value, y[:4])
    np.testing.assert_array_equal(y, ref_y)
    assert inc.value == ref_inc and dt_out.value == ref_dt
    if model == "hh":
        assert rc == 0 and it.value == ref_it == 436094
        rates = L.eval_rhs(lib, y[:, None], p0[:, None], exp_mode="libm")
        assert np.max(np.abs(rates)) < 1e-4
        # the gates at their steady state alpha/(alpha+beta) of the rest voltage,
        # to the stopping bar (|dy/dt| < 1e-5 per ms) times the slowest gate's
        # time constant (h: about 8 ms at rest)
        gates = cases_steady_state(y[0])
        np.testing.assert_allclose(y[1:], gates, rtol=0, atol=2e-4)
    else:
        assert rc == L.FMS_ERR_STATE and ref_it == -1 and it.value == max_it
        assert "not found" in L.last_error(lib)


def cases_steady_state(v):
    from cases import hh_steady_state_gates
    return np.array(hh_steady_state_gates(v))


def test_equilibrium_search_reports_non_finite_states():
    """a state that turns NaN never raises the search's maximum increment; the
    library reports it instead of returning the NaN states as converged"""
    import ctypes
    from opendihu_amd import _lib as L
    lib = L.load_library("libfms_hodgkin_huxley_1952.so")
    ns, _, npar, _ = L.model_info(lib)
    y, p0 = np.zeros(ns), np.zeros(max(npar, 1))
    L.check(lib, lib.fms_model_defaults(L.as_ptr(y, L._dp), L.as_ptr(p0, L._dp)), "defaults")
    y[2] = np.nan   # h: after one RK4 step every state is NaN and every |increment| compares false
    it, inc, dt_out = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
    rc = lib.fms_equilibrium_states(L.as_ptr(y, L._dp), L.as_ptr(p0, L._dp), 1e-4, 1000, ctypes.byref(it),
                                    ctypes.byref(inc), ctypes.byref(dt_out))
    assert rc == L.FMS_ERR_STATE
    assert "non-finite" in L.last_error(lib)


def test_front_end_equilibrium_output_and_transfer(tmp_path):
    """initializeStatesToEquilibrium, statesForTransfer / algebraicsForTransfer
    and PythonFile output through the settings front end: the run starts from
    the equilibrium states (as the oracle run does), the output files hold the
    mesh and Vm of every output step and compare to an identical run with
    error 0 under the reference's cmp metric"""
    from opendihu_amd import _lib as L
    from opendihu_amd.fast_monodomain_solver import FastMonodomainSolver
    from opendihu_amd.output import compare_directories, read_file, solution_values
    runs = []
    for r in range(2):
        cfg = settings(n_fibers=2, end_time=1.0)
        for inst in cfg["RepeatedCall"]["MultipleInstances"]["instances"][0]["StrangSplitting"]["Term1"][
                "MultipleInstances"]["instances"]:
            c = inst["Heun"]["CellML"]
            c["initializeStatesToEquilibrium"] = True
            c["statesForTransfer"] = [0, 2]
            c["algebraicsForTransfer"] = [3]
        strang = cfg["RepeatedCall"]["MultipleInstances"]["instances"][0]["StrangSplitting"]
        strang["OutputWriter"] = [{"format": "PythonFile", "outputInterval": 100, "binary": r == 1,
                                   "filename": str(tmp_path / f"run{r}" / "fibers"), "fileNumbering": "incremental"}]
        s = FastMonodomainSolver(cfg)
        s.run()
        runs.append(s)
    s = runs[0]
    eq = s.equilibrium
    assert eq["max_increment"] < 1e-5
    tr = s.transfer()
    assert sorted(tr["states"]) == [0, 2] and sorted(tr["algebraics"]) == [3]
    np.testing.assert_array_equal(tr["states"][0][1], s.vm()[1])
    files = s._writers()[0].files
    assert len(files) == 2 * (1 + 500 // 100)
    d = read_file(files[-1])
    for key in ("meshType", "dimension", "nElementsGlobal", "nElementsLocal", "beginNodeGlobalNatural",
                "hasFullNumberOfNodes", "basisFunction", "basisOrder", "onlyNodalValues", "nRanks", "ownRankNo",
                "data", "timeStepNo", "currentTime"):
        assert key in d
    np.testing.assert_array_equal(solution_values(d), s.vm()[1])
    d0 = read_file(files[0])
    np.testing.assert_array_equal(solution_values(d0), np.full(s.n_nodes[0], eq["states"][0]))
    err, per_file = compare_directories(str(tmp_path / "run0"), str(tmp_path / "run1"))
    assert err == 0.0 and len(per_file) == len(files)
    # the oracle from the same equilibrium states
    n = s.n_nodes
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    sb, se = zip(*[vc_stimulation_range(int(o), int(k) // 2, int(k), 4) for o, k in zip(off, n)])
    o = OracleBatch(n_nodes=n, element_lengths=s.batch.element_lengths, prefactor=0.03, scheme="IE",
                    initial_states=eq["states"])
    o.set_firing_table_stimulation(np.array(sb), np.array(se), s.stim["firing_events"], s.stim["motor_unit"],
                                   call_frequency=0.1, repeat_after_first_call=0.1,
                                   call_enable_begin=0.5 * np.arange(2), jitter_lists=[[0.1, -0.1]] * 2)
    for t0, dts, ns, dt0, n0, dt1, n1 in s.schedule:
        o.advance(t0, dts, int(ns), dt0, int(n0), dt1, int(n1))
    np.testing.assert_array_equal(s.batch.vm(), o.vm())
    L.check(s.batch.lib, 0, "")


def test_front_end_warns_on_keys_it_does_not_apply():
    from opendihu_amd.fast_monodomain_solver import FastMonodomainSolver
    cfg = settings(n_fibers=1, end_time=0.02)
    cfg["RepeatedCall"]["useSinglePrecision"] = True
    c = cfg["RepeatedCall"]["MultipleInstances"]["instances"][0]["StrangSplitting"]["Term1"]["MultipleInstances"][
        "instances"][0]["Heun"]["CellML"]
    c["mappings"] = {("parameter", 0): "membrane/i_Stim"}
    c["someUnknownKey"] = 1
    c["compilerFlags"] = "-O3"          # no effect here: accepted silently
    with pytest.warns(UserWarning) as rec:
        s = FastMonodomainSolver(cfg)
        s.initialize()
    msgs = [str(w.message) for w in rec]
    assert any("CellML.mappings" in m for m in msgs) and any("CellML.someUnknownKey" in m for m in msgs)
    assert any("useSinglePrecision" in m for m in msgs) and not any("compilerFlags" in m for m in msgs)
    assert s.batch.config["precision"] == "mixed"
