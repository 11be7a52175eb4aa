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
