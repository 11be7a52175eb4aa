"""Settings-dict front end of the MI355X fast fibre solver.

Accepts the reference's settings for
``RepeatedCall<FastMonodomainSolver<MultipleInstances<StrangSplitting<
MultipleInstances<Heun<CellmlAdapter>>, MultipleInstances<ImplicitEuler|CrankNicolson<FEM>>>>>>``
(examples/electrophysiology/monodomain/hodgkin_huxley_fast/settings_fast_fiber.py:77-220,
doc/sphinx/settings/fast_monodomain_solver.rst:59-208) and runs them on the GPU
through the C ABI. Fibres are the Heun/diffusion instance pairs of all outer
instances, in order; each Heun instance's CellML block and each diffusion
instance's FiniteElementMethod block describe that fibre.

Keys read (defaults as in the reference,
core/src/specialized_solver/fast_monodomain_solver/fast_monodomain_solver_initialization.tpp:66-72):
  RepeatedCall: timeStepWidth, endTime, fiberDistributionFile, firingTimesFile,
    onlyComputeIfHasBeenStimulated (True), disableComputationWhenStatesAreCloseToEquilibrium
    (True; not applied, a warning says so), valueForStimulatedPoint (20),
    neuromuscularJunctionRelativeSize (0)
  StrangSplitting: timeStepWidth, endTime
  Heun: timeStepWidth; CellML: modelFilename, optimizationType, approximateExponentialFunction,
    statesInitialValues, setSpecificStatesCallFrequency, setSpecificStatesFrequencyJitter,
    setSpecificStatesRepeatAfterFirstCall, setSpecificStatesCallEnableBegin, additionalArgument,
    parametersUsedAsAlgebraic, parametersUsedAsConstant, parametersInitialValues,
    algebraicsForTransfer, stimulationLogFilename, meshName, libraryFilename
  ImplicitEuler | CrankNicolson: timeStepWidth; FiniteElementMethod: meshName, prefactor
  Meshes.<name>: nElements, physicalExtent, physicalOffset | nodePositions
Extension keys (ours): "vcVectorWidth" (stimulated lanes of the vc path, default 4),
"neuromuscularJunctionSeed" (the reference draws NMJ positions unseeded),
"kernels" ("fused" | "split"), "thomas" ("partitioned" | "serial"), "precision"
("fp64" | "mixed").
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from . import build as B
from .fibers import FiberBatch
from .settings import (element_lengths, gpu_stimulation_range, node_positions, parse_fiber_distribution,
                       parse_firing_times, vc_stimulation_range, write_stimulation_log)

HH_CONSTANT_PARAMETERS = [2, 3, 4, 5]   # i_Stim, g_Na, g_K, g_L of hodgkin_huxley_1952


def _list(v):
    if v is None:
        return []
    return list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v]


def _read(path):
    with open(path) as f:
        return f.read()


class FastMonodomainSolver:
    """settings dict -> fibres on the GPU; ``run()`` = initialize() + advanceTimeSpan()"""

    def __init__(self, config: dict, base_dir: str | None = None):
        self.config = config
        self.base_dir = base_dir or os.getcwd()
        self.rc = config["RepeatedCall"]
        self.initialized = False
        self.current_time = 0.0

    # -- settings -----------------------------------------------------------------
    def _path(self, p):
        return p if os.path.isabs(p) else os.path.join(self.base_dir, p)

    def _fibers(self):
        outer = self.rc["MultipleInstances"]["instances"]
        fibres = []
        for inst in outer:
            strang = inst["StrangSplitting"]
            t1 = strang["Term1"]["MultipleInstances"]["instances"]
            t2 = strang["Term2"]["MultipleInstances"]["instances"]
            if len(t1) != len(t2):
                raise ValueError("Term1 and Term2 must have the same number of instances")
            for a, b in zip(t1, t2):
                heun = a["Heun"]
                key = "ImplicitEuler" if "ImplicitEuler" in b else "CrankNicolson" if "CrankNicolson" in b else None
                if key is None:
                    raise ValueError("Timestepping scheme of diffusion in FastMonodomainSolver must be either "
                                     "ImplicitEuler or CrankNicolson")
                fibres.append((strang, heun, key, b[key]))
        return fibres

    def initialize(self):
        if self.initialized:
            return
        cfg, rc = self.config, self.rc
        fibres = self._fibers()
        if not fibres:
            raise ValueError("no fibres in the settings")
        strang0, heun0, scheme_key, diff0 = fibres[0]
        cellml0 = heun0["CellML"]
        schemes = {k for _, _, k, _ in fibres}
        if len(schemes) != 1:
            raise ValueError("all fibres must use the same diffusion scheme")
        self.scheme = "IE" if scheme_key == "ImplicitEuler" else "CN"

        # model library
        model_file = cellml0.get("modelFilename", "")
        base = os.path.splitext(os.path.basename(model_file))[0]
        p_alg = _list(cellml0.get("parametersUsedAsAlgebraic", []))
        p_const = _list(cellml0.get("parametersUsedAsConstant", []))
        p_init = [float(v) for v in _list(cellml0.get("parametersInitialValues", []))]
        library = cellml0.get("libraryFilename")
        if library is None:
            library = self._model_library(model_file, base, p_alg, p_const)
        self.library = library

        # fibres: meshes, element lengths, prefactors
        meshes = cfg.get("Meshes", {})
        n_nodes, lengths, prefactors, fiber_no = [], [], [], []
        for k, (_, heun, _, diff) in enumerate(fibres):
            fem = diff["FiniteElementMethod"]
            mesh_name = fem.get("meshName", heun["CellML"].get("meshName"))
            pos = node_positions(meshes[mesh_name])
            n_nodes.append(len(pos))
            lengths.append(element_lengths(pos))
            prefactors.append(float(fem.get("prefactor", 1.0)))
            fiber_no.append(int(heun["CellML"].get("additionalArgument", k) or 0))
        self.n_nodes = np.array(n_nodes, np.int32)
        self.fiber_no = np.array(fiber_no, np.int32)
        F = len(fibres)

        # parameters per fibre
        self._batch_kw = dict(
            scheme=self.scheme,
            precision=rc.get("precision", "fp64"),
            exp_mode="approx" if cellml0.get("approximateExponentialFunction", True) else "libm",
            thomas=rc.get("thomas", "partitioned"),
            kernels=rc.get("kernels", "fused"),
            value_for_stimulated_point=float(rc.get("valueForStimulatedPoint", 20.0)),
            library=library,
        )
        y0 = cellml0.get("statesInitialValues")
        from . import _lib as L
        lib = L.load_library(library)
        n_states, _, n_params, _ = L.model_info(lib)
        defaults = np.zeros(max(n_params, 1))
        L.check(lib, lib.fms_model_defaults(None, L.as_ptr(defaults, L._dp)), "fms_model_defaults")
        params = np.tile(defaults[:n_params], (F, 1))
        slots = self._parameter_slots(base, p_alg, p_const)
        for j, v in enumerate(p_init):
            if j < len(slots) and slots[j] is not None:
                params[:, slots[j]] = v
        self.batch = FiberBatch(self.n_nodes, np.concatenate(lengths), np.array(prefactors), parameters=params,
                                initial_states=(None if not y0 else np.asarray(y0, float)), **self._batch_kw)
        self.parameters = params

        # stimulation from the motor-unit files
        self.stim = None
        if rc.get("firingTimesFile") and rc.get("fiberDistributionFile"):
            table = parse_firing_times(_read(self._path(rc["firingTimesFile"])))
            mu_list = parse_fiber_distribution(_read(self._path(rc["fiberDistributionFile"])))
            mus = mu_list[self.fiber_no % len(mu_list)]
            rel = float(rc.get("neuromuscularJunctionRelativeSize", 0.0))
            if rel <= 0.0:
                nmj = self.n_nodes // 2
            else:
                rng = np.random.default_rng(rc.get("neuromuscularJunctionSeed"))
                nmj = (self.n_nodes * rng.uniform(0.5 - rel / 2, 0.5 + rel / 2, F)).astype(np.int32)
            offsets = np.concatenate([[0], np.cumsum(self.n_nodes)[:-1]])
            if cellml0.get("optimizationType", "vc") == "gpu":
                rng_ = [gpu_stimulation_range(int(m), int(n)) for m, n in zip(nmj, self.n_nodes)]
            else:
                w = int(rc.get("vcVectorWidth", 4))
                rng_ = [vc_stimulation_range(int(o), int(m), int(n), w) for o, m, n in zip(offsets, nmj, self.n_nodes)]
            sb = np.array([r[0] for r in rng_], np.int32)
            se = np.array([r[1] for r in rng_], np.int32)
            freq, jit, rep, beg = [], [], [], []
            for _, heun, _, _ in fibres:
                c = heun["CellML"]
                freq.append(float(c.get("setSpecificStatesCallFrequency", 0.0)))
                j = c.get("setSpecificStatesFrequencyJitter", [])
                jit.append([float(x) for x in _list(j)])
                rep.append(float(c.get("setSpecificStatesRepeatAfterFirstCall", 0.0)))
                beg.append(float(c.get("setSpecificStatesCallEnableBegin", 0.0)))
            self.motor_units = mus
            self.batch.set_firing_table_stimulation(sb, se, table, mus, np.array(freq), np.array(rep),
                                                    np.array(beg), jitter_lists=jit, log_capacity=1 << 16)
            self.stim = dict(stim_begin=sb, stim_end=se, firing_events=table, motor_unit=mus)
        if rc.get("onlyComputeIfHasBeenStimulated", True):
            self.batch.set_only_compute_if_stimulated(True)
        if rc.get("disableComputationWhenStatesAreCloseToEquilibrium", True):
            warnings.warn("disableComputationWhenStatesAreCloseToEquilibrium is not applied: every node is "
                          "computed every step (set it to False to silence this warning)", stacklevel=2)
        alg = _list(cellml0.get("algebraicsForTransfer", []))
        if alg:
            self.batch.set_algebraics_for_transfer(alg)
        self.stimulation_log_file = cellml0.get("stimulationLogFilename", "out/stimulation.log")

        # time stepping
        from . import _lib as L
        sched = L.schedule(self.batch.lib, float(rc["endTime"]), float(rc["timeStepWidth"]),
                           float(strang0.get("endTime", rc["timeStepWidth"])), float(strang0["timeStepWidth"]),
                           0.5 * float(strang0["timeStepWidth"]), float(heun0["timeStepWidth"]),
                           float(strang0["timeStepWidth"]), float(diff0["timeStepWidth"]))
        self.schedule = sched
        self.initialized = True

    def _model_library(self, model_file, base, p_alg, p_const):
        if base == "hodgkin_huxley_1952":
            if p_alg or any(k not in HH_CONSTANT_PARAMETERS for k in p_const):
                raise ValueError("hodgkin_huxley_1952: only CONSTANTS[2..5] can be parameters")
            return B.HH_LIBRARY
        lib = f"libfms_{base}.so"
        if os.path.exists(os.path.join(B.LIB_DIR, lib)):
            return lib
        path = self._path(model_file)
        if not os.path.exists(path):
            raise RuntimeError(f"no library for model {base!r} and model file {path} not found")
        return os.path.basename(B.build_model_library(path, p_alg, p_const, name=base))

    def _parameter_slots(self, base, p_alg, p_const):
        """library parameter index of each settings parameter (parametersUsedAsAlgebraic
        first, then parametersUsedAsConstant, 00_parse_source_code_file.cpp:284-333)"""
        if base == "hodgkin_huxley_1952":
            return [None] * len(p_alg) + [HH_CONSTANT_PARAMETERS.index(k) for k in p_const]
        return list(range(len(p_alg) + len(p_const)))

    # -- run ------------------------------------------------------------------------
    def advance_time_span(self):
        """all RepeatedCall calls; each is one computeMonodomain with its own widths"""
        self.initialize()
        s = self.schedule
        n0 = np.unique(s[:, 4])
        if len(n0) != 1:
            raise ValueError("the number of Heun steps per half step changes between calls")
        rows = []
        for t0, dts, ns, dt0, _, dt1, _ in s:
            for k in range(int(ns)):
                rows.append((t0 + k * dts, dts, dt0, dt1))
        self.batch.advance_schedule(np.array(rows), int(n0[0]))
        self.batch.synchronize()
        self.current_time = float(s[-1, 0] + s[-1, 1] * s[-1, 2])

    def run(self):
        self.initialize()
        self.advance_time_span()
        if self.stim is not None and self.stimulation_log_file:
            self.write_stimulation_log(self._path(self.stimulation_log_file))

    # -- results ----------------------------------------------------------------------
    def vm(self) -> list[np.ndarray]:
        v = self.batch.vm()
        off = np.concatenate([[0], np.cumsum(self.n_nodes)])
        return [v[off[f]:off[f + 1]] for f in range(len(self.n_nodes))]

    def write_stimulation_log(self, path):
        times, fibers = self.batch.stimulation_log()
        if len(times) == 0:
            return
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        write_stimulation_log(path, times, self.fiber_no[fibers], self.motor_units[fibers])
