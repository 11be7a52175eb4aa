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
  CellML: initializeStatesToEquilibrium (False), initializeStatesToEquilibriumTimestepWidth (1e-4),
    statesForTransfer ([0]), useSinglePrecision at RepeatedCall level (gpu back end: runs the
    mixed form, FP32 gates with FP64 Vm and diffusion; a warning says so)
  StrangSplitting.OutputWriter: {"format": "PythonFile", "outputInterval", "filename", "binary",
    "fileNumbering"} (opendihu_amd.output)
Keys that change nothing here (log keys, callbacks the fast solver does not call,
boundary conditions of the nested solvers) are accepted silently; every other key
this front end does not apply raises a warning that names it.
Extension keys (ours): "vcVectorWidth" (stimulated lanes of the vc path, default 4),
"neuromuscularJunctionSeed" (the reference draws NMJ positions unseeded),
"kernels" ("auto": the autotuner's fastest bit-identical schedule, timed on the run's first
split steps at initialisation | "fused" | "split"), "thomas" ("partitioned" | "serial"), "precision"
("fp64" | "mixed"), "pointBufferOrder" (True: onlyComputeIfHasBeenStimulated in the
reference's sweep over buffers of vcVectorWidth nodes, compute.tpp:131-216).
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

# keys per settings block: applied here, or without effect on this path (logging,
# callbacks and nested-solver options the fast solver does not use)
_APPLIED = {
    "RepeatedCall": {"timeStepWidth", "endTime", "MultipleInstances", "fiberDistributionFile", "firingTimesFile",
                     "onlyComputeIfHasBeenStimulated", "disableComputationWhenStatesAreCloseToEquilibrium",
                     "valueForStimulatedPoint", "neuromuscularJunctionRelativeSize", "useSinglePrecision",
                     "vcVectorWidth", "neuromuscularJunctionSeed", "kernels", "thomas", "precision",
                     "pointBufferOrder"},
    "StrangSplitting": {"timeStepWidth", "endTime", "Term1", "Term2", "OutputWriter"},
    "Heun": {"timeStepWidth", "CellML"},
    "CellML": {"modelFilename", "optimizationType", "approximateExponentialFunction", "statesInitialValues",
               "initializeStatesToEquilibrium", "initializeStatesToEquilibriumTimestepWidth",
               "setSpecificStatesCallFrequency", "setSpecificStatesFrequencyJitter",
               "setSpecificStatesRepeatAfterFirstCall", "setSpecificStatesCallEnableBegin", "additionalArgument",
               "parametersUsedAsAlgebraic", "parametersUsedAsConstant", "parametersInitialValues",
               "algebraicsForTransfer", "statesForTransfer", "stimulationLogFilename", "meshName", "libraryFilename"},
    "Diffusion": {"timeStepWidth", "FiniteElementMethod"},
    "FiniteElementMethod": {"meshName", "prefactor"},
}
_NO_EFFECT = {"logKey", "ranksAllComputedInstances", "nInstances", "ranks", "numberTimeSteps",
              "logTimeStepWidthAsKey", "durationLogKey", "timeStepOutputInterval", "initialValues",
              "dirichletBoundaryConditions", "dirichletOutputFilename", "inputMeshIsGlobal", "checkForNanInf",
              "nAdditionalFieldVariables", "additionalSlotNames", "compilerFlags", "maximumNumberOfThreads",
              "setSpecificStatesFunction", "setSpecificStatesCallInterval", "setSpecificParametersFunction",
              "setSpecificParametersCallInterval", "handleResultFunction", "handleResultCallInterval",
              "connectedSlotsTerm1To2", "connectedSlotsTerm2To1", "solverName", "slotName", "slotNames",
              "connectorSlotsTerm1To2", "connectorSlotsTerm2To1", "OutputWriter", "solverType", "preconditionerType",
              "relativeTolerance", "absoluteTolerance", "maxIterations", "dumpFilename", "dumpFormat",
              "timeStepWidthRelativeTolerance", "timeStepWidthRelativeToleranceAsKey", "description",
              "reuseResultsFromPreviousRun", "nElements", "inputMeshIsGlobal", "enableFiberDistribution"}


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

    def _check_keys(self, fibres):
        """warn about every settings key that this front end neither applies nor
        knows to be without effect on the fibre path"""
        blocks = [("RepeatedCall", self.rc)]
        for strang, heun, key, diff in fibres[:1]:
            blocks += [("StrangSplitting", strang), ("Heun", heun), ("CellML", heun.get("CellML", {})),
                       ("Diffusion", diff), ("FiniteElementMethod", diff.get("FiniteElementMethod", {}))]
        for name, block in blocks:
            unknown = sorted(k for k in block if k not in _APPLIED[name] and k not in _NO_EFFECT)
            for k in unknown:
                msg = f"{name}.{k} is not applied by the MI355X fibre solver"
                if k == "mappings":
                    msg += (" (give parametersUsedAsAlgebraic / parametersUsedAsConstant / statesForTransfer / "
                            "algebraicsForTransfer as index lists instead)")
                warnings.warn(msg, stacklevel=3)

    def initialize(self):
        if self.initialized:
            return
        cfg, rc = self.config, self.rc
        fibres = self._fibers()
        self._check_keys(fibres)
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
        n_nodes, lengths, prefactors, fiber_no, positions = [], [], [], [], []
        for k, (_, heun, _, diff) in enumerate(fibres):
            fem = diff["FiniteElementMethod"]
            mesh_name = fem.get("meshName", heun["CellML"].get("meshName"))
            pos = node_positions(meshes[mesh_name])
            n_nodes.append(len(pos))
            positions.append(np.asarray(pos, np.float64).reshape(len(pos), -1))
            lengths.append(element_lengths(pos))
            prefactors.append(float(fem.get("prefactor", 1.0)))
            fiber_no.append(int(heun["CellML"].get("additionalArgument", k) or 0))
        self.n_nodes = np.array(n_nodes, np.int32)
        self.node_positions = [np.pad(p, ((0, 0), (0, 3 - p.shape[1]))) if p.shape[1] < 3 else p[:, :3]
                               for p in positions]
        self.fiber_no = np.array(fiber_no, np.int32)
        F = len(fibres)

        # parameters per fibre
        precision = rc.get("precision", "fp64")
        if rc.get("useSinglePrecision", False):
            # gpu.tpp:264-272 makes the whole gpu back end single precision; here
            # the states are FP32 while Vm and the diffusion solve stay FP64
            warnings.warn("useSinglePrecision: running the mixed form (FP32 gates, FP64 Vm and diffusion)",
                          stacklevel=2)
            precision = "mixed"
        self._batch_kw = dict(
            scheme=self.scheme,
            precision=precision,
            exp_mode="approx" if cellml0.get("approximateExponentialFunction", True) else "libm",
            thomas=rc.get("thomas", "partitioned"),
            kernels="fused" if rc.get("kernels", "auto") == "auto" else rc["kernels"],
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
        if cellml0.get("initializeStatesToEquilibrium", False):
            # 00_cellml_adapter_base.tpp:148-151, 03_cellml_adapter.tpp:446-520
            start = np.zeros(n_states)
            L.check(lib, lib.fms_model_defaults(L.as_ptr(start, L._dp), None), "fms_model_defaults")
            if y0:
                start = np.asarray(y0, float)
            eq = L.equilibrium_states(lib, start, params[0],
                                      float(cellml0.get("initializeStatesToEquilibriumTimestepWidth", 1e-4)))
            self.equilibrium = eq
            y0 = list(eq["states"])
            self._write_equilibrium_file(os.path.splitext(model_file)[0] if model_file else base, eq)
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
            self.stim_nmj = nmj
        if rc.get("onlyComputeIfHasBeenStimulated", True):
            self.batch.set_only_compute_if_stimulated(True)
        if rc.get("disableComputationWhenStatesAreCloseToEquilibrium", True):
            warnings.warn("disableComputationWhenStatesAreCloseToEquilibrium is not applied: every node is "
                          "computed every step (set it to False to silence this warning)", stacklevel=2)
        if rc.get("pointBufferOrder", False):
            nmj_ = self.stim_nmj if self.stim is not None else self.n_nodes // 2
            self.batch.set_point_buffer_order(int(rc.get("vcVectorWidth", 4)), nmj_)

        # time stepping
        sched = L.schedule(self.batch.lib, float(rc["endTime"]), float(rc["timeStepWidth"]),
                           float(strang0.get("endTime", rc["timeStepWidth"])), float(strang0["timeStepWidth"]),
                           0.5 * float(strang0["timeStepWidth"]), float(heun0["timeStepWidth"]),
                           float(strang0["timeStepWidth"]), float(diff0["timeStepWidth"]))
        self.schedule = sched
        self.tuned_schedule = None
        if rc.get("kernels", "auto") == "auto" and not rc.get("pointBufferOrder", False):
            # the candidates give the same bits; the trial restores every state
            rows = self._step_rows()
            if len(rows):
                self.tuned_schedule = self.batch.autotune(rows[:64], n0D=int(sched[0, 4]), allow_split=True)

        alg = _list(cellml0.get("algebraicsForTransfer", []))
        if alg:
            self.batch.set_algebraics_for_transfer(alg)
        self.algebraics_for_transfer = alg
        self.states_for_transfer = _list(cellml0.get("statesForTransfer", [0])) or [0]
        if any(not 0 <= k < n_states for k in self.states_for_transfer):
            raise ValueError(f"statesForTransfer {self.states_for_transfer}: state numbers are 0..{n_states - 1}")
        self.output_writers = [w for w in _list(strang0.get("OutputWriter", [])) if isinstance(w, dict)]
        for w in self.output_writers:
            if w.get("format") != "PythonFile":
                warnings.warn(f"OutputWriter format {w.get('format')!r} is not written (PythonFile only)",
                              stacklevel=2)
        self.output_writers = [w for w in self.output_writers if w.get("format") == "PythonFile"]
        self.stimulation_log_file = cellml0.get("stimulationLogFilename", "out/stimulation.log")

        self.initialized = True

    def _write_equilibrium_file(self, stem, eq):
        """<model>_equilibrium_values.txt as the reference writes it (03_cellml_adapter.tpp:527-560)"""
        path = self._path(f"{stem}_equilibrium_values.txt")
        try:
            with open(path, "w") as f:
                f.write(f"// equilibrium values of the states, {eq['iterations']} RK-4 iterations, dt: {eq['dt']}\n"
                        f"// Maximum du/dt = {eq['max_increment']}\n\n")
                for k, v in enumerate(eq["states"]):
                    f.write(f"state[{k}] = {v!r};\n")
                f.write("\n  Line to copy for settings:\n  \"statesInitialValues\": ["
                        + ", ".join(repr(float(v)) for v in eq["states"]) + "],\n")
        except OSError as e:
            warnings.warn(f"could not write {path}: {e}", stacklevel=3)

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
        rows = self._step_rows()
        writers = self._writers()
        if not writers:
            self.batch.advance_schedule(rows, int(n0[0]))
        else:
            # output after every outputInterval split steps, and the initial values
            interval = min(w.interval for w in writers)
            for w in writers:
                w.write(self, step=0, time=float(rows[0, 0]))
            for k0 in range(0, len(rows), interval):
                chunk = rows[k0:k0 + interval]
                self.batch.advance_schedule(chunk, int(n0[0]))
                step = k0 + len(chunk)
                for w in writers:
                    if step % w.interval == 0 or step == len(rows):
                        w.write(self, step=step, time=float(chunk[-1, 0] + chunk[-1, 1]))
        self.batch.synchronize()
        self.current_time = float(s[-1, 0] + s[-1, 1] * s[-1, 2])

    def _step_rows(self) -> np.ndarray:
        """one record (t, dt_split, dt0D, dt1D) per split step of all RepeatedCall calls"""
        rows = [(t0 + k * dts, dts, dt0, dt1) for t0, dts, ns, dt0, _, dt1, _ in self.schedule
                for k in range(int(ns))]
        return np.array(rows, np.float64).reshape(-1, 4)

    def _writers(self):
        from .output import PythonFileWriter
        if not hasattr(self, "_writer_objs"):
            self._writer_objs = [PythonFileWriter.from_settings(w, self._path) for w in self.output_writers]
        return self._writer_objs

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

    def transfer(self) -> dict:
        """statesForTransfer and algebraicsForTransfer after the last call, per
        fibre (the vc back end's updateFiberData, communication.tpp:237-296)"""
        off = np.concatenate([[0], np.cumsum(self.n_nodes)])
        states = [self.batch.get_state(k) for k in self.states_for_transfer]
        out = dict(states={k: [v[off[f]:off[f + 1]] for f in range(len(self.n_nodes))]
                           for k, v in zip(self.states_for_transfer, states)})
        if self.algebraics_for_transfer:
            a = self.batch.get_algebraics()
            out["algebraics"] = {k: [a[q, off[f]:off[f + 1]] for f in range(len(self.n_nodes))]
                                 for q, k in enumerate(self.algebraics_for_transfer)}
        return out

    def write_stimulation_log(self, path):
        times, fibers = self.batch.stimulation_log()
        if len(times) == 0:
            return
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        write_stimulation_log(path, times, self.fiber_no[fibers], self.motor_units[fibers])
