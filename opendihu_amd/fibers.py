"""Low-level handle over one batch of fibres on the GPU (native fms_* API)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


class FiberBatch:
    """Fibres of (possibly) uneven length, state-major on the device.

    Parameters mirror FmsConfig (include/opendihu_fms.h). ``n_nodes`` and
    ``element_lengths`` describe the 1D meshes; ``prefactor`` is sigma/(Am*Cm)
    per fibre; ``parameters`` is (n_fibers, n_params) or None for the model's
    defaults.
    """

    def __init__(self, n_nodes, element_lengths, prefactor, parameters=None, initial_states=None,
                 scheme="IE", precision="fp64", exp_mode="approx", thomas="partitioned", kernels="split",
                 value_for_stimulated_point=20.0, library=L.HH_LIBRARY):
        self.lib = L.load_library(library)
        self.n_states, self.n_algebraics, self.n_params, self.model_name = L.model_info(self.lib)
        self.n_nodes = np.ascontiguousarray(n_nodes, dtype=np.int32)
        self.n_fibers = int(self.n_nodes.size)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_nodes)]).astype(np.int64)
        self.n_total = int(self.offsets[-1])
        self.element_lengths = np.ascontiguousarray(element_lengths, dtype=np.float64)
        if self.element_lengths.size != self.n_total - self.n_fibers:
            raise ValueError("element_lengths must hold sum(n_nodes-1) values")
        pf = np.asarray(prefactor, dtype=np.float64)
        self.prefactor = np.ascontiguousarray(np.broadcast_to(pf, (self.n_fibers,)), dtype=np.float64)
        if parameters is None:
            defaults = np.zeros(self.n_params)
            self.lib.fms_model_defaults(None, L.as_ptr(defaults, L._dp)) if self.n_params else None
            parameters = np.tile(defaults, (self.n_fibers, 1))
        self.parameters = np.ascontiguousarray(np.asarray(parameters, dtype=np.float64).reshape(self.n_fibers, -1))
        if self.parameters.shape[1] != self.n_params:
            raise ValueError(f"model {self.model_name} has {self.n_params} parameters")
        self._keep = []
        cfg = L.FmsConfig()
        cfg.n_fibers = self.n_fibers
        cfg.n_nodes = L.as_ptr(self.n_nodes, L._ip)
        cfg.element_lengths = L.as_ptr(self.element_lengths, L._dp)
        cfg.prefactor = L.as_ptr(self.prefactor, L._dp)
        cfg.parameters = L.as_ptr(self.parameters, L._dp)
        if initial_states is not None:
            y0 = np.ascontiguousarray(initial_states, dtype=np.float64)
            self._keep.append(y0)
            cfg.initial_states = L.as_ptr(y0, L._dp)
        cfg.scheme = {"IE": L.FMS_IMPLICIT_EULER, "ImplicitEuler": L.FMS_IMPLICIT_EULER,
                      "CN": L.FMS_CRANK_NICOLSON, "CrankNicolson": L.FMS_CRANK_NICOLSON}[scheme]
        cfg.precision = {"fp64": L.FMS_FP64, "mixed": L.FMS_MIXED}[precision]
        cfg.exp_mode = {"approx": L.FMS_EXP_APPROXIMATE, "libm": L.FMS_EXP_LIBM}[exp_mode]
        cfg.thomas = {"serial": L.FMS_THOMAS_SERIAL, "partitioned": L.FMS_THOMAS_PARTITIONED}[thomas]
        cfg.kernels = {"split": L.FMS_KERNELS_SPLIT, "fused": L.FMS_KERNELS_FUSED}[kernels]
        cfg.value_for_stimulated_point = float(value_for_stimulated_point)
        handle = ctypes.c_void_p()
        L.check(self.lib, self.lib.fms_create(ctypes.byref(cfg), ctypes.byref(handle)), "fms_create")
        self.handle = handle
        self.config = dict(scheme=scheme, precision=precision, exp_mode=exp_mode, thomas=thomas, kernels=kernels)

    # -- lifetime -----------------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None):
            self.lib.fms_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        L.check(self.lib, rc, what)

    # -- stimulation ----------------------------------------------------------------
    def set_firing_table_stimulation(self, stim_begin, stim_end, firing_events, motor_unit, call_frequency,
                                     repeat_after_first_call, call_enable_begin, jitter_lists=None,
                                     last_check_time=None, log_capacity=0):
        F = self.n_fibers
        fe = np.ascontiguousarray(firing_events, dtype=np.uint8)
        if fe.ndim != 2:
            raise ValueError("firing_events must be (rows, motor units)")
        jitter_lists = jitter_lists if jitter_lists is not None else [[] for _ in range(F)]
        joff = np.zeros(F, np.int32)
        jcnt = np.zeros(F, np.int32)
        flat = []
        for f, lst in enumerate(jitter_lists):
            joff[f] = len(flat)
            jcnt[f] = len(lst)
            flat.extend(lst)
        arrs = dict(
            sb=np.ascontiguousarray(np.broadcast_to(stim_begin, (F,)), np.int32),
            se=np.ascontiguousarray(np.broadcast_to(stim_end, (F,)), np.int32),
            mu=np.ascontiguousarray(np.broadcast_to(motor_unit, (F,)), np.int32),
            fr=np.ascontiguousarray(np.broadcast_to(call_frequency, (F,)), np.float64),
            rp=np.ascontiguousarray(np.broadcast_to(repeat_after_first_call, (F,)), np.float64),
            bg=np.ascontiguousarray(np.broadcast_to(call_enable_begin, (F,)), np.float64),
            jit=np.ascontiguousarray(flat if flat else [0.0], np.float64), joff=joff, jcnt=jcnt, fe=fe)
        st = L.FmsStimulation()
        st.mode = L.FMS_STIM_FIRING_TABLE
        st.stim_begin = L.as_ptr(arrs["sb"], L._ip)
        st.stim_end = L.as_ptr(arrs["se"], L._ip)
        st.firing_rows, st.firing_cols = fe.shape
        st.firing_events = L.as_ptr(fe, L._ucp)
        st.motor_unit = L.as_ptr(arrs["mu"], L._ip)
        st.call_frequency = L.as_ptr(arrs["fr"], L._dp)
        st.repeat_after_first_call = L.as_ptr(arrs["rp"], L._dp)
        st.call_enable_begin = L.as_ptr(arrs["bg"], L._dp)
        st.jitter = L.as_ptr(arrs["jit"], L._dp)
        st.jitter_offset = L.as_ptr(joff, L._ip)
        st.jitter_count = L.as_ptr(jcnt, L._ip)
        if last_check_time is not None:
            arrs["lc"] = np.ascontiguousarray(np.broadcast_to(last_check_time, (F,)), np.float64)
            st.last_check_time = L.as_ptr(arrs["lc"], L._dp)
        st.log_capacity = int(log_capacity)
        self._chk(self.lib.fms_set_stimulation(self.handle, ctypes.byref(st)), "fms_set_stimulation")
        self._stim_arrays = arrs

    def set_current_stimulation(self, stim_begin, stim_end, onset, duration, amplitude, stim_param=0):
        F = self.n_fibers
        arrs = dict(
            sb=np.ascontiguousarray(np.broadcast_to(stim_begin, (F,)), np.int32),
            se=np.ascontiguousarray(np.broadcast_to(stim_end, (F,)), np.int32),
            on=np.ascontiguousarray(np.broadcast_to(onset, (F,)), np.float64),
            du=np.ascontiguousarray(np.broadcast_to(duration, (F,)), np.float64),
            am=np.ascontiguousarray(np.broadcast_to(amplitude, (F,)), np.float64))
        st = L.FmsStimulation()
        st.mode = L.FMS_STIM_CURRENT
        st.stim_begin = L.as_ptr(arrs["sb"], L._ip)
        st.stim_end = L.as_ptr(arrs["se"], L._ip)
        st.onset = L.as_ptr(arrs["on"], L._dp)
        st.duration = L.as_ptr(arrs["du"], L._dp)
        st.amplitude = L.as_ptr(arrs["am"], L._dp)
        st.stim_param = int(stim_param)
        self._chk(self.lib.fms_set_stimulation(self.handle, ctypes.byref(st)), "fms_set_stimulation")
        self._stim_arrays = arrs

    # -- time stepping --------------------------------------------------------------
    def advance(self, start_time, dt_split, n_split, dt0D=None, n0D=1, dt1D=None, n1D=1):
        dt0D = 0.5 * dt_split if dt0D is None else dt0D
        dt1D = dt_split if dt1D is None else dt1D
        self._chk(self.lib.fms_advance(self.handle, float(start_time), float(dt_split), int(n_split), float(dt0D),
                                       int(n0D), float(dt1D), int(n1D)), "fms_advance")

    def advance_schedule(self, steps, n0D=1):
        steps = np.ascontiguousarray(steps, dtype=np.float64).reshape(-1, 4)
        self._chk(self.lib.fms_advance_schedule(self.handle, L.as_ptr(steps, L._dp), steps.shape[0], int(n0D)),
                  "fms_advance_schedule")

    def set_only_compute_if_stimulated(self, on: bool) -> None:
        """onlyComputeIfHasBeenStimulated: skip a fibre's reaction until its first stimulation"""
        self._chk(self.lib.fms_set_tuning(self.handle, L.FMS_TUNE_ONLY_IF_STIMULATED, int(bool(on))),
                  "fms_set_tuning")

    def set_point_buffer_order(self, width: int, nmj=None) -> None:
        """onlyComputeIfHasBeenStimulated in the reference's sweep over buffers of
        `width` nodes (compute.tpp:131-216); nmj: each fibre's stimulation
        point (fibre-local); width 0 returns to the per-fibre rule"""
        arr = np.ascontiguousarray(np.broadcast_to(nmj if nmj is not None else 0, (self.n_fibers,)), np.int32)
        self._chk(self.lib.fms_set_point_buffer_order(self.handle, int(width), L.as_ptr(arr, L._ip)),
                  "fms_set_point_buffer_order")

    def set_tuning(self, steps_per_launch: int | None = None, resident_gates: int | None = None,
                   threads: int | None = None, graph_steps: int | None = None,
                   reaction_per_fibre: int | None = None, pair_halves: int | None = None,
                   reaction_threads: int | None = None, adjoint: int | None = None,
                   adjoint_threads: int | None = None, adjoint_tape: int | None = None) -> None:
        """Schedule knobs. Fused kernel: split steps per launch; gates in registers (1),
        streamed (0) or auto (-1); threads per fibre workgroup (128 or 256). Split
        kernels: steps per captured hipGraph (0 launches eagerly); for large models
        the reaction with one workgroup per fibre (1) or one lane per node (0); the
        two reaction halves between two solves in one launch (1, default) or one
        launch per half (0); threads per fibre of the per-fibre reaction (256 or 64).
        Adjoint backward: segment recompute beside the reverse sweep (0), the same
        with the reverse sweep's priority raised (1) or the recompute's (3), or one
        after the other (2); threads per fibre workgroup of the reverse sweep (256,
        512 or 1024); the adjoint forward records the whole tape (1) or keeps
        checkpoints that the backward recomputes from (0), or the whole tape when it
        fits in half of the free device memory (-1, default)."""
        for key, value in ((L.FMS_TUNE_STEPS_PER_LAUNCH, steps_per_launch), (L.FMS_TUNE_RESIDENT_GATES, resident_gates),
                           (L.FMS_TUNE_THREADS, threads), (L.FMS_TUNE_GRAPH_STEPS, graph_steps),
                           (L.FMS_TUNE_REACTION_PER_FIBRE, reaction_per_fibre),
                           (L.FMS_TUNE_PAIR_HALVES, pair_halves), (L.FMS_TUNE_REACTION_THREADS, reaction_threads),
                           (L.FMS_TUNE_ADJOINT, adjoint), (L.FMS_TUNE_ADJOINT_THREADS, adjoint_threads),
                           (L.FMS_TUNE_ADJOINT_TAPE, adjoint_tape)):
            if value is not None:
                self._chk(self.lib.fms_set_tuning(self.handle, key, int(value)), "fms_set_tuning")

    def profile_kernels(self, on: bool = True) -> None:
        """time every kernel launch with a start/stop event pair bound to it (for
        measurement runs; kernel_times() reads the sums back)"""
        self._chk(self.lib.fms_set_tuning(self.handle, L.FMS_TUNE_PROFILE_KERNELS, int(bool(on))), "fms_set_tuning")

    def kernel_times(self, reset: bool = False) -> dict:
        """{class: (total ms, launches)} for reaction, diffusion, fused, other, reaction_pair"""
        ms = np.zeros(len(L.FMS_KERNEL_CLASSES))
        n = np.zeros(len(L.FMS_KERNEL_CLASSES), np.int64)
        self._chk(self.lib.fms_kernel_times(self.handle, L.as_ptr(ms, L._dp),
                                            n.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), int(reset)),
                  "fms_kernel_times")
        return {k: (float(ms[j]), int(n[j])) for j, k in enumerate(L.FMS_KERNEL_CLASSES)}

    def graph_counts(self) -> dict:
        c, r = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self.lib.fms_graph_counts(self.handle, ctypes.byref(c), ctypes.byref(r)), "fms_graph_counts")
        return dict(captures=c.value, replays=r.value)

    def autotune(self, steps, n0D: int = 1, allow_split: bool = True) -> dict:
        """Time every candidate schedule over the step records `steps` (n, 4)
        on this batch's own states, restore the states, keep the fastest of
        those whose end state is bit-identical to the current schedule's
        (`rejected` counts the others; their `candidates_ms` entry is -1)."""
        steps = np.ascontiguousarray(steps, dtype=np.float64).reshape(-1, 4)
        res = L.FmsTuneResult()
        ms = np.full(64, np.nan)
        self._chk(self.lib.fms_autotune(self.handle, L.as_ptr(steps, L._dp), steps.shape[0], int(n0D),
                                        int(bool(allow_split)), ctypes.byref(res), L.as_ptr(ms, L._dp), ms.size),
                  "fms_autotune")
        self.config["kernels"] = "fused" if res.kernels == L.FMS_KERNELS_FUSED else "split"
        return dict(kernels=self.config["kernels"], steps_per_launch=res.steps_per_launch,
                    resident_gates=res.resident_gates, threads=res.threads, graph_steps=res.graph_steps,
                    reaction_per_fibre=res.reaction_per_fibre, reaction_threads=res.reaction_threads,
                    ms_per_step=res.ms_per_step, candidates_ms=ms[:res.n_candidates].tolist(),
                    rejected=res.n_rejected)

    def record_event(self, slot: int) -> None:
        self._chk(self.lib.fms_record_event(self.handle, int(slot)), "fms_record_event")

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = ctypes.c_double()
        self._chk(self.lib.fms_event_elapsed_ms(self.handle, int(a), int(b), ctypes.byref(ms)), "fms_event_elapsed_ms")
        return ms.value

    def launch_counts(self) -> dict:
        f, r, t = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self.lib.fms_launch_counts(self.handle, ctypes.byref(f), ctypes.byref(r), ctypes.byref(t)),
                  "fms_launch_counts")
        return dict(fused=f.value, reaction=r.value, thomas=t.value)

    def synchronize(self):
        self._chk(self.lib.fms_synchronize(self.handle), "fms_synchronize")

    # -- data ---------------------------------------------------------------------
    def get_state(self, k: int) -> np.ndarray:
        out = np.empty(self.n_total, np.float64)
        self._chk(self.lib.fms_get_states(self.handle, int(k), L.as_ptr(out, L._dp)), "fms_get_states")
        return out

    def get_states(self) -> np.ndarray:
        return np.stack([self.get_state(k) for k in range(self.n_states)])

    def set_state(self, k: int, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._chk(self.lib.fms_set_states(self.handle, int(k), L.as_ptr(v, L._dp)), "fms_set_states")

    def vm(self) -> np.ndarray:
        return self.get_state(0)

    def set_algebraics_for_transfer(self, indices):
        idx = np.ascontiguousarray(indices, np.int32)
        self._chk(self.lib.fms_set_algebraics_for_transfer(self.handle, L.as_ptr(idx, L._ip), idx.size),
                  "fms_set_algebraics_for_transfer")
        self._alg_count = idx.size

    def get_algebraics(self) -> np.ndarray:
        out = np.empty((self._alg_count, self.n_total), np.float64)
        self._chk(self.lib.fms_get_algebraics(self.handle, L.as_ptr(out, L._dp)), "fms_get_algebraics")
        return out

    def enable_spike_detection(self, threshold=0.0):
        self._chk(self.lib.fms_enable_spike_detection(self.handle, float(threshold)), "fms_enable_spike_detection")

    def spikes(self):
        first = np.empty(self.n_total, np.int32)
        count = np.empty(self.n_total, np.int32)
        self._chk(self.lib.fms_get_spikes(self.handle, L.as_ptr(first, L._ip), L.as_ptr(count, L._ip)),
                  "fms_get_spikes")
        return first, count

    def stimulation_log(self, capacity=100000):
        cnt = ctypes.c_int()
        times = np.empty(capacity, np.float64)
        fibers = np.empty(capacity, np.int32)
        self._chk(self.lib.fms_get_stimulation_log(self.handle, ctypes.byref(cnt), L.as_ptr(times, L._dp),
                                                   L.as_ptr(fibers, L._ip), capacity), "fms_get_stimulation_log")
        n = min(cnt.value, capacity)
        return times[:n].copy(), fibers[:n].copy()
