"""ctypes binding of the C ABI in include/opendihu_fms.h.

The libraries live in ``opendihu_amd/lib``; loading fails loudly when the
requested library was not built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import LIB_DIR, HH_LIBRARY

FMS_IMPLICIT_EULER, FMS_CRANK_NICOLSON = 0, 1
FMS_FP64, FMS_MIXED = 0, 1
FMS_EXP_APPROXIMATE, FMS_EXP_LIBM = 0, 1
FMS_THOMAS_SERIAL, FMS_THOMAS_PARTITIONED = 0, 1
FMS_KERNELS_SPLIT, FMS_KERNELS_FUSED = 0, 1
FMS_STIM_NONE, FMS_STIM_FIRING_TABLE, FMS_STIM_CURRENT = 0, 1, 2
FMS_OK, FMS_ERR_ARGUMENT, FMS_ERR_HIP, FMS_ERR_STATE, FMS_ERR_UNSUPPORTED, FMS_ERR_INTERNAL = 0, -1, -2, -3, -4, -5
FMS_TUNE_STEPS_PER_LAUNCH, FMS_TUNE_RESIDENT_GATES, FMS_TUNE_ONLY_IF_STIMULATED = 1, 2, 3
FMS_TUNE_THREADS, FMS_TUNE_GRAPH_STEPS, FMS_TUNE_REACTION_PER_FIBRE, FMS_TUNE_PROFILE_KERNELS = 4, 5, 6, 7
FMS_TUNE_PAIR_HALVES, FMS_TUNE_REACTION_THREADS, FMS_TUNE_ADJOINT = 8, 9, 10
FMS_ADJOINT_OVERLAP, FMS_ADJOINT_OVERLAP_PRIORITY, FMS_ADJOINT_SERIAL = 0, 1, 2
FMS_ADJOINT_OVERLAP_RECOMPUTE_PRIORITY = 3
FMS_TUNE_ADJOINT_THREADS = 11
FMS_TUNE_ADJOINT_TAPE = 12
FMS_ADJOINT_TAPE_AUTO, FMS_ADJOINT_TAPE_SEGMENTS, FMS_ADJOINT_TAPE_WHOLE = -1, 0, 1
FMS_KERNEL_CLASSES = ("reaction", "diffusion", "fused", "other", "reaction_pair")

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_ucp = ctypes.POINTER(ctypes.c_ubyte)


class FmsTuneResult(ctypes.Structure):
    _fields_ = [
        ("kernels", ctypes.c_int),
        ("steps_per_launch", ctypes.c_int),
        ("resident_gates", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("graph_steps", ctypes.c_int),
        ("ms_per_step", ctypes.c_double),
        ("n_candidates", ctypes.c_int),
        ("reaction_per_fibre", ctypes.c_int),
        ("reaction_threads", ctypes.c_int),
        ("n_rejected", ctypes.c_int),
    ]


class FmsConfig(ctypes.Structure):
    _fields_ = [
        ("n_fibers", ctypes.c_int),
        ("n_nodes", _ip),
        ("element_lengths", _dp),
        ("prefactor", _dp),
        ("parameters", _dp),
        ("initial_states", _dp),
        ("scheme", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("exp_mode", ctypes.c_int),
        ("thomas", ctypes.c_int),
        ("kernels", ctypes.c_int),
        ("value_for_stimulated_point", ctypes.c_double),
    ]


class FmsStimulation(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("stim_begin", _ip),
        ("stim_end", _ip),
        ("firing_rows", ctypes.c_int),
        ("firing_cols", ctypes.c_int),
        ("firing_events", _ucp),
        ("motor_unit", _ip),
        ("call_frequency", _dp),
        ("repeat_after_first_call", _dp),
        ("call_enable_begin", _dp),
        ("jitter", _dp),
        ("jitter_offset", _ip),
        ("jitter_count", _ip),
        ("last_check_time", _dp),
        ("log_capacity", ctypes.c_int),
        ("onset", _dp),
        ("duration", _dp),
        ("amplitude", _dp),
        ("stim_param", ctypes.c_int),
    ]


class FmsReferenceDims(ctypes.Structure):
    _fields_ = [
        ("n_fibers_to_compute", ctypes.c_int),
        ("n_instances_per_fiber", ctypes.c_int),
        ("n_parameters_per_instance", ctypes.c_int),
        ("firing_events_rows", ctypes.c_int),
        ("firing_events_cols", ctypes.c_int),
        ("frequency_jitter_cols", ctypes.c_int),
        ("n_algebraics_for_transfer", ctypes.c_int),
        ("n_states_for_transfer", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("exp_mode", ctypes.c_int),
    ]


# every symbol declared in include/opendihu_fms.h
EXPORTED_SYMBOLS = [
    "fms_model_info", "fms_model_defaults", "fms_model_ops", "fms_create", "fms_destroy", "fms_set_stimulation",
    "fms_advance", "fms_advance_schedule", "fms_synchronize", "fms_get_states", "fms_set_states",
    "fms_get_algebraics", "fms_set_algebraics_for_transfer", "fms_total_nodes",
    "fms_enable_spike_detection", "fms_get_spikes", "fms_get_stimulation_log", "fms_schedule",
    "fms_last_error", "fms_set_tuning", "fms_record_event", "fms_event_elapsed_ms", "fms_launch_counts",
    "fms_set_parameters", "fms_set_sensors", "fms_adjoint_forward", "fms_adjoint_backward",
    "fms_reference_set_dims", "initializeArrays", "computeMonodomain", "fms_graph_counts", "fms_autotune",
    "fms_eval_rhs", "fms_eval_heun", "fms_adjoint_forward_device", "fms_adjoint_backward_device",
    "fms_kernel_times", "fms_set_point_buffer_order", "fms_equilibrium_states", "fms_advance_schedule_noalloc",
    "fms_adjoint_whole_tape",
]

_LIBS: dict[str, ctypes.CDLL] = {}


def _share_hip_runtime_with_torch() -> None:
    """PyTorch-ROCm ships its own libamdhip64.so.7. Loading torch first makes the
    library's libamdhip64.so.7 dependency resolve to that same runtime (same
    soname), so torch tensors and the library share one HIP runtime and device
    context in the process; loaded the other way round, the second runtime finds
    no GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def library_path(name: str = HH_LIBRARY) -> str:
    return name if os.path.isabs(name) else os.path.join(LIB_DIR, name)


def load_library(name: str = HH_LIBRARY) -> ctypes.CDLL:
    """Load (once) a built fibre library; raises if it does not exist."""
    path = library_path(name)
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        raise RuntimeError(f"fibre library {path} is not built; run `python -m opendihu_amd.build` "
                           "(or __graft_entry__.build()) first")
    _share_hip_runtime_with_torch()
    lib = ctypes.CDLL(path)
    d, i, v = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    lib.fms_model_info.argtypes = [_ip, _ip, _ip, ctypes.c_char_p, i]
    lib.fms_model_defaults.argtypes = [_dp, _dp]
    lib.fms_model_ops.argtypes = [_dp, _dp, _dp]
    lib.fms_create.argtypes = [ctypes.POINTER(FmsConfig), ctypes.POINTER(v)]
    lib.fms_destroy.argtypes = [v]
    lib.fms_set_stimulation.argtypes = [v, ctypes.POINTER(FmsStimulation)]
    lib.fms_advance.argtypes = [v, d, d, i, d, i, d, i]
    lib.fms_advance_schedule.argtypes = [v, _dp, i, i]
    lib.fms_synchronize.argtypes = [v]
    lib.fms_get_states.argtypes = [v, i, _dp]
    lib.fms_set_states.argtypes = [v, i, _dp]
    lib.fms_get_algebraics.argtypes = [v, _dp]
    lib.fms_set_algebraics_for_transfer.argtypes = [v, _ip, i]
    lib.fms_total_nodes.argtypes = [v, ctypes.POINTER(ctypes.c_longlong)]
    lib.fms_enable_spike_detection.argtypes = [v, d]
    lib.fms_get_spikes.argtypes = [v, _ip, _ip]
    lib.fms_get_stimulation_log.argtypes = [v, _ip, _dp, _ip, i]
    lib.fms_schedule.argtypes = [d, d, d, d, d, d, d, d, i, _dp]
    lib.fms_last_error.argtypes = [ctypes.c_char_p, i]
    lib.fms_set_tuning.argtypes = [v, i, i]
    lib.fms_set_point_buffer_order.argtypes = [v, i, _ip]
    lib.fms_advance_schedule_noalloc.argtypes = [v, _dp, i, i]
    lib.fms_equilibrium_states.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_longlong,
                                           ctypes.POINTER(ctypes.c_longlong), _dp, _dp]
    lib.fms_record_event.argtypes = [v, i]
    lib.fms_event_elapsed_ms.argtypes = [v, i, i, _dp]
    lib.fms_launch_counts.argtypes = [v, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_longlong)]
    lib.fms_set_parameters.argtypes = [v, _dp, _dp]
    lib.fms_set_sensors.argtypes = [v, _ip, i]
    lib.fms_adjoint_forward.argtypes = [v, _dp, i, i, i, _dp]
    lib.fms_adjoint_backward.argtypes = [v, _dp, _dp, _dp]
    lib.fms_adjoint_forward_device.argtypes = [v, _dp, i, i, i, v]
    lib.fms_adjoint_backward_device.argtypes = [v, v, v, v]
    lib.fms_adjoint_whole_tape.argtypes = [v, _ip]
    lib.fms_eval_rhs.argtypes = [_dp, _dp, d, i, i, _dp, _dp]
    lib.fms_eval_heun.argtypes = [_dp, _dp, d, d, i, i, i]
    lib.fms_kernel_times.argtypes = [v, _dp, ctypes.POINTER(ctypes.c_longlong), i]
    lib.fms_graph_counts.argtypes = [v, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    lib.fms_autotune.argtypes = [v, _dp, i, i, i, ctypes.POINTER(FmsTuneResult), _dp, i]
    lib.fms_reference_set_dims.argtypes = [ctypes.POINTER(FmsReferenceDims)]
    lib.initializeArrays.argtypes = [_dp, _ip, _ip, ctypes.c_char_p, _dp, _ip, _ip, _dp, _dp, _dp, _dp]
    lib.initializeArrays.restype = None
    lib.computeMonodomain.argtypes = [ctypes.POINTER(ctypes.c_float), _dp, _dp, ctypes.POINTER(ctypes.c_float),
                                      d, d, i, d, i, d, i, d, d]
    lib.computeMonodomain.restype = None
    for name_ in EXPORTED_SYMBOLS:
        if name_ not in ("initializeArrays", "computeMonodomain"):
            getattr(lib, name_).restype = ctypes.c_int
    _LIBS[path] = lib
    return lib


def last_error(lib: ctypes.CDLL) -> str:
    buf = ctypes.create_string_buffer(4096)
    lib.fms_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(lib: ctypes.CDLL, rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error(lib)}")


def as_ptr(a: np.ndarray, ptype):
    return a.ctypes.data_as(ptype)


def model_info(lib: ctypes.CDLL) -> tuple[int, int, int, str]:
    ns, na, nparam = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    buf = ctypes.create_string_buffer(256)
    check(lib, lib.fms_model_info(ctypes.byref(ns), ctypes.byref(na), ctypes.byref(nparam), buf, 256), "fms_model_info")
    return ns.value, na.value, nparam.value, buf.value.decode()


def eval_rhs(lib: ctypes.CDLL, states: np.ndarray, params: np.ndarray, t: float = 0.0, exp_mode: str = "approx",
             algebraics: bool = False):
    """The model's right-hand side on the device: states (n_states, n), params (n_params, n)
    -> rates (n_states, n) [, algebraics (n_algebraics, n)]"""
    ns, na, nparams, _ = model_info(lib)
    y = np.ascontiguousarray(states, np.float64).reshape(ns, -1)
    n = y.shape[1]
    p = np.ascontiguousarray(params, np.float64).reshape(max(nparams, 1), -1) if nparams else np.zeros((1, n))
    r = np.empty((ns, n))
    a = np.empty((na, n)) if algebraics else None
    check(lib, lib.fms_eval_rhs(as_ptr(y, _dp), as_ptr(p, _dp), float(t), n,
                                FMS_EXP_APPROXIMATE if exp_mode == "approx" else FMS_EXP_LIBM, as_ptr(r, _dp),
                                as_ptr(a, _dp) if a is not None else None), "fms_eval_rhs")
    return (r, a) if algebraics else r


def eval_heun(lib: ctypes.CDLL, states: np.ndarray, params: np.ndarray, t0: float, dt: float, n_steps: int,
              exp_mode: str = "approx") -> np.ndarray:
    """n_steps Heun steps of independent points on the device: states (n_states, n) -> (n_states, n)"""
    ns, _, nparams, _ = model_info(lib)
    y = np.array(states, np.float64, copy=True).reshape(ns, -1)
    n = y.shape[1]
    p = np.ascontiguousarray(params, np.float64).reshape(max(nparams, 1), -1) if nparams else np.zeros((1, n))
    check(lib, lib.fms_eval_heun(as_ptr(y, _dp), as_ptr(p, _dp), float(t0), float(dt), int(n_steps), n,
                                 FMS_EXP_APPROXIMATE if exp_mode == "approx" else FMS_EXP_LIBM), "fms_eval_heun")
    return y


def model_ops(lib: ctypes.CDLL) -> dict:
    o, d, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib, lib.fms_model_ops(ctypes.byref(o), ctypes.byref(d), ctypes.byref(e)), "fms_model_ops")
    return dict(ops_rhs=o.value, divs_rhs=d.value, exps_rhs=e.value)


# algorithmic ops per node of one diffusion solve, as the reference's row
# assembly and Thomas recurrence are written (fast_monodomain_solver_compute.tpp:336-486)
THOMAS_OPS = {"IE": 60, "CN": 68}


def flops_per_node_step(lib: ctypes.CDLL, n_states: int, n0D: int = 1, scheme: str = "IE") -> float:
    """algorithmic ops per fiber-node-step: two reaction halves of n0D Heun steps
    (two rhs evaluations + 5 ops per state each) plus one diffusion solve"""
    ops = model_ops(lib)["ops_rhs"]
    return 2 * n0D * (2 * ops + 5 * n_states) + THOMAS_OPS[scheme]


def schedule(lib: ctypes.CDLL, rc_end, rc_width, strang_end, strang_width, heun_end, heun_width, ie_end,
             ie_width) -> np.ndarray:
    """Per-call (startTime, dtSplit, nSplit, dt0D, n0D, dt1D, n1D) of a RepeatedCall run."""
    n = lib.fms_schedule(rc_end, rc_width, strang_end, strang_width, heun_end, heun_width, ie_end, ie_width, 0, None)
    if n < 0:
        raise RuntimeError(last_error(lib))
    out = np.zeros((max(n, 1), 7))
    lib.fms_schedule(rc_end, rc_width, strang_end, strang_width, heun_end, heun_width, ie_end, ie_width, n,
                     as_ptr(out, _dp))
    return out[:n]


def equilibrium_states(lib: ctypes.CDLL, states, params, dt: float = 1e-4, max_iterations: int = 10_000_000) -> dict:
    """initializeStatesToEquilibrium of one cell (core/src/cellml/03_cellml_adapter.tpp:446-520):
    {states, iterations, max_increment, dt}; raises if no equilibrium was found"""
    ns, _, nparams, _ = model_info(lib)
    y = np.ascontiguousarray(states, np.float64).copy().reshape(ns)
    p = np.ascontiguousarray(params if nparams > 0 else [0.0], np.float64)
    it, inc, dt_out = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
    check(lib, lib.fms_equilibrium_states(as_ptr(y, _dp), as_ptr(p, _dp), float(dt), int(max_iterations),
                                          ctypes.byref(it), ctypes.byref(inc), ctypes.byref(dt_out)),
          "fms_equilibrium_states")
    return dict(states=y, iterations=it.value, max_increment=inc.value, dt=dt_out.value)
