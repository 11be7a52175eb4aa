"""ctypes binding of the CPU oracle (oracle/fms_oracle_core.h), for tests and the CPU baseline.

``OracleBatch`` mirrors ``opendihu_amd.fibers.FiberBatch`` so parity tests read the
same on both sides.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_BUILD = os.path.join(ORACLE_DIR, "build")

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_ucp = ctypes.POINTER(ctypes.c_ubyte)


class fmso_problem(ctypes.Structure):
    _fields_ = [
        ("n_fibers", ctypes.c_int), ("n_nodes", _ip), ("elem_len", _dp), ("prefactor", _dp), ("params", _dp),
        ("scheme", ctypes.c_int), ("mass", ctypes.c_int), ("exp_mode", ctypes.c_int),
        ("value_for_stimulated_point", ctypes.c_double), ("stim_mode", ctypes.c_int),
        ("stim_begin", _ip), ("stim_end", _ip),
        ("fe_rows", ctypes.c_int), ("fe_cols", ctypes.c_int), ("firing_events", _ucp), ("motor_unit", _ip),
        ("call_frequency", _dp), ("repeat_after_first_call", _dp), ("call_enable_begin", _dp),
        ("jitter", _dp), ("jitter_offset", _ip), ("jitter_count", _ip),
        ("stim_onset", _dp), ("stim_duration", _dp), ("stim_amplitude", _dp), ("stim_param", ctypes.c_int),
        ("spike_threshold", ctypes.c_double),
        ("only_if_stimulated", ctypes.c_int),
        ("n_alg_transfer", ctypes.c_int), ("alg_transfer_idx", _ip),
        ("vc_width", ctypes.c_int), ("nmj", _ip), ("equilibrium_skip", ctypes.c_int),
    ]


class fmso_state(ctypes.Structure):
    _fields_ = [
        ("states", _dp), ("last_check_time", _dp), ("current_jitter", _dp), ("jitter_index", _ip),
        ("currently_stimulating", _ucp), ("has_been_stimulated", _ucp), ("first_spike_step", _ip),
        ("spike_count", _ip), ("step_counter", ctypes.c_int64), ("stim_log_count", _ip),
        ("stim_log_time", _dp), ("stim_log_fiber", _ip), ("stim_log_capacity", ctypes.c_int),
        ("alg_transfer", _dp), ("vc_eq", _ucp), ("vc_counters", ctypes.POINTER(ctypes.c_int64)), ("vc_pad", _dp),
    ]


_LIBS = {}


def oracle_library(model: str = "hh") -> ctypes.CDLL:
    path = os.path.join(ORACLE_BUILD, f"libfmso_{model}.so")
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)
    lib = ctypes.CDLL(path)
    d, i = ctypes.c_double, ctypes.c_int
    lib.fmso_advance.argtypes = [ctypes.POINTER(fmso_problem), ctypes.POINTER(fmso_state), d, d, i, d, i, d, i, i]
    lib.fmso_rhs.argtypes = [_dp, _dp, d, _dp, _dp, i]
    lib.fmso_heun_cell.argtypes = [_dp, _dp, d, d, i, i]
    lib.fmso_euler_cell.argtypes = [_dp, _dp, d, d, ctypes.c_long, i]
    lib.fmso_thomas.argtypes = [_dp, i, _dp, d, d, i, i]
    lib.fmso_model_info.argtypes = [_ip, _ip, _ip]
    lib.fmso_initial_states.argtypes = [_dp]
    lib.fmso_default_params.argtypes = [_dp]
    lib.fmso_schedule.argtypes = [d, d, d, d, d, d, d, d, i, _dp]
    lib.fmso_model_name.restype = ctypes.c_char_p
    lib.fmso_advance_vc.argtypes = [ctypes.POINTER(fmso_problem), ctypes.POINTER(fmso_state), d, d, i, d, i, d, i]
    lib.fmso_equilibrium.argtypes = [_dp, _dp, d, ctypes.c_long, _dp, _dp]
    lib.fmso_equilibrium.restype = ctypes.c_long
    _LIBS[path] = lib
    return lib


def P(a, t):
    return a.ctypes.data_as(t)


def model_info(lib):
    ns, na, np_ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.fmso_model_info(ctypes.byref(ns), ctypes.byref(na), ctypes.byref(np_))
    return ns.value, na.value, np_.value


def equilibrium(lib, y0, params, dt=1e-4, max_iter=10_000_000):
    """initializeStatesToEquilibrium of one cell: (states, iterations or -1,
    last largest increment, final step width)"""
    y = np.ascontiguousarray(y0, np.float64).copy()
    prm = np.ascontiguousarray(params, np.float64)
    inc, dt_out = ctypes.c_double(), ctypes.c_double()
    it = lib.fmso_equilibrium(P(y, _dp), P(prm, _dp), float(dt), int(max_iter), ctypes.byref(inc), ctypes.byref(dt_out))
    return y, int(it), inc.value, dt_out.value


def schedule(lib, *args) -> np.ndarray:
    n = lib.fmso_schedule(*[float(a) for a in args], 0, None)
    out = np.zeros((max(n, 1), 7))
    lib.fmso_schedule(*[float(a) for a in args], n, P(out, _dp))
    return out[:n]


class OracleBatch:
    """CPU oracle over the same inputs as FiberBatch (FP64, reference order)."""

    def __init__(self, n_nodes, element_lengths, prefactor, parameters=None, initial_states=None, scheme="IE",
                 exp_mode="approx", value_for_stimulated_point=20.0, model="hh", mass="consistent", threads=None):
        self.lib = oracle_library(model)
        self.n_states, self.n_algebraics, self.n_params = model_info(self.lib)
        self.n_nodes = np.ascontiguousarray(n_nodes, np.int32)
        self.n_fibers = self.n_nodes.size
        self.n_total = int(self.n_nodes.sum())
        self.elem = np.ascontiguousarray(element_lengths, np.float64)
        self.pf = np.ascontiguousarray(np.broadcast_to(np.asarray(prefactor, np.float64), (self.n_fibers,)))
        if parameters is None:
            p0 = np.zeros(self.n_params)
            self.lib.fmso_default_params(P(p0, _dp))
            parameters = np.tile(p0, (self.n_fibers, 1))
        self.params = np.ascontiguousarray(np.asarray(parameters, np.float64).reshape(self.n_fibers, -1))
        y0 = np.zeros(self.n_states)
        self.lib.fmso_initial_states(P(y0, _dp))
        if initial_states is not None:
            y0 = np.asarray(initial_states, np.float64)
        self.states = np.ascontiguousarray(np.repeat(y0[:, None], self.n_total, axis=1))
        F = self.n_fibers
        self.threads = threads or 1
        self.prob = fmso_problem()
        p = self.prob
        p.n_fibers = F
        p.n_nodes = P(self.n_nodes, _ip)
        p.elem_len = P(self.elem, _dp)
        p.prefactor = P(self.pf, _dp)
        p.params = P(self.params, _dp)
        p.scheme = {"IE": 0, "CN": 1}[scheme]
        p.mass = {"consistent": 0, "lumped": 1}[mass]
        p.exp_mode = {"approx": 0, "libm": 1}[exp_mode]
        p.value_for_stimulated_point = value_for_stimulated_point
        p.stim_mode = 0
        self._zi = np.zeros(F, np.int32)
        p.stim_begin = P(self._zi, _ip)
        p.stim_end = P(self._zi, _ip)
        p.spike_threshold = float("nan")
        self.st = fmso_state()
        self.st.states = P(self.states, _dp)
        self._sm = dict(last=np.zeros(F), jit=np.zeros(F), jidx=np.zeros(F, np.int32), curr=np.zeros(F, np.uint8),
                        has=np.zeros(F, np.uint8))
        self.st.last_check_time = P(self._sm["last"], _dp)
        self.st.current_jitter = P(self._sm["jit"], _dp)
        self.st.jitter_index = P(self._sm["jidx"], _ip)
        self.st.currently_stimulating = P(self._sm["curr"], _ucp)
        self.st.has_been_stimulated = P(self._sm["has"], _ucp)
        self.st.step_counter = 0

    def set_firing_table_stimulation(self, stim_begin, stim_end, firing_events, motor_unit, call_frequency,
                                     repeat_after_first_call, call_enable_begin, jitter_lists=None,
                                     last_check_time=None, log_capacity=0):
        F = self.n_fibers
        fe = np.ascontiguousarray(firing_events, np.uint8)
        jitter_lists = jitter_lists if jitter_lists is not None else [[] for _ in range(F)]
        joff = np.zeros(F, np.int32)
        jcnt = np.zeros(F, np.int32)
        flat = []
        for f, lst in enumerate(jitter_lists):
            joff[f], jcnt[f] = len(flat), len(lst)
            flat.extend(lst)
        a = self._stim = dict(
            sb=np.ascontiguousarray(np.broadcast_to(stim_begin, (F,)), np.int32),
            se=np.ascontiguousarray(np.broadcast_to(stim_end, (F,)), np.int32),
            mu=np.ascontiguousarray(np.broadcast_to(motor_unit, (F,)), np.int32),
            fr=np.ascontiguousarray(np.broadcast_to(call_frequency, (F,)), np.float64),
            rp=np.ascontiguousarray(np.broadcast_to(repeat_after_first_call, (F,)), np.float64),
            bg=np.ascontiguousarray(np.broadcast_to(call_enable_begin, (F,)), np.float64),
            jit=np.ascontiguousarray(flat if flat else [0.0], np.float64), joff=joff, jcnt=jcnt, fe=fe)
        p = self.prob
        p.stim_mode = 1
        p.stim_begin, p.stim_end = P(a["sb"], _ip), P(a["se"], _ip)
        p.fe_rows, p.fe_cols = fe.shape
        p.firing_events = P(fe, _ucp)
        p.motor_unit = P(a["mu"], _ip)
        p.call_frequency, p.repeat_after_first_call, p.call_enable_begin = P(a["fr"], _dp), P(a["rp"], _dp), P(a["bg"], _dp)
        p.jitter, p.jitter_offset, p.jitter_count = P(a["jit"], _dp), P(joff, _ip), P(jcnt, _ip)
        # lastStimulationCheckTime = enableBegin - 1e-13 - 1/(f + 0), initialization.tpp:344
        if last_check_time is None:
            self._sm["last"][:] = a["bg"] - 1e-13 - 1.0 / (a["fr"] + 0.0)
        else:
            self._sm["last"][:] = last_check_time
        if log_capacity:
            self._log = dict(n=np.zeros(1, np.int32), t=np.zeros(log_capacity), f=np.zeros(log_capacity, np.int32))
            self.st.stim_log_count = P(self._log["n"], _ip)
            self.st.stim_log_time = P(self._log["t"], _dp)
            self.st.stim_log_fiber = P(self._log["f"], _ip)
            self.st.stim_log_capacity = log_capacity

    def set_current_stimulation(self, stim_begin, stim_end, onset, duration, amplitude, stim_param=0):
        F = self.n_fibers
        a = self._stim = dict(
            sb=np.ascontiguousarray(np.broadcast_to(stim_begin, (F,)), np.int32),
            se=np.ascontiguousarray(np.broadcast_to(stim_end, (F,)), np.int32),
            on=np.ascontiguousarray(np.broadcast_to(onset, (F,)), np.float64),
            du=np.ascontiguousarray(np.broadcast_to(duration, (F,)), np.float64),
            am=np.ascontiguousarray(np.broadcast_to(amplitude, (F,)), np.float64))
        p = self.prob
        p.stim_mode = 2
        p.stim_begin, p.stim_end = P(a["sb"], _ip), P(a["se"], _ip)
        p.stim_onset, p.stim_duration, p.stim_amplitude = P(a["on"], _dp), P(a["du"], _dp), P(a["am"], _dp)
        p.stim_param = stim_param

    def set_only_compute_if_stimulated(self, on: bool) -> None:
        self.prob.only_if_stimulated = int(bool(on))

    def set_algebraics_for_transfer(self, indices) -> None:
        """algebraicsForTransfer: the last reaction half of every advance() stores
        these algebraics of its last Heun step's second evaluation"""
        self._alg_idx = np.ascontiguousarray(indices, np.int32)
        self._alg = np.zeros((self._alg_idx.size, self.n_total))
        self.prob.n_alg_transfer = self._alg_idx.size
        self.prob.alg_transfer_idx = P(self._alg_idx, _ip)
        self.st.alg_transfer = P(self._alg, _dp)

    def algebraics(self) -> np.ndarray:
        return self._alg.copy()

    def enable_vc_sweep(self, nmj, width=4, equilibrium_skip=False) -> None:
        """advance() runs the reference's point-buffer sweep (W nodes per
        buffer, stimulation and skip decisions per buffer in buffer order)"""
        n_buf = (self.n_total + width - 1) // width
        self._vc = dict(nmj=np.ascontiguousarray(np.broadcast_to(nmj, (self.n_fibers,)), np.int32),
                        eq=np.full(n_buf, 2, np.uint8), cnt=np.zeros(4, np.int64),
                        pad=np.ascontiguousarray(np.repeat(self.states[:, :1], width, axis=1)))
        p = self.prob
        p.vc_width = width
        p.nmj = P(self._vc["nmj"], _ip)
        p.equilibrium_skip = int(bool(equilibrium_skip))
        self.st.vc_eq = P(self._vc["eq"], _ucp)
        self.st.vc_counters = self._vc["cnt"].ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        self.st.vc_pad = P(self._vc["pad"], _dp)

    def vc_counters(self) -> dict:
        c = self._vc["cnt"]
        return dict(inactive_counter=int(c[0]), computed=int(c[1]), skipped_equilibrium=int(c[2]),
                    skipped_not_stimulated=int(c[3]))

    def vc_equilibrium_states(self) -> np.ndarray:
        return self._vc["eq"].copy()

    def enable_spike_detection(self, threshold=0.0):
        self.prob.spike_threshold = threshold
        self._spk = dict(first=np.full(self.n_total, -1, np.int32), count=np.zeros(self.n_total, np.int32))
        self.st.first_spike_step = P(self._spk["first"], _ip)
        self.st.spike_count = P(self._spk["count"], _ip)

    def advance(self, start_time, dt_split, n_split, dt0D=None, n0D=1, dt1D=None, n1D=1):
        dt0D = 0.5 * dt_split if dt0D is None else dt0D
        dt1D = dt_split if dt1D is None else dt1D
        if getattr(self, "_vc", None) is not None:
            rc = self.lib.fmso_advance_vc(ctypes.byref(self.prob), ctypes.byref(self.st), float(start_time),
                                          float(dt_split), int(n_split), float(dt0D), int(n0D), float(dt1D), int(n1D))
            assert rc == 0, "fmso_advance_vc: the point-buffer sweep takes firing-table stimulation"
            return
        self.lib.fmso_advance(ctypes.byref(self.prob), ctypes.byref(self.st), float(start_time), float(dt_split),
                              int(n_split), float(dt0D), int(n0D), float(dt1D), int(n1D), int(self.threads))

    def advance_schedule(self, steps, n0D=1):
        for row in np.asarray(steps).reshape(-1, 4):
            self.advance(row[0], row[1], 1, row[2], n0D, row[3], 1)

    def get_state(self, k):
        return self.states[k].copy()

    def vm(self):
        return self.states[0].copy()

    def get_states(self):
        return self.states.copy()

    def spikes(self):
        return self._spk["first"].copy(), self._spk["count"].copy()

    def stimulation_log(self):
        n = int(self._log["n"][0])
        return self._log["t"][:n].copy(), self._log["f"][:n].copy()
