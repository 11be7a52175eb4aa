"""Gradients of a Vm-trace misfit with respect to per-fibre conductances and
conductivity, through the hand-written discrete adjoint of the split step
(opendihu_amd/csrc/fms_adjoint.h), exposed as a torch.autograd.Function.

    theta[f] = (g_Na, g_K, g_L, sigma)          per fibre, FP64
    traces = MonodomainTraces.apply(theta, problem)   (n_steps, n_fibers, n_sensors)

``problem`` (:class:`AdjointProblem`) holds the fibre batch (Hodgkin-Huxley,
current-pulse stimulation), the initial states, the split-step table, the
sensor nodes, the checkpoint interval and where the tape lives: the forward
records the whole tape in device memory (72 B per node and split step) and the
backward only reverses it, or the forward keeps a checkpoint per interval and
the backward recomputes each segment's tape. Forward and backward run on the GPU;
with theta on the GPU the traces and their adjoints stay in device memory, with
theta on the CPU they are CPU float64 tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .fibers import FiberBatch


class AdjointProblem:
    TAPES = {"auto": L.FMS_ADJOINT_TAPE_AUTO, "segments": L.FMS_ADJOINT_TAPE_SEGMENTS,
             "whole": L.FMS_ADJOINT_TAPE_WHOLE}

    def __init__(self, batch: FiberBatch, steps, sensor_nodes, checkpoint_interval=100, n0D=1, am=500.0, cm=1.0,
                 tape: str = "auto"):
        """tape: "whole" (the forward records the whole tape), "segments"
        (checkpoints; the backward recomputes each segment) or "auto" (the whole
        tape when it fits in half of the free device memory)"""
        if batch.model_name != "hodgkin_huxley_1952":
            raise ValueError("the adjoint is implemented for the Hodgkin-Huxley model")
        self.batch = batch
        self.steps = np.ascontiguousarray(steps, np.float64).reshape(-1, 4)
        self.n_steps = self.steps.shape[0]
        self.sensor_nodes = np.ascontiguousarray(sensor_nodes, np.int32).reshape(batch.n_fibers, -1)
        self.n_sensors = self.sensor_nodes.shape[1]
        self.interval = int(checkpoint_interval)
        self.n0D = int(n0D)
        self.am_cm = am * cm
        self.initial = batch.get_states()
        self.base_params = batch.parameters.copy()
        # forwards run so far: the solver keeps the checkpoints, step table and
        # parameters of the last one only, which the backward reverses
        self.generation = 0
        self.set_tape(tape)
        lib = batch.lib
        L.check(lib, lib.fms_set_sensors(batch.handle, L.as_ptr(self.sensor_nodes, L._ip), self.n_sensors),
                "fms_set_sensors")

    def set_theta(self, theta: np.ndarray) -> None:
        theta = np.asarray(theta, np.float64).reshape(self.batch.n_fibers, 4)
        params = self.base_params.copy()
        params[:, 1:4] = theta[:, 0:3]
        pf = np.ascontiguousarray(theta[:, 3] / self.am_cm)
        params = np.ascontiguousarray(params)
        L.check(self.batch.lib, self.batch.lib.fms_set_parameters(self.batch.handle, L.as_ptr(params, L._dp),
                                                                  L.as_ptr(pf, L._dp)), "fms_set_parameters")

    def forward(self, theta) -> np.ndarray:
        self.generation += 1
        self.set_theta(theta)
        for k in range(self.batch.n_states):
            self.batch.set_state(k, self.initial[k])
        out = np.empty((self.n_steps, self.batch.n_fibers, self.n_sensors))
        lib = self.batch.lib
        L.check(lib, lib.fms_adjoint_forward(self.batch.handle, L.as_ptr(self.steps, L._dp), self.n_steps, self.n0D,
                                             self.interval, L.as_ptr(out, L._dp)), "fms_adjoint_forward")
        return out

    def forward_tensor(self, theta):
        """forward with the traces in a torch tensor on theta's device (no host
        copy of the traces when theta is a GPU tensor)"""
        torch = _torch()
        if not theta.is_cuda:
            return torch.from_numpy(self.forward(theta.detach().numpy()))
        self.generation += 1
        self.set_theta(theta.detach().cpu().numpy())
        for k in range(self.batch.n_states):
            self.batch.set_state(k, self.initial[k])
        out = torch.empty((self.n_steps, self.batch.n_fibers, self.n_sensors), dtype=torch.float64,
                          device=theta.device)
        torch.cuda.synchronize(theta.device)
        lib = self.batch.lib
        L.check(lib, lib.fms_adjoint_forward_device(self.batch.handle, L.as_ptr(self.steps, L._dp), self.n_steps,
                                                    self.n0D, self.interval, ctypes.c_void_p(out.data_ptr())),
                "fms_adjoint_forward_device")
        return out

    def backward_tensor(self, dtraces):
        """dL/dtheta for the last forward as a torch tensor on dtraces' device"""
        torch = _torch()
        if not dtraces.is_cuda:
            return torch.from_numpy(self.backward(dtraces.detach().contiguous().numpy()))
        d = dtraces.detach().to(torch.float64).contiguous()
        F = self.batch.n_fibers
        gp = torch.empty((F, 3), dtype=torch.float64, device=d.device)
        gpf = torch.empty(F, dtype=torch.float64, device=d.device)
        torch.cuda.synchronize(d.device)
        lib = self.batch.lib
        L.check(lib, lib.fms_adjoint_backward_device(self.batch.handle, ctypes.c_void_p(d.data_ptr()),
                                                     ctypes.c_void_p(gp.data_ptr()), ctypes.c_void_p(gpf.data_ptr())),
                "fms_adjoint_backward_device")
        return torch.cat([gp, (gpf / self.am_cm)[:, None]], dim=1)

    def set_tape(self, tape: str) -> None:
        """where the next forward keeps the tape ("whole", "segments" or "auto")"""
        if tape not in self.TAPES:
            raise ValueError(f"tape is one of {sorted(self.TAPES)}")
        self.batch.set_tuning(adjoint_tape=self.TAPES[tape])

    def whole_tape(self) -> bool:
        """whether the last forward recorded the whole tape (else checkpoints)"""
        w = ctypes.c_int(0)
        L.check(self.batch.lib, self.batch.lib.fms_adjoint_whole_tape(self.batch.handle, ctypes.byref(w)),
                "fms_adjoint_whole_tape")
        return bool(w.value)

    def tune_backward(self, theta, dtraces, threads=(256, 512)) -> dict:
        """time fms_adjoint_backward for one forward at theta under each
        workgroup size of the reverse sweep and, with checkpoints, each schedule
        of the segment recompute (before the reverse sweep, or beside it with
        equal priority, the sweep's or the recompute's priority raised), and
        keep the fastest. The schedules give the same gradient; the workgroup
        size changes the order of the per-fibre prefactor-gradient sum (last
        bits)."""
        import time
        from . import _lib
        times = {}
        self.forward(theta)
        whole = self.whole_tape()
        names = {0: "overlap", 1: "overlap_priority", 2: "serial", 3: "overlap_recompute_priority"}
        if whole:
            names = {2: "whole_tape"}
        for tt in threads:
            for mode in ((_lib.FMS_ADJOINT_SERIAL,) if whole else
                         (_lib.FMS_ADJOINT_SERIAL, _lib.FMS_ADJOINT_OVERLAP, _lib.FMS_ADJOINT_OVERLAP_PRIORITY,
                          _lib.FMS_ADJOINT_OVERLAP_RECOMPUTE_PRIORITY)):
                self.batch.set_tuning(adjoint=mode, adjoint_threads=tt)
                self.backward(dtraces)   # untimed: buffers, code objects
                t0 = time.perf_counter()
                self.backward(dtraces)
                times[(tt, mode)] = time.perf_counter() - t0
        best = min(times, key=times.get)
        self.batch.set_tuning(adjoint_threads=best[0], adjoint=best[1])
        return dict(schedule=names[best[1]], threads=best[0], tape="whole" if whole else "segments",
                    seconds={f"{names[m]}/{tt}": v for (tt, m), v in times.items()})

    def backward(self, dtraces) -> np.ndarray:
        """dL/dtheta for the last forward; theta columns (g_Na, g_K, g_L, sigma)"""
        d = np.ascontiguousarray(dtraces, np.float64)
        F = self.batch.n_fibers
        gp = np.empty((F, 3))
        gpf = np.empty(F)
        lib = self.batch.lib
        L.check(lib, lib.fms_adjoint_backward(self.batch.handle, L.as_ptr(d, L._dp), L.as_ptr(gp, L._dp),
                                              L.as_ptr(gpf, L._dp)), "fms_adjoint_backward")
        return np.concatenate([gp, (gpf / self.am_cm)[:, None]], axis=1)


def _torch():
    import torch
    return torch


class MonodomainTraces:
    """torch.autograd.Function: theta (n_fibers, 4) -> sensor traces (n_steps, n_fibers, n_sensors)"""

    _fn = None

    @classmethod
    def apply(cls, theta, problem: AdjointProblem):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, th, prob):
                    ctx.prob = prob
                    out = prob.forward_tensor(th)
                    ctx.generation = prob.generation
                    ctx.theta = th.detach().clone()
                    return out

                @staticmethod
                def backward(ctx, g):
                    prob = ctx.prob
                    if prob.generation != ctx.generation:
                        # another forward ran on the solver since this output
                        # was made (two outputs in one loss, accumulation,
                        # gradcheck): recompute this output's forward so the
                        # reverse sweep uses its own checkpoints and parameters
                        prob.forward_tensor(ctx.theta)
                    return prob.backward_tensor(g), None

            cls._fn = _Fn
        return cls._fn.apply(theta, problem)
