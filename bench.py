"""Throughput of the MI355X FastMonodomainSolver fibre path (fiber-node-steps/s).

Default workload (one GPU): the Shorten-type muscle model (56 states / 71
algebraics, the reference's OpenCOR CellML input) on 1024 equal fibres x 1481
nodes (14.8 cm), Am=500 /cm, Cm=1 uF/cm^2, sigma=3.828 mS/cm, Heun reaction +
implicit-Euler Thomas diffusion, dt_0D=dt_1D=dt_split=2.5e-5 ms, fibres in 10
motor units (seeded uniform draw) firing at seeded Poisson times (20 Hz) at the
centre node. A step is one Strang split step over all fibres; inputs are
resident in HBM when the timed region starts.

The JSON line carries:
  value            FP64 fiber-node-steps/s over the timed steps
  roofline         the dominant kernel (the kernel class with the largest summed
                   launch time: on the default workload the paired per-fibre
                   reaction) against the FP64 vector peak (the kernels are FP64
                   VALU work): algorithmic ops per launch (the model's op count,
                   each + - * / exp log = 1, per node and Heun step x nodes x
                   Heun steps in the launch) / mean launch time from a start/stop
                   event pair bound to each launch; "hbm" holds the same launch
                   against the HBM roof with the algorithmic bytes (every state
                   read and written once), "traffic" the rocprofv3-measured bytes
  cpu_baseline     the FP64 CPU oracle (the reference's vc-path restatement, four
                   nodes per vector operation as Vc's double_v on AVX2) on the host
                   cores, over a bounded sample of the same workload
  mixed            the FP32-state / FP64-Vm form on the same workload
  workloads        batched uneven HH under the autotuned schedule (FP64, mixed,
                   spike-time error, CPU oracle), the adjoint's gradient cost
                   against one forward solve, the large mixed-precision
                   Shorten batch (16384 uneven fibres) against its FP64 run, and
                   the 57-state TK2014 model on the default batch (FP64, mixed,
                   CPU oracle)
usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c3|c1|tk2014] [--no-extras]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"

# HBM bytes per node and launch of the dominant kernel, measured with rocprofv3
# (FETCH_SIZE and WRITE_SIZE scaled by what they read for a 1 GiB device copy),
# one record per (workload, kernel class, threads per fibre of the per-fibre
# reaction) in profiles/traffic.json, each naming the profile summary it comes from
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic.json")


def measured_traffic(workload: str, kclass: str, threads: int, nodes: int, node_steps_per_launch: float):
    """(HBM bytes per launch, source) of the dominant kernel from its record
    (bytes per node and launch, or per node-step for a fused launch of several
    split steps), or (None, None)"""
    try:
        with open(TRAFFIC_FILE) as f:
            for r in json.load(f)["records"]:
                if (r["workload"], r["kernel_class"], r["threads"]) == (workload, kclass, threads):
                    if "bytes_per_node_step" in r:
                        return r["bytes_per_node_step"] * node_steps_per_launch, r["source"]
                    return r["bytes_per_node_launch"] * nodes, r["source"]
    except (OSError, ValueError, KeyError):
        pass
    return None, None


def bytes_per_node_step(n_states: int, precision: str) -> int:
    """Algorithmic HBM bytes per fiber-node-step: every state read and written once."""
    if precision == "fp64":
        return 2 * 8 * n_states
    return 2 * 8 + 2 * 4 * (n_states - 1)


def step_table(t0, dt, n):
    """n split-step records (startTime, dtSplit, dt0D, dt1D): one Heun step of
    dt/2 per reaction half, the reference's step-count rule for dt0D = dt
    (00_time_stepping_scheme.cpp:19-42)"""
    return np.stack([t0 + np.arange(n) * dt, np.full(n, dt), np.full(n, dt / 2), np.full(n, dt)], 1)


def make_workload(name: str, precision: str, kernels: str):
    kernels = "fused" if kernels == "auto" else kernels
    from cases import c2, uneven, c1
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import SHORTEN_LIBRARY, HH_LIBRARY
    if name == "c2":
        mesh, stim = c2()
        g = FiberBatch(**mesh, scheme="IE", precision=precision, kernels=kernels, thomas="partitioned",
                       library=SHORTEN_LIBRARY)
        g.set_firing_table_stimulation(**stim)
        desc = dict(workload="shorten56_1024x1481_IE_dt2.5e-5", model=SHORTEN, fibers=1024, nodes_per_fiber=1481,
                    scheme="ImplicitEuler", dt_split_ms=2.5e-5, motor_units=10, firing_hz=20,
                    motor_units_firing_in_5ms=int(stim["firing_events"].any(axis=0).sum()),
                    stimulation="firing table, Vm clamp (reference vc semantics, W=4)",
                    work="every node's reaction and diffusion every step (onlyComputeIfHasBeenStimulated and "
                         "the equilibrium skip off)")
        return g, mesh, stim, desc, 2.5e-5, "IE", SHORTEN
    if name == "tk2014":
        from opendihu_amd.build import TK2014_LIBRARY
        mesh, stim = c2()
        g = FiberBatch(**mesh, scheme="IE", precision=precision, kernels=kernels, thomas="partitioned",
                       library=TK2014_LIBRARY)
        g.set_firing_table_stimulation(**stim)
        desc = dict(workload="tk2014_57_1024x1481_IE_dt2.5e-5", model="tk2014 (57 states, 71 algebraics)",
                    fibers=1024, nodes_per_fiber=1481, scheme="ImplicitEuler", dt_split_ms=2.5e-5, motor_units=10,
                    firing_hz=20, motor_units_firing_in_5ms=int(stim["firing_events"].any(axis=0).sum()),
                    stimulation="firing table, Vm clamp (reference vc semantics, W=4)")
        return g, mesh, stim, desc, 2.5e-5, "IE", "tk2014"
    if name == "c3":
        mesh, stim = uneven(4096, seed=3)
        g = FiberBatch(**mesh, scheme="CN", precision=precision, kernels=kernels, thomas="partitioned",
                       library=HH_LIBRARY)
        g.set_current_stimulation(**stim)
        desc = dict(workload="hh_4096_uneven_400-2000_CN_dt1e-3", model="hodgkin_huxley_1952", fibers=4096,
                    nodes=int(mesh["n_nodes"].sum()), scheme="CrankNicolson", dt_split_ms=1e-3)
        return g, mesh, stim, desc, 1e-3, "CN", "hh"
    if name == "c1":
        mesh, stim = c1()
        g = FiberBatch(**mesh, scheme="CN", precision=precision, kernels=kernels, thomas="partitioned",
                       library=HH_LIBRARY)
        g.set_current_stimulation(**stim)
        desc = dict(workload="hh_1x1001_CN_dt1e-3", model="hodgkin_huxley_1952", fibers=1, nodes_per_fiber=1001)
        return g, mesh, stim, desc, 1e-3, "CN", "hh"
    raise SystemExit(f"unknown workload {name}")


def time_gpu(g, dt, steps, warmup):
    g.advance(0.0, dt, warmup)
    g.synchronize()
    c0 = g.launch_counts()
    t_start = warmup * dt
    g.record_event(0)
    w0 = time.perf_counter()
    g.advance(t_start, dt, steps)
    g.record_event(1)
    g.synchronize()
    wall = time.perf_counter() - w0
    ev_ms = g.elapsed_ms(0, 1)
    c1 = g.launch_counts()
    launches = {k: c1[k] - c0[k] for k in c1}
    return wall, ev_ms, launches


def host_cpu() -> dict:
    """the host's CPU model, logical CPUs, the CPUs this process may run on, and AVX-512"""
    model, avx512 = "unknown", False
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                if line.startswith("flags") and " avx512f" in line:
                    avx512 = True
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return dict(model=model, logical_cpus=os.cpu_count(), allowed_cpus=allowed, avx512=avx512)


def cpu_baseline(mesh, stim, scheme, model, dt, budget_s=10.0):
    """FP64 oracle (the reference's vc-path restatement) on the host over a
    bounded sample of the workload: W = 4 nodes per vector operation (the
    reference's Vc double_v on AVX2) and, on an AVX-512 host, W = 8; value is the
    faster. Threads: every CPU this process may run on, capped by OMP_NUM_THREADS
    where that is set (the CPU share of one GPU on a shared host)."""
    from oracle_binding import OracleBatch, ORACLE_BUILD
    cpu = host_cpu()
    threads = cpu["allowed_cpus"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = max(1, min(threads, int(os.environ["OMP_NUM_THREADS"])))
    variants = [("", 4)]
    if cpu["avx512"] and os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{model}_w8.so")):
        variants.append(("_w8", 8))
    runs = {}
    for suffix, width in variants:
        o = OracleBatch(**mesh, scheme=scheme, model=model + suffix, threads=threads)
        if "firing_events" in stim:
            o.set_firing_table_stimulation(**stim)
        else:
            o.set_current_stimulation(**stim)
        n_total = o.n_total
        t0 = time.perf_counter()
        o.advance(0.0, dt, 1)
        probe = time.perf_counter() - t0
        steps = max(1, min(2000, int(budget_s / len(variants) / max(probe, 1e-6))))
        t0 = time.perf_counter()
        o.advance(dt, dt, steps)
        el = time.perf_counter() - t0
        runs[width] = dict(value=n_total * steps / el, steps=steps, seconds=el)
        del o
    best = max(runs, key=lambda w: runs[w]["value"])
    n_total = int(np.sum(mesh["n_nodes"]))
    return dict(value=runs[best]["value"], unit="fiber-node-steps/s", cores=threads, kind="port",
                lanes=best, by_lanes={f"W{w}": r["value"] for w, r in runs.items()}, host=cpu,
                sample=f"all {len(mesh['n_nodes'])} fibres ({n_total} nodes) x "
                       + ", ".join(f"{r['steps']} split steps ({r['seconds']:.1f} s, W={w})" for w, r in runs.items())
                       + f"; FP64 oracle, W nodes per vector operation as the reference's Vc path, {threads} "
                       f"OpenMP threads on {cpu['model']} ({cpu['allowed_cpus']} CPUs allowed, "
                       f"{cpu['logical_cpus']} logical)")


def _spike_shift(first_a, first_b, n_steps):
    """first-spike step differences: largest shift over nodes that fired in both
    runs, and the nodes that fired in one run only (with the earliest such step,
    which shows whether they are crossings at the end of the window)"""
    both = (first_a >= 0) & (first_b >= 0)
    one = (first_a >= 0) != (first_b >= 0)
    shift = int(np.max(np.abs(first_a[both] - first_b[both]))) if both.any() else 0
    only = np.where(first_a >= 0, first_a, first_b)[one]
    return dict(spike_time_error_steps=shift, nodes_fired_both=int(both.sum()), nodes_fired_one_run_only=int(one.sum()),
                earliest_one_run_only_step=int(only.min()) if one.any() else None, steps=n_steps)


def bench_c3(cpu=True):
    """Batched uneven HH (4096 fibres of 400-2000 nodes, CN, dt=1e-3 ms, 10 ms):
    schedule chosen by the autotuner, FP64 and mixed, spike-time error of the
    mixed run against the FP64 run, FP64 CPU oracle on a bounded sample."""
    from cases import uneven
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import HH_LIBRARY
    dt, n_steps = 1e-3, 10000
    mesh, stim = uneven(4096, seed=3)
    # as long as one fused launch of the timed run: shorter trials understate the
    # fused kernel (0.181 ms/step in 64-step launches, 0.166 in 512-step ones)
    trial = step_table(0.0, dt, 512)
    res = {}
    for precision in ("fp64", "mixed"):
        g = FiberBatch(**mesh, scheme="CN", precision=precision, kernels="fused", thomas="partitioned",
                       library=HH_LIBRARY)
        g.set_current_stimulation(**stim)
        g.enable_spike_detection(0.0)
        choice = g.autotune(trial, allow_split=True)
        g.synchronize()
        g.record_event(0)
        w0 = time.perf_counter()
        g.advance(0.0, dt, n_steps)
        g.record_event(1)
        g.synchronize()
        wall = time.perf_counter() - w0
        res[precision] = dict(value=g.n_total * n_steps / wall, ms_per_step=wall / n_steps * 1e3,
                              event_ms=g.elapsed_ms(0, 1), schedule=choice, vm=g.vm(), first=g.spikes()[0])
        g.close()
    spikes = _spike_shift(res["mixed"]["first"], res["fp64"]["first"], n_steps)
    vm_err = float(np.max(np.abs(res["mixed"]["vm"] - res["fp64"]["vm"])))
    out = dict(workload="hh_4096_uneven_400-2000_CN_dt1e-3_10ms", nodes=int(mesh["n_nodes"].sum()), steps=n_steps,
               value_fp64=res["fp64"]["value"], value_mixed=res["mixed"]["value"],
               ms_per_step_fp64=res["fp64"]["ms_per_step"], ms_per_step_mixed=res["mixed"]["ms_per_step"],
               schedule_fp64={k: v for k, v in res["fp64"]["schedule"].items()},
               schedule_mixed={k: v for k, v in res["mixed"]["schedule"].items()},
               mixed_vs_fp64=dict(spikes, vm_max_abs_err_mV=vm_err))
    if cpu:
        c = cpu_baseline(mesh, stim, "CN", "hh", dt, budget_s=10.0)
        out["cpu_baseline"] = c
        out["speedup_mixed_vs_cpu"] = res["mixed"]["value"] / c["value"]
        out["speedup_fp64_vs_cpu"] = res["fp64"]["value"] / c["value"]
    return out


def bench_c4():
    """Adjoint: HH on 256 fibres x 1000 nodes, 5 ms at dt=1e-3 ms (5000 split
    steps), 32 sensors per fibre. Cost of the gradient (recording forward +
    backward) against one plain forward solve, with the whole tape in device
    memory (92 GB: the forward records it, the backward only reverses) and with
    a checkpoint every 100 steps (the backward recomputes each segment's tape),
    and one directional central difference (h=1e-6) at full size."""
    from cases import uneven, AM, CM
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.adjoint import AdjointProblem
    F, n_nodes, n_steps, dt = 256, 1000, 5000, 1e-3
    mesh, stim = uneven(F, lo=n_nodes, hi=n_nodes, seed=41)
    b = FiberBatch(**mesh, scheme="CN", kernels="fused", thomas="partitioned")
    b.set_current_stimulation(**stim)
    steps = step_table(0.0, dt, n_steps)
    sensors = np.tile(np.round(np.linspace(0, n_nodes - 1, 32)).astype(np.int32), (F, 1))
    prob = AdjointProblem(b, steps, sensors, checkpoint_interval=100, am=AM, cm=CM)
    rng = np.random.default_rng(42)
    sigma = mesh["prefactor"] * AM * CM
    theta_true = np.stack([rng.uniform(100, 140, F), rng.uniform(30, 42, F), rng.uniform(0.2, 0.4, F), sigma], 1)
    theta0 = theta_true * rng.uniform(0.95, 1.05, theta_true.shape)
    import torch
    dev = torch.device("cuda")
    target = prob.forward_tensor(torch.tensor(theta_true, device=dev))
    th0 = torch.tensor(theta0, device=dev)
    # warm-up of every path (code objects, buffers); the reverse sweep's
    # workgroup size (and with checkpoints the recompute's schedule) timed and
    # chosen here
    # the whole tape when it fits in half of the free device memory (92 GB: on
    # one MI355X it does), else checkpoints
    prob.set_tape("auto")
    prob.backward_tensor(prob.forward_tensor(th0) - target)
    whole = prob.whole_tape()
    tuned = prob.tune_backward(theta0, (prob.forward(theta0) - target.cpu().numpy()))
    # each phase three times, the fastest counted (all are reported)
    reps = 3
    # plain forward solve (no traces, no checkpoints) at theta0
    plain = []
    for _ in range(reps):
        prob.set_theta(theta0)
        for k in range(b.n_states):
            b.set_state(k, prob.initial[k])
        b.synchronize()
        w0 = time.perf_counter()
        b.advance_schedule(steps)
        b.synchronize()
        plain.append(time.perf_counter() - w0)
    # gradient: recording forward + backward, traces and adjoints in device memory
    def gradient_times():
        fwd, bwd = [], []
        for _ in range(reps):
            w0 = time.perf_counter()
            tr = prob.forward_tensor(th0)
            torch.cuda.synchronize()
            fwd.append(time.perf_counter() - w0)
            resid = tr - target
            torch.cuda.synchronize()
            w0 = time.perf_counter()
            grad = prob.backward_tensor(resid).cpu().numpy()
            bwd.append(time.perf_counter() - w0)
        return fwd, bwd, grad
    fwd, bwd, grad = gradient_times()
    t_plain, t_fwd, t_bwd = min(plain), min(fwd), min(bwd)
    # the same with checkpoints: the backward recomputes each segment
    prob.set_tape("segments")
    tuned_ck = prob.tune_backward(theta0, (prob.forward(theta0) - target.cpu().numpy()))
    fwd_ck, bwd_ck, _ = gradient_times()
    assert not prob.whole_tape()
    # the same gradient as the whole tape's, bit for bit, at its reverse-sweep size
    prob.batch.set_tuning(adjoint_threads=tuned["threads"])
    grad_ck = prob.backward_tensor(prob.forward_tensor(th0) - target).cpu().numpy()
    prob.set_tape("auto")
    loss = lambda th: 0.5 * float(((prob.forward_tensor(torch.tensor(th, device=dev)) - target) ** 2).sum())  # noqa: E731
    v = np.random.default_rng(43).standard_normal(theta0.shape) * theta0
    h = 1e-6
    fd = (loss(theta0 + h * v) - loss(theta0 - h * v)) / (2 * h)
    fd2 = (loss(theta0 + 2 * h * v) - loss(theta0 - 2 * h * v)) / (4 * h)
    fdr = (4 * fd - fd2) / 3
    ad = float(np.sum(grad * v))
    b.close()
    ck = dict(recording_forward_s=min(fwd_ck), backward_s=min(bwd_ck),
              repetitions_s=dict(recording_forward=fwd_ck, backward=bwd_ck),
              gradient_cost_vs_forward=(min(fwd_ck) + min(bwd_ck)) / t_plain, adjoint_schedule=tuned_ck,
              max_abs_grad_difference_vs_whole_tape=float(np.max(np.abs(grad_ck - grad))))
    return dict(workload="hh_adjoint_256x1000_CN_dt1e-3_5ms_32sensors", steps=n_steps,
                tape="whole (92 GB)" if whole else "checkpoints every 100 steps (the whole tape did not fit)",
                plain_forward_s=t_plain, recording_forward_s=t_fwd, backward_s=t_bwd,
                repetitions_s=dict(plain_forward=plain, recording_forward=fwd, backward=bwd),
                gradient_cost_vs_forward=(t_fwd + t_bwd) / t_plain, backward_vs_forward=t_bwd / t_plain,
                adjoint_schedule=tuned, checkpoints_every_100_steps=ck,
                directional_derivative=dict(adjoint=ad, central_fd=fd, h=h, rel_err=abs(ad - fd) / abs(fd),
                                            central_fd_2h=fd2, richardson=fdr, rel_err_richardson=abs(ad - fdr) / abs(fdr)),
                note="wall-clock per call, fastest of three; traces and their adjoints stay in device memory (torch tensors)")


def bench_c5(steps=200, warmup=20):
    """Large mixed-precision batch: Shorten on 16384 fibres of U{1000..1481}
    nodes, 100 motor units (Poisson 10-40 Hz, every unit firing at t=0),
    IE, dt=2.5e-5 ms; bounded to `steps` timed split steps; the mixed run's Vm
    and spike-time error against the FP64 run after the same steps."""
    from cases import c5
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import SHORTEN_LIBRARY
    dt = 2.5e-5
    mesh, stim = c5(n_rows=steps + warmup + 1)
    res = {}
    for precision in ("mixed", "fp64"):
        g = FiberBatch(**mesh, scheme="IE", precision=precision, kernels="fused", thomas="partitioned",
                       library=SHORTEN_LIBRARY)
        g.set_firing_table_stimulation(**stim)
        g.enable_spike_detection(0.0)
        sched = g.autotune(step_table(0.0, dt, 8), allow_split=True)
        g.advance(0.0, dt, warmup)
        g.synchronize()
        w0 = time.perf_counter()
        g.advance(warmup * dt, dt, steps)
        g.synchronize()
        wall = time.perf_counter() - w0
        res[precision] = dict(value=g.n_total * steps / wall, ms_per_step=wall / steps * 1e3, vm=g.vm(),
                              first=g.spikes()[0], schedule=sched)
        g.close()
    spikes = _spike_shift(res["mixed"]["first"], res["fp64"]["first"], steps + warmup)
    return dict(workload="shorten56_16384_uneven_1000-1481_IE_dt2.5e-5_100MU", nodes=int(mesh["n_nodes"].sum()),
                steps=steps, warmup=warmup, value_mixed=res["mixed"]["value"], value_fp64=res["fp64"]["value"],
                ms_per_step_mixed=res["mixed"]["ms_per_step"], ms_per_step_fp64=res["fp64"]["ms_per_step"],
                schedule_mixed=res["mixed"]["schedule"], schedule_fp64=res["fp64"]["schedule"],
                mixed_vs_fp64=dict(spikes, vm_max_abs_err_mV=float(np.max(np.abs(res["mixed"]["vm"] - res["fp64"]["vm"]))),
                                   after_steps=steps + warmup))


def bench_tk2014(cpu=True, steps=1000, warmup=50):
    """The 57-state / 71-algebraic TK2014 model (converted from the reference's
    hand-written Vc kernel) on the headline batch: 1024 fibres x 1481 nodes, IE,
    dt=2.5e-5 ms, 10 motor units at 20 Hz. FP64 and mixed under the autotuned
    schedule, FP64 CPU oracle on a bounded sample."""
    from cases import c2
    from opendihu_amd.fibers import FiberBatch
    from opendihu_amd.build import LIB_DIR, TK2014_LIBRARY
    if not os.path.exists(os.path.join(LIB_DIR, TK2014_LIBRARY)):
        return dict(skipped=f"{TK2014_LIBRARY} not built (the reference checkout is needed at build time)")
    dt = 2.5e-5
    mesh, stim = c2()
    res = {}
    for precision in ("fp64", "mixed"):
        g = FiberBatch(**mesh, scheme="IE", precision=precision, kernels="fused", thomas="partitioned",
                       library=TK2014_LIBRARY)
        g.set_firing_table_stimulation(**stim)
        sched = g.autotune(step_table(0.0, dt, 64), allow_split=True)
        wall, _, _ = time_gpu(g, dt, steps, warmup)
        res[precision] = dict(value=g.n_total * steps / wall, ms_per_step=wall / steps * 1e3, schedule=sched)
        g.close()
    out = dict(workload="tk2014_57_1024x1481_IE_dt2.5e-5", nodes=int(mesh["n_nodes"].sum()), steps=steps,
               warmup=warmup, value_fp64=res["fp64"]["value"], value_mixed=res["mixed"]["value"],
               ms_per_step_fp64=res["fp64"]["ms_per_step"], ms_per_step_mixed=res["mixed"]["ms_per_step"],
               schedule_fp64=res["fp64"]["schedule"], schedule_mixed=res["mixed"]["schedule"])
    if cpu:
        c = cpu_baseline(mesh, stim, "IE", "tk2014", dt)
        out["cpu_baseline"] = c
        out["speedup_fp64_vs_cpu"] = res["fp64"]["value"] / c["value"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c1", "tk2014"])
    ap.add_argument("--kernels", default="auto", choices=["auto", "fused", "split"],
                    help="auto: the autotuner's fastest bit-identical schedule for this batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-HH (autotuned), adjoint and large mixed-precision Shorten lines")
    ap.add_argument("--extras", default="c3,c4,c5,tk2014", help="comma list of extra workloads to run")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the fibre path shards whole fibres across GPUs with no per-step "
              "exchange, and multi-GPU runs are not part of this build (use --gpus 1)", file=sys.stderr)
        return 2

    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8}))
        return 0

    from opendihu_amd import _lib as L
    g, mesh, stim, desc, dt, scheme, oracle_model = make_workload(args.workload, "fp64", args.kernels)
    schedule = None
    if args.kernels == "auto":
        schedule = g.autotune(step_table(0.0, dt, 64), allow_split=True)
    wall, ev_ms, launches = time_gpu(g, dt, args.steps, args.warmup)
    n_total = g.n_total
    value = n_total * args.steps / wall
    # dominant kernel: launch durations from a start/stop event pair bound to each
    # launch on the solver's stream (the kernel's own begin and end), over 64 more
    # steps of the same schedule right after the timed ones (eager: graph replays
    # are not timed per launch)
    g.profile_kernels(True)
    g.set_tuning(graph_steps=0)   # graph replays are not timed per launch; the kernels are the same
    g.kernel_times(reset=True)
    t_prof = (args.warmup + args.steps) * dt
    # 64 split steps, or one whole launch of a fused schedule (its per-step cost
    # depends on the steps per launch)
    fused = (schedule or {}).get("kernels") == "fused"
    prof_steps = max(64, schedule["steps_per_launch"]) if fused else 64
    g.advance(t_prof, dt, prof_steps)
    g.synchronize()
    times = g.kernel_times()
    g.profile_kernels(False)
    kclass = max(("reaction", "fused", "reaction_pair"), key=lambda k: times[k][0])
    total_ms, n_launch = times[kclass]
    mean_launch_ms = total_ms / max(1, n_launch)
    ops = L.model_ops(g.lib)["ops_rhs"]
    if kclass in ("reaction", "reaction_pair"):
        # one launch = one reaction half (n0D = 1 Heun step) over all nodes, or
        # two (the second half of a step and the first of the next) that read
        # and write the states once; unit = one node-half
        halves = 2 if kclass == "reaction_pair" else 1
        per_fibre = (schedule or {}).get("reaction_per_fibre", 1)
        rt = (schedule or {}).get("reaction_threads", 256)
        kernel = f"k_reaction_fibre<{rt} threads>" if g.n_states > 8 and per_fibre else "k_reaction"
        if halves == 2:
            kernel += " (PAIR: two reaction halves per launch)"
        units_per_launch = n_total * halves
        f_unit = 2 * ops + 5 * g.n_states
        b_unit = bytes_per_node_step(g.n_states, "fp64") / halves
    else:
        kernel = "k_fused"
        units_per_launch = n_total * prof_steps / max(1, n_launch)
        f_unit = L.flops_per_node_step(g.lib, g.n_states, 1, scheme)
        b_unit = bytes_per_node_step(g.n_states, "fp64")
    achieved_tflops = f_unit * units_per_launch / (mean_launch_ms * 1e-3) / 1e12
    achieved_gbs = b_unit * units_per_launch / (mean_launch_ms * 1e-3) / 1e9
    rthreads = (schedule or {}).get("reaction_threads", 256)
    node_steps_per_launch = n_total * prof_steps / max(1, n_launch)
    traffic, traffic_source = measured_traffic(args.workload, kclass, rthreads, n_total, node_steps_per_launch)
    kernel_times = {k: dict(total_ms=v[0], launches=v[1]) for k, v in times.items()}
    # the same batch under the reference's default onlyComputeIfHasBeenStimulated
    # (a fibre's reaction starts with its first stimulation): fibres whose motor
    # unit has not fired skip their reaction work; not the headline figure, which
    # computes every node's reaction every step
    g.profile_kernels(False)
    g.set_only_compute_if_stimulated(True)
    t_oc = t_prof + prof_steps * dt
    g.advance(t_oc, dt, 16)
    g.synchronize()
    w0 = time.perf_counter()
    n_oc = min(args.steps, 1000)
    g.advance(t_oc + 16 * dt, dt, n_oc)
    g.synchronize()
    only_if_stimulated = dict(ms_per_step=(time.perf_counter() - w0) / n_oc * 1e3, steps=n_oc,
                              note="reference default onlyComputeIfHasBeenStimulated; no fibre stimulated yet when "
                                   "no motor unit has fired (config.motor_units_firing_in_5ms)")
    g.close()

    mixed = None
    if not args.no_mixed:
        gm, *_ = make_workload(args.workload, "mixed", args.kernels)
        mixed_schedule = gm.autotune(step_table(0.0, dt, 64), allow_split=True) if args.kernels == "auto" else None
        wall_m, ev_m, _ = time_gpu(gm, dt, args.steps, args.warmup)
        mixed = dict(value=n_total * args.steps / wall_m, ms_per_step=wall_m / args.steps * 1e3,
                     bytes_per_node_step=bytes_per_node_step(gm.n_states, "mixed"), dtype="f32 states + f64 Vm/Thomas",
                     schedule=mixed_schedule)
        gm.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mesh, stim, scheme, oracle_model, dt)

    extras = {}
    if not args.no_extras:
        for name in [x for x in args.extras.split(",") if x]:
            t0 = time.perf_counter()
            if name == "c3":
                extras[name] = bench_c3(cpu=not args.no_cpu_baseline)
            elif name == "c4":
                extras[name] = bench_c4()
            elif name == "c5":
                extras[name] = bench_c5()
            elif name == "tk2014":
                extras[name] = bench_tk2014(cpu=not args.no_cpu_baseline)
            else:
                raise SystemExit(f"unknown extra workload {name}")
            extras[name]["bench_wall_s"] = time.perf_counter() - t0
            print(f"# {name} done in {extras[name]['bench_wall_s']:.1f} s", file=sys.stderr, flush=True)

    out = {
        "metric": "fiber-node-steps/s",
        "value": value,
        "unit": "fiber-node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded fibre meshes, motor units and firing times; no checkpoints)",
        "config": dict(desc, kernels=args.kernels, schedule=schedule, thomas="partitioned",
                       exp="approximate (1+x/1024)^1024"),
        "roofline": {
            "bound": "valu",
            "kernel": kernel,
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "ops_per_node_and_launch_unit": f_unit,
            "nodes_x_units_per_launch": units_per_launch,
            "mean_launch_ms": mean_launch_ms,
            "launches_timed": n_launch,
            "kernel_times": dict(kernel_times, split_steps=prof_steps),
            "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved_gbs / HBM_PEAK_GBS, "bytes_per_node_and_launch_unit": b_unit},
        },
        "cpu_baseline": cpu,
        "mixed": mixed,
        "only_if_stimulated": only_if_stimulated,
        "workloads": extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
