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
  roofline         dominant kernel (k_fused) against the FP64 vector peak (the
                   kernel is VALU-bound): algorithmic ops per launch (F_alg per
                   node-step x nodes x steps in the launch; F_alg from the model's
                   op count, each + - * / exp log = 1) / mean launch time from HIP
                   events on the solver's stream; "hbm" holds the same launch
                   against the HBM roof with B_alg bytes per node-step
  cpu_baseline     the FP64 CPU oracle (the reference's vc-path restatement) on the
                   host cores, over a bounded sample of the same workload
  mixed            the FP32-state / FP64-Vm form on the same workload
usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c3|c1]
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

# HBM bytes per fiber-node-step of k_fused measured with rocprofv3 (FETCH_SIZE x 2.000 +
# WRITE_SIZE x 1.000, factors from a 1 GiB copy), profiles/k_fused_shorten56_1024x1481.md
MEASURED_TRAFFIC_PER_NODE_STEP = {("c2", "fused"): 941.9}


def bytes_per_node_step(n_states: int, precision: str) -> int:
    """Algorithmic HBM bytes per fiber-node-step: every state read and written once."""
    if precision == "fp64":
        return 2 * 8 * n_states
    return 2 * 8 + 2 * 4 * (n_states - 1)


def make_workload(name: str, precision: str, kernels: str):
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
                    stimulation="firing table, Vm clamp (reference vc semantics, W=4)")
        return g, mesh, stim, desc, 2.5e-5, "IE", SHORTEN
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


def cpu_baseline(mesh, stim, scheme, model, dt, budget_s=12.0):
    """FP64 oracle on the host cores over a bounded sample of the workload."""
    from oracle_binding import OracleBatch
    threads = min(os.cpu_count() or 1, 16)
    o = OracleBatch(**mesh, scheme=scheme, model=model, threads=threads)
    if "firing_events" in stim:
        o.set_firing_table_stimulation(**stim)
    else:
        o.set_current_stimulation(**stim)
    n_total = o.n_total
    t0 = time.perf_counter()
    o.advance(0.0, dt, 1)
    probe = time.perf_counter() - t0
    steps = max(1, min(2000, int(budget_s / max(probe, 1e-6))))
    t0 = time.perf_counter()
    o.advance(dt, dt, steps)
    el = time.perf_counter() - t0
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(value=n_total * steps / el, unit="fiber-node-steps/s", cores=threads, kind="port",
                sample=f"all {len(mesh['n_nodes'])} fibres ({n_total} nodes) x {steps} split steps "
                       f"({el:.1f} s, FP64 oracle, {threads} OpenMP threads on {cpu})")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c1"])
    ap.add_argument("--kernels", default="fused", choices=["fused", "split"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
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

    g, mesh, stim, desc, dt, scheme, oracle_model = make_workload(args.workload, "fp64", args.kernels)
    wall, ev_ms, launches = time_gpu(g, dt, args.steps, args.warmup)
    n_total = g.n_total
    value = n_total * args.steps / wall
    from opendihu_amd import _lib as L
    b_alg = bytes_per_node_step(g.n_states, "fp64")
    f_alg = L.flops_per_node_step(g.lib, g.n_states, 1, scheme)
    kernel = "k_fused" if args.kernels == "fused" else "k_reaction"
    n_launch = launches["fused"] if args.kernels == "fused" else launches["reaction"]
    mean_launch_ms = ev_ms / max(1, n_launch)
    steps_per_launch = args.steps / max(1, n_launch)
    node_steps_per_launch = n_total * (steps_per_launch if args.kernels == "fused" else 0.5)
    achieved_gbs = b_alg * node_steps_per_launch / (mean_launch_ms * 1e-3) / 1e9
    achieved_tflops = f_alg * node_steps_per_launch / (mean_launch_ms * 1e-3) / 1e12
    g.close()

    mixed = None
    if not args.no_mixed:
        gm, *_ = make_workload(args.workload, "mixed", args.kernels)
        wall_m, ev_m, _ = time_gpu(gm, dt, args.steps, args.warmup)
        mixed = dict(value=n_total * args.steps / wall_m, ms_per_step=wall_m / args.steps * 1e3,
                     bytes_per_node_step=bytes_per_node_step(gm.n_states, "mixed"), dtype="f32 states + f64 Vm/Thomas")
        gm.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mesh, stim, scheme, oracle_model, dt)

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
        "config": dict(desc, kernels=args.kernels, thomas="partitioned", exp="approximate (1+x/1024)^1024"),
        "roofline": {
            "bound": "valu",
            "kernel": kernel,
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": (MEASURED_TRAFFIC_PER_NODE_STEP[(args.workload, args.kernels)] * node_steps_per_launch
                        if (args.workload, args.kernels) in MEASURED_TRAFFIC_PER_NODE_STEP else None),
            "flops_per_node_step": f_alg,
            "launches": n_launch,
            "mean_launch_ms": mean_launch_ms,
            "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved_gbs / HBM_PEAK_GBS, "bytes_per_node_step": b_alg},
        },
        "cpu_baseline": cpu,
        "mixed": mixed,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
