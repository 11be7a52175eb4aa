"""Ahead-of-time build of the fibre libraries for gfx950.

One shared library per cell model (``libfms_<model>.so`` in ``opendihu_amd/lib``),
the counterpart of the reference compiling one library per CellML file at run
time (core/src/specialized_solver/fast_monodomain_solver/
fast_monodomain_solver_gpu.tpp:12-206). The hand-written Hodgkin-Huxley model
is always built; models generated from OpenCOR C files by
``opendihu_amd.cellml`` are built with :func:`build_model_library`.
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")
GEN_DIR = os.path.join(PKG_DIR, "_gen")
INCLUDE = os.path.join(REPO_DIR, "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FMS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every + - * / rounds as written, as in the reference's
# g++ -std=c++17 build of its generated code (bit-exact parity).
# -packed-fp32-ops: no packed FP32 instructions are generated.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# Large models (more than 8 states): the machine scheduler that maximises
# instruction-level parallelism for one wave per SIMD. Their kernels run at one
# wave per SIMD anyway (400+ VGPR+AGPR per lane), and with the default
# occupancy-driven scheduler the fused kernel of the 56-state model computed
# wrong and run-to-run different states on gfx950 (hipcc of ROCm 7.2), while
# the same source scheduled for ILP is bit-identical to the oracle
# (tests/test_gpu_models.py) and faster. Small models keep the default
# scheduler, which is faster for them (occupancy matters there).
LARGE_MODEL_FLAGS = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
LARGE_MODEL_STATES = 8

HH_LIBRARY = "libfms_hodgkin_huxley_1952.so"


def _sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".h", ".hip")))


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd)}):\n{out}")


def compile_library(out_name: str, model_header: str | None = None, model_type: str | None = None,
                    extra_include: str | None = None, n_states: int = 4, ilp_schedule: bool | None = None) -> str:
    """Compile fms_lib.hip for one model into ``LIB_DIR/out_name``. ilp_schedule:
    the machine scheduler for ILP (None: for models of more than 8 states)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, out_name)
    if ilp_schedule is None:
        ilp_schedule = n_states > LARGE_MODEL_STATES
    flags = HIP_FLAGS + (LARGE_MODEL_FLAGS if ilp_schedule else [])
    cmd = [HIPCC, *flags, f"-I{INCLUDE}", f"-I{CSRC}"]
    if extra_include:
        cmd.append(f"-I{extra_include}")
    if model_header:
        cmd += [f'-DFMS_MODEL_HEADER="{model_header}"', f"-DFMS_MODEL_TYPE={model_type}"]
    tmp = out + ".tmp"
    cmd += ["-o", tmp, os.path.join(CSRC, "fms_lib.hip")]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_hh() -> str:
    return compile_library(HH_LIBRARY)


def build_model_library(model_c_file: str, parameters_used_as_algebraic=(), parameters_used_as_constant=(),
                        name: str | None = None, lib_name: str | None = None, ilp_schedule: bool | None = None) -> str:
    """Generate HIP code for an OpenCOR C model file and compile its library."""
    from .cellml import generate_hip_model

    os.makedirs(GEN_DIR, exist_ok=True)
    with open(model_c_file) as f:
        text = f.read()
    base = name or os.path.splitext(os.path.basename(model_c_file))[0]
    key = hashlib.sha1((text + repr(tuple(parameters_used_as_algebraic)) + repr(tuple(parameters_used_as_constant)))
                       .encode()).hexdigest()[:8]
    header_name = f"model_{base}_{key}.h"
    type_name = f"fms::Model_{base}_{key}".replace("-", "_")
    code = generate_hip_model(text, base, list(parameters_used_as_algebraic), list(parameters_used_as_constant),
                              struct_name=type_name.split("::")[1])
    with open(os.path.join(GEN_DIR, header_name), "w") as f:
        f.write(code)
    from .cellml import parse_opencor_c
    n_states = parse_opencor_c(text).n_states
    return compile_library(lib_name or f"libfms_{base}.so", header_name, type_name, extra_include=GEN_DIR,
                           n_states=n_states, ilp_schedule=ilp_schedule)


# CellML models shipped with the reference's tests (testing/unit_testing/input),
# with the parameter mappings the reference's settings use for them:
#   Shorten 2007: parametersUsedAsAlgebraic=[32] (I_Stim), parametersUsedAsConstant=[65] (l_hs)
#     (testing/unit_testing/src/1_rank/cellml.cpp:193-195)
#   Hodgkin-Huxley: parametersUsedAsConstant=[2,3,4,5] (i_Stim, g_Na, g_K, g_L); the
#     generated library cross-checks the generator against the hand-written model
REFERENCE_MODELS = [
    ("testing/unit_testing/input/shorten_ocallaghan_davidson_soboleva_2007.c",
     "shorten_ocallaghan_davidson_soboleva_2007", [32], [65], None),
    ("testing/unit_testing/input/hodgkin_huxley_1952.c", "hodgkin_huxley_1952", [], [2, 3, 4, 5],
     "libfms_gen_hodgkin_huxley_1952.so"),
]
SHORTEN_LIBRARY = "libfms_shorten_ocallaghan_davidson_soboleva_2007.so"
GEN_HH_LIBRARY = "libfms_gen_hodgkin_huxley_1952.so"

# Models the reference ships only as hand-written Vc kernels, converted to the
# OpenCOR form by opendihu_amd.cellml.vc_kernel (no parameters: the kernels
# take none, stimulation clamps Vm):
#   TK2014, 57 states / 71 algebraics (the kernel's new_slow_TK_2014_12_08).
#   Built with the default machine scheduler: under the ILP scheduler its
#   per-fibre reaction kernel (k_reaction_fibre) computed wrong states that
#   changed from run to run (NaN in one run, rates off by 100 % in another) on
#   gfx950 with hipcc of ROCm 7.2, while the default scheduler's build is
#   bit-identical to the oracle in every kernel variant
#   (tests/test_gpu_models.py).
VC_KERNEL_MODELS = [
    ("core/src/specialized_solver/fast_monodomain_solver/tk2014.cpp.disabled", "tk2014", "libfms_tk2014.so", False),
]
TK2014_LIBRARY = "libfms_tk2014.so"


def build_vc_kernel_library(kernel_file: str, name: str, lib_name: str, ilp_schedule: bool | None = None) -> str:
    """Convert a hand-written Vc kernel to the OpenCOR form and compile its library."""
    from .cellml import convert_vc_kernel
    os.makedirs(GEN_DIR, exist_ok=True)
    with open(kernel_file) as f:
        text = convert_vc_kernel(f.read(), name)
    path = os.path.join(GEN_DIR, f"{name}_opencor.c")
    with open(path, "w") as f:
        f.write(text)
    return build_model_library(path, [], [], name, lib_name, ilp_schedule)


def build_reference_models(reference_root: str) -> list[str]:
    """Libraries for the CellML models shipped in the reference's test inputs
    (compiled concurrently, one hipcc process per model)."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = [(os.path.join(reference_root, rel), pa, pc, name, lib)
            for rel, name, pa, pc, lib in REFERENCE_MODELS if os.path.exists(os.path.join(reference_root, rel))]
    vc_jobs = [(os.path.join(reference_root, rel), name, lib, ilp)
               for rel, name, lib, ilp in VC_KERNEL_MODELS if os.path.exists(os.path.join(reference_root, rel))]
    with ThreadPoolExecutor(max_workers=max(1, len(jobs) + len(vc_jobs))) as ex:
        futs = [ex.submit(build_model_library, path, pa, pc, name, lib) for path, pa, pc, name, lib in jobs]
        futs += [ex.submit(build_vc_kernel_library, path, name, lib, ilp) for path, name, lib, ilp in vc_jobs]
        return [f.result() for f in futs]


def main(argv=None) -> int:
    print("building", build_hh(), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
