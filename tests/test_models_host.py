"""CPU checks of the product's model front end against the oracle's independent
routes to the same models:

* the generated model headers (opendihu_amd.cellml) compiled for the host with
  g++ give the oracle's rates bit for bit at random states (approximate
  exponential). The oracle pastes the model files' lines under macros; the
  product tokenizes and rewrites them;
* the TK2014 conversion of the reference's hand-written Vc kernel
  (opendihu_amd.cellml.vc_kernel) has the kernel's 57 rates and 71 algebraics,
  and the kernel's second right-hand-side evaluation is its first one applied
  to the intermediate states (the converter reads only the first);
* the accurate logarithm both sides use for TK2014's Vc::log is within 2 ulp of
  the host's log, and the product's and the oracle's are bit-identical.
The reference checkout is needed to generate the models; without it they skip.
"""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

from oracle_binding import ORACLE_BUILD, P, _dp, oracle_library

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "opendihu_amd", "csrc")
REFERENCE = os.environ.get("OPENDIHU_REFERENCE", "/root/reference")
TK_KERNEL = os.path.join(REFERENCE, "core/src/specialized_solver/fast_monodomain_solver/tk2014.cpp.disabled")
SHORTEN_C = os.path.join(REFERENCE, "testing/unit_testing/input/shorten_ocallaghan_davidson_soboleva_2007.c")
HH_C = os.path.join(REFERENCE, "testing/unit_testing/input/hodgkin_huxley_1952.c")
SHORTEN = "shorten_ocallaghan_davidson_soboleva_2007"

DRIVER = r"""
#include "model_under_test.h"
using M = fms::Model_under_test;
extern "C" int ns() { return M::NS; }
extern "C" void rhs_n(const double *Y, const double *Pm, int n, double *R) {
  for (int j = 0; j < n; j++) {
    fms::FD y[M::NS], r[M::NS], a[M::NA], p[M::NP > 0 ? M::NP : 1];
    for (int k = 0; k < M::NS; k++) y[k] = fms::FD(Y[(long)k * n + j]);
    for (int k = 0; k < M::NP; k++) p[k] = fms::FD(Pm[(long)k * n + j]);
    M::rhs<fms::FD, fms::ExpApprox>(y, p, 0.0, r, a);
    for (int k = 0; k < M::NS; k++) R[(long)k * n + j] = r[k].v;
  }
}
extern "C" void log_n(const double *x, double *y, long n) {
  for (long i = 0; i < n; i++) y[i] = fms::log_accurate(x[i]);
}
"""


def _host_model(tmp_path, text, p_alg=(), p_const=()):
    from opendihu_amd.cellml import generate_hip_model
    src = generate_hip_model(text, "under_test", list(p_alg), list(p_const), struct_name="Model_under_test")
    (tmp_path / "model_under_test.h").write_text(src)
    (tmp_path / "driver.cpp").write_text(DRIVER)
    so = tmp_path / "libhostmodel.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{CSRC}",
                    f"-I{tmp_path}", "-o", str(so), str(tmp_path / "driver.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    lib.rhs_n.argtypes = [_dp, _dp, ctypes.c_int, _dp]
    lib.log_n.argtypes = [_dp, _dp, ctypes.c_long]
    return lib


def _tk_text():
    if not os.path.exists(TK_KERNEL):
        pytest.skip("the reference's TK2014 kernel is not present")
    from opendihu_amd.cellml import convert_vc_kernel
    return convert_vc_kernel(open(TK_KERNEL).read(), "tk2014")


def _oracle_or_skip(model):
    if not os.path.exists(os.path.join(ORACLE_BUILD, f"libfmso_{model}.so")):
        pytest.skip(f"oracle model {model} not built")
    return oracle_library(model)


@pytest.mark.parametrize("model", ["tk2014", SHORTEN, "hodgkin_huxley_gen"])
def test_generated_model_on_host_matches_oracle(tmp_path, model):
    if model == "tk2014":
        text, p_alg, p_const = _tk_text(), (), ()
    else:
        path = SHORTEN_C if model == SHORTEN else HH_C
        if not os.path.exists(path):
            pytest.skip("the reference's model file is not present")
        text = open(path).read()
        p_alg, p_const = ((32,), (65,)) if model == SHORTEN else ((), (2, 3, 4, 5))
    o = _oracle_or_skip(model)
    lib = _host_model(tmp_path, text, p_alg, p_const)
    ns = lib.ns()
    y0 = np.zeros(ns)
    o.fmso_initial_states(P(y0, _dp))
    npar = len(p_alg) + len(p_const)
    p0 = np.zeros(max(npar, 1))
    o.fmso_default_params(P(p0, _dp))
    n = 2000
    rng = np.random.default_rng(7)
    Y = np.ascontiguousarray(y0[:, None] * (1 + 0.02 * rng.standard_normal((ns, n))))
    Y[0] = rng.uniform(-90, 40, n)
    Pm = np.ascontiguousarray(np.repeat(p0[:max(npar, 1), None], n, axis=1))
    R = np.zeros((ns, n))
    lib.rhs_n(P(Y, _dp), P(Pm, _dp), n, P(R, _dp))
    Ro = np.zeros((ns, n))
    rr, aa = np.zeros(ns), np.zeros(256)
    for j in range(n):
        yy = np.ascontiguousarray(Y[:, j])
        pp = np.ascontiguousarray(Pm[:, j])
        o.fmso_rhs(P(yy, _dp), P(pp, _dp), 0.0, P(rr, _dp), P(aa, _dp), 0)
        Ro[:, j] = rr
    assert np.isfinite(Ro).all()
    np.testing.assert_array_equal(R, Ro)


def test_tk2014_conversion_structure():
    from opendihu_amd.cellml import parse_opencor_c
    m = parse_opencor_c(_tk_text())
    assert (m.n_states, m.n_algebraics) == (57, 71)
    assert sorted(k for kind, k, _ in m.lines if kind == "RATES") == list(range(57))
    assert sorted(k for kind, k, _ in m.lines if kind == "ALGEBRAIC") == list(range(71))
    # every algebraic is assigned before it is read
    seen = set()
    for kind, k, expr in m.lines:
        for a in re.findall(r"ALGEBRAIC\[(\d+)\]", expr):
            assert int(a) in seen, f"ALGEBRAIC[{a}] read before it is assigned"
        if kind == "ALGEBRAIC":
            seen.add(k)
    # initial values of the kernel's initializeStates (tk2014.cpp.disabled:800-858)
    assert m.state_initial[0] == -79.974 and m.state_initial[30] == 1500.0 and m.state_initial[56] == 0.23
    # pow(constant57, n) is the C library's pow (std::pow in the kernel), not a product of rounded squares
    cv = m.constant_values()
    assert cv[57] == 0.2 and math.pow(0.2, 4.0) in cv and (0.2 * 0.2) * (0.2 * 0.2) != math.pow(0.2, 4.0)


def test_tk2014_second_evaluation_is_the_first_on_intermediate_states():
    if not os.path.exists(TK_KERNEL):
        pytest.skip("the reference's TK2014 kernel is not present")
    text = open(TK_KERNEL).read()
    k0 = text.index("compute0DInstance")
    first = dict((f"{kind}{k}", e) for kind, k, e in
                 re.findall(r"^\s*const double_v (algebraic|rate)(\d+)\s*=\s*(.*?);\s*$", text[k0:], re.M))
    second = dict((f"{kind.lower()}{k}", e) for kind, k, e in
                  re.findall(r"^\s*const double_v intermediate(Algebraic|Rate)(\d+)\s*=\s*(.*?);\s*$", text[k0:], re.M))
    assert len(first) == 128 and set(first) == set(second)

    def norm(e):
        e = re.sub(r"intermediateState(\d+)", r"states[\1]", e)
        e = re.sub(r"intermediate(Algebraic|Rate)(\d+)", lambda m: m.group(1).lower() + m.group(2), e)
        return re.sub(r"\s+", "", e)

    for k in first:
        assert norm(first[k]) == norm(second[k]), k


def test_accurate_log(tmp_path):
    o = oracle_library("hh")
    o.fmso_log_accurate_n.argtypes = [_dp, _dp, ctypes.c_long]
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0.5, 2.0, 100000), 10 ** rng.uniform(-300, 300, 100000),
                        rng.uniform(0.02, 0.1, 50000), 1 + rng.standard_normal(50000) * 1e-8,
                        np.array([1.0, 2.0, 0.5, 1e-310, 5e-324, 1.7976931348623157e308])])
    y = np.empty_like(x)
    o.fmso_log_accurate_n(P(x, _dp), P(y, _dp), x.size)
    ref = np.log(x)
    nz = ref != 0
    ulp = np.abs(y - ref)[nz] / np.spacing(np.abs(ref[nz]))
    assert ulp.max() <= 2.0, ulp.max()
    assert y[x == 1.0][0] == 0.0 and y[x == 2.0][0] == np.log(2.0) and y[x == 5e-324][0] == np.log(5e-324)
    # the product's version (fms_math.h) compiled for the host: the same bits
    if not os.path.exists(TK_KERNEL):
        return
    lib = _host_model(tmp_path, _tk_text())
    yp = np.empty_like(x)
    lib.log_n(P(x, _dp), P(yp, _dp), x.size)
    np.testing.assert_array_equal(yp, y)


def test_tk2014_four_lane_reaction_matches_one_node_at_a_time():
    """lane selects, broadcast constants and the per-lane accurate log of the
    four-lane TK2014 oracle: bit-identical to the one-node-at-a-time build"""
    from cases import c2
    from oracle_binding import OracleBatch
    from test_oracle import avx512
    models = ("tk2014", "tk2014_scalar") + (("tk2014_w8",) if avx512() else ())
    for m in models:
        _oracle_or_skip(m)
    mesh, stim = c2(n_fibers=2, n_nodes=103, length_cm=1.02, rate_hz=4000.0)
    out = []
    for model in models:
        o = OracleBatch(**mesh, scheme="IE", model=model, threads=2)
        o.set_firing_table_stimulation(**stim)
        o.enable_spike_detection(-40.0)
        o.advance(0.0, 2.5e-5, 6000)
        out.append((o.states.copy(), o.spikes()[0].copy()))
    for other in out[:1] + out[2:]:
        np.testing.assert_array_equal(other[0], out[1][0])
        np.testing.assert_array_equal(other[1], out[1][1])
    assert (out[1][1] >= 0).sum() >= 3, "the clamped nodes must fire"


@pytest.mark.parametrize("model", ["hh", SHORTEN, "tk2014"])
def test_model_code_under_address_and_undefined_behaviour_sanitizers(model, tmp_path):
    """the device model code (right-hand side, Heun step with FP64 and FP32
    gates, the clamp and the algebraics, row assembly, serial solve) built for
    the host with -fsanitize=address,undefined and run over random states:
    no out-of-bounds access and no undefined behaviour (tests/sanitize_heun.cpp)"""
    import glob
    import shutil
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    if model == "hh":
        header, mtype = "model_hh.h", "fms::ModelHH"
        inc = []
    else:
        gen = os.path.join(REPO, "opendihu_amd", "_gen")
        found = sorted(glob.glob(os.path.join(gen, f"model_{model}_*.h")))
        if not found:
            pytest.skip(f"generated {model} header not present (built where the reference checkout exists)")
        header = os.path.basename(found[0])
        mtype = "fms::Model_" + header[len("model_"):-2]
        inc = [f"-I{gen}"]
    exe = tmp_path / "sanitize_heun"
    cmd = [hipcc, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer",
           f"-I{CSRC}", f"-I{os.path.join(REPO, 'include')}", *inc,
           f'-DFMS_MODEL_HEADER="{header}"', f"-DFMS_MODEL_TYPE={mtype}",
           "-o", str(exe), "-x", "hip", os.path.join(REPO, "tests", "sanitize_heun.cpp")]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    run = subprocess.run([str(exe), "1500" if model != "hh" else "20000"], capture_output=True, text=True, env=env,
                         timeout=600)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "checksum" in run.stdout
    print(model, run.stdout.strip())
