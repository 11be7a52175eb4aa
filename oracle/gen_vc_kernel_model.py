"""Oracle cell model from a hand-written Vc Heun kernel (test infrastructure only).

The reference ships the 57-state TK2014 model only as the hand-written kernel
core/src/specialized_solver/fast_monodomain_solver/tk2014.cpp.disabled. This
script copies that file's lines, without parsing their expressions, into a C
translation unit in oracle/_gen (not versioned) that includes fms_oracle_core.h.
The lines it takes:

  * the constant lines of compute0DInstance (``const double constantK = ...``);
  * the first right-hand-side evaluation (``const double_v algebraicK|rateK = ...``,
    up to the Heun predictor);
  * initializeStates.

The oracle's own Heun step does the predictor, the clamp and the second
evaluation exactly as the kernel writes them (tk2014.cpp.disabled:298-566).
Names are bound by macros:

  * ``double_v`` -> double, or the four-lane fmso_vw for the lane path;
  * ``exponential`` -> exponential_approx (the kernel's own, lines 23-34);
  * ``sqr``, ``pow3``, ``pow4`` -> x*x, x*x*x, sqr(sqr(x)) (lines 5-18);
  * ``log`` (Vc::log) -> fmso_log_accurate;
  * ``Vc::iif(c, a, b)`` -> c ? a : b (lane select);
  * ``double_v(Vc::One|Zero)`` -> 1 | 0;
  * ``pow`` stays the C library's pow, the kernel's std::pow on constants.

This is a route to the model independent of the product's tokenizing converter
(opendihu_amd/cellml/vc_kernel.py).

usage: python gen_vc_kernel_model.py KERNEL.cpp NAME [MARCH [scalar|w8]]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
# build variants: the W = 4 lane path (AVX2, default), one node at a time, and
# the W = 8 lane path of an AVX-512 host (the reference's Vc double_v there)
VARIANT_FLAGS = {"": [], "scalar": ["-DFMSO_SCALAR_ONLY"], "w8": ["-DFMSO_W=8", "-march=x86-64-v4"]}


def main(path, name, march="x86-64-v3", variant=""):
    text = open(path).read()
    k0 = text.index("compute0DInstance")
    k1 = text.index("intermediateState0", k0)
    kernel = text[k0:k1]
    consts = re.findall(r"^\s*(const double constant\d+\s*=.*;)\s*$", kernel, re.M)
    lines = re.findall(r"^\s*(const double_v (algebraic|rate)(\d+)\s*=\s*(.*);)\s*$", kernel, re.M)
    i0 = text.index("initializeStates")
    states = {int(k): v for k, v in re.findall(r"^\s*states\[(\d+)\]\s*=\s*(.*?);", text[i0:], re.M)}
    n_states = max(states) + 1
    alg = sorted(int(k) for _, kind, k, _ in lines if kind == "algebraic")
    rates = sorted(int(k) for _, kind, k, _ in lines if kind == "rate")
    n_alg = max(alg) + 1
    assert rates == list(range(n_states)), "every state needs a rate"
    for _, _, _, e in lines:
        for b in re.findall(r"\bpow\(([^,]*),", e):
            assert re.fullmatch(r"\s*constant\d+\s*", b), f"pow of a non-constant: {b}"

    def lane_form(stmt, v4):
        s = stmt.replace("Vc::iif(", "Vc_iif(").replace("double_v(Vc::One)", "1.0").replace("double_v(Vc::Zero)", "0.0")
        if v4:   # the right-hand side may be a scalar expression: broadcast it
            s = re.sub(r"^(const double_v \w+\s*=\s*)(.*);$", r"\1FMSO_BC(\2);", s)
        return s

    def body(v4):
        out = [lane_form(s, v4) for s, _, _, _ in lines]
        out += [f"RATES[{k}] = rate{k};" for k in rates]
        out += [f"ALGEBRAIC[{k}] = algebraic{k};" for k in alg]
        return "\n  ".join(out)

    y0 = ", ".join(states.get(i, "0") for i in range(n_states))
    src = f"""/* oracle model generated from {os.path.basename(path)} (test infrastructure) */
#define NS {n_states}
#define NA {n_alg}
#define NP 0
#define MODEL_NAME "{name}"
static const double model_y0[NS] = {{{y0}}};
static void model_rhs_a(const double *states, const double *P_, double VOI, double *RATES, double *ALGEBRAIC);
static void model_rhs_e(const double *states, const double *P_, double VOI, double *RATES, double *ALGEBRAIC);
#ifndef FMSO_SCALAR_ONLY
#define FMSO_MODEL_RHS_VW model_rhs_a_vw
#endif
#include "fms_oracle_core.h"

/* the kernel's helpers (tk2014.cpp.disabled:5-18) */
static inline double sqr_s(double x) {{ return x * x; }}
static inline double pow3_s(double x) {{ return x * x * x; }}
static inline double pow4_s(double x) {{ return sqr_s(sqr_s(x)); }}

#define Vc_iif(c, a, b) ((c) ? (a) : (b))
#define double_v double
#define sqr(x) sqr_s(x)
#define pow3(x) pow3_s(x)
#define pow4(x) pow4_s(x)
#define log(x) fmso_log_accurate(x)
#define exponential(x) exponential_approx(x)
static void model_rhs_a(const double *states, const double *P_, double VOI, double *RATES, double *ALGEBRAIC)
{{
  (void)VOI; (void)P_;
  {chr(10).join("  " + c for c in consts).strip()}
  {body(False)}
}}
#undef exponential
#define exponential(x) exp(x)
static void model_rhs_e(const double *states, const double *P_, double VOI, double *RATES, double *ALGEBRAIC)
{{
  (void)VOI; (void)P_;
  {chr(10).join("  " + c for c in consts).strip()}
  {body(False)}
}}
#undef exponential
#undef log
#undef sqr
#undef pow3
#undef pow4
#undef double_v
#undef Vc_iif

#ifdef FMSO_MODEL_RHS_VW
static inline fmso_vw sqr_v(fmso_vw x) {{ return x * x; }}
static inline fmso_vw pow3_v(fmso_vw x) {{ return x * x * x; }}
static inline fmso_vw pow4_v(fmso_vw x) {{ return sqr_v(sqr_v(x)); }}
#define Vc_iif(c, a, b) FMSO_SEL((c), (a), (b))
#define double_v fmso_vw
#define sqr(x) _Generic((x), fmso_vw: sqr_v, default: sqr_s)(x)
#define pow3(x) _Generic((x), fmso_vw: pow3_v, default: pow3_s)(x)
#define pow4(x) _Generic((x), fmso_vw: pow4_v, default: pow4_s)(x)
#define log(x) _Generic((x), fmso_vw: fmso_log_accurate_vw, default: fmso_log_accurate)(x)
#define exponential(x) _Generic((x), fmso_vw: exponential_approx_vw, default: exponential_approx)(x)
static void model_rhs_a_vw(const fmso_vw *states, const fmso_vw *P_, double VOI, fmso_vw *RATES, fmso_vw *ALGEBRAIC)
{{
  (void)VOI; (void)P_;
  {chr(10).join("  " + c for c in consts).strip()}
  {body(True)}
}}
#undef exponential
#undef log
#undef sqr
#undef pow3
#undef pow4
#undef double_v
#undef Vc_iif
#endif

void fmso_default_params(double *p) {{ (void)p; }}
"""
    os.makedirs(os.path.join(HERE, "_gen"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "build"), exist_ok=True)
    sfx = "_scalar" if variant == "scalar" else ("_w8" if variant == "w8" else "")
    cfile = os.path.join(HERE, "_gen", f"model_{name}{sfx}.c")
    with open(cfile, "w") as f:
        f.write(src)
    out = os.path.join(HERE, "build", f"libfmso_{name}{sfx}.so")
    subprocess.run(["gcc", "-O3", f"-march={march}", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                    "-shared", "-std=gnu11", "-Wno-unused-function", "-Wno-unused-variable", "-I", HERE, "-o", out,
                    cfile, "-lm"] + VARIANT_FLAGS[variant], check=True)
    return out


if __name__ == "__main__":
    march = sys.argv[3] if len(sys.argv) > 3 else "x86-64-v3"
    scalar = sys.argv[4] if len(sys.argv) > 4 else ""
    print(main(sys.argv[1], sys.argv[2], march, scalar))
