"""Oracle cell model from an OpenCOR C export (test infrastructure only).

Writes a C translation unit that compiles the model file's own computeRates
lines verbatim, with exp/log/pow bound by preprocessor macros to the
reference generator's helpers (core/src/cellml/source_code_generator/
03_generator_vc.cpp:385-611), and includes fms_oracle_core.h. This is an
independent route to the model from the product's tokenizing code generator
(opendihu_amd/cellml): nothing is parsed beyond selecting lines and renaming
the parameter entries (00_parse_source_code_file.cpp:284-333: the
parametersUsedAsAlgebraic entries come first, then parametersUsedAsConstant).

The generated source contains the model file's text, so it is written to
oracle/_gen (not versioned) and compiled to oracle/build/libfmso_<name>.so.

usage: python gen_opencor_model.py MODEL.c NAME "alg,..." "const,..." [MARCH [scalar|w8]]
("scalar": without the W-lane function, one node at a time; "w8": eight lanes,
AVX-512)
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
# build variants: the W = 4 lane path (AVX2, default), one node at a time, and
# the W = 8 lane path of an AVX-512 host (the reference's Vc double_v there)
VARIANT_FLAGS = {"": [], "scalar": ["-DFMSO_SCALAR_ONLY"], "w8": ["-DFMSO_W=8", "-march=x86-64-v4"]}


def lane_selects(line):
    """(C ? A : B) -> FMSO_SEL((C), (A), (B)) for the four-lane function (the
    reference's Vc generator turns the same conditionals into lane selects)."""
    while "?" in line:
        q = line.rindex("?")
        depth, lo = 0, q - 1
        while lo >= 0 and not (line[lo] == "(" and depth == 0):
            depth += {")": 1, "(": -1}.get(line[lo], 0)
            lo -= 1
        depth, colon, hi = 0, None, q + 1
        while hi < len(line) and not (line[hi] == ")" and depth == 0):
            depth += {"(": 1, ")": -1}.get(line[hi], 0)
            if line[hi] == ":" and depth == 0:
                colon = hi
            hi += 1
        if lo < 0 or hi >= len(line) or colon is None:
            raise ValueError(f"conditional outside parentheses: {line}")
        line = (line[:lo] + f"FMSO_SEL(({line[lo + 1:q]}), ({line[q + 1:colon]}), ({line[colon + 1:hi]}))"
                + line[hi + 1:])
    return line


def main(path, name, p_alg, p_const, march="x86-64-v3", variant=""):
    text = open(path).read()
    n_alg = int(re.search(r"There are a total of (\d+) entries in the algebraic", text).group(1))
    n_states = int(re.search(r"There are a total of (\d+) entries in each of the rate and state", text).group(1))
    n_const = int(re.search(r"There are a total of (\d+) entries in the constant", text).group(1))
    lines = text.splitlines()
    section, y0, init, rates = None, {}, [], []
    for ln in lines:
        s = ln.strip()
        if s.startswith("initConsts"):
            section = "init"; continue
        if s.startswith("computeRates"):
            section = "rates"; continue
        if s.startswith("computeVariables"):
            section = None; continue
        if section == "init":
            m = re.match(r"STATES\[(\d+)\]\s*=\s*(.*);", s)
            if m:
                y0[int(m.group(1))] = m.group(2)
            elif s.startswith("CONSTANTS["):
                init.append(s)
        elif section == "rates" and (s.startswith("ALGEBRAIC[") or s.startswith("RATES[")):
            rates.append(s)
    palg = {k: j for j, k in enumerate(p_alg)}
    pconst = {k: len(p_alg) + j for j, k in enumerate(p_const)}
    body = []
    for s in rates:
        m = re.match(r"ALGEBRAIC\[(\d+)\]", s)
        if m and int(m.group(1)) in palg:
            continue
        s = re.sub(r"ALGEBRAIC\[(\d+)\]", lambda m: f"P_[{palg[int(m.group(1))]}]" if int(m.group(1)) in palg else m.group(0), s)
        s = re.sub(r"CONSTANTS\[(\d+)\]", lambda m: f"P_[{pconst[int(m.group(1))]}]" if int(m.group(1)) in pconst else m.group(0), s)
        body.append("  " + s)
    np_ = len(p_alg) + len(p_const)
    src = f"""/* oracle model generated from {os.path.basename(path)} (test infrastructure) */
#define NS {n_states}
#define NA {n_alg}
#define NP {np_}
#define MODEL_NAME "{name}"
{f"#define MODEL_PARAM_ALGEBRAICS {', '.join(str(k) for k in p_alg)}" if p_alg else ""}
static const double model_y0[NS] = {{{", ".join(y0.get(i, "0") for i in range(n_states))}}};
static void model_rhs_a(const double *STATES, const double *P_, double VOI, double *RATES, double *ALGEBRAIC);
static void model_rhs_e(const double *STATES, const double *P_, double VOI, double *RATES, double *ALGEBRAIC);
#ifndef FMSO_SCALAR_ONLY
#define FMSO_MODEL_RHS_VW model_rhs_a_vw
#endif
#include "fms_oracle_core.h"

static double CONSTANTS[{n_const}];
static int consts_ready = 0;
static void init_consts(void)
{{
{chr(10).join("  " + s for s in init)}
  consts_ready = 1;
}}
__attribute__((constructor)) static void init_consts_ctor(void) {{ init_consts(); }}

/* pow(x, n) for integer n as the reference generator builds it (03_generator_vc.cpp:541-611) */
static double ipow_(double x, int n)
{{
  if (n == 1) return x;
  if (n == 2) return x * x;
  int e0 = n / 2, e1 = n - n / 2;
  if (e0 == 1) return x * (ipow_(x, e1));
  return ipow_(ipow_(x, e1), e0);
}}
static double oracle_pow(double x, double y)
{{
  if (y == (double)(int)y && y != 0) {{
    int n = (int)y;
    return n > 0 ? ipow_(x, n) : 1. / (ipow_(x, -n));
  }}
  return pow(x, y);
}}

#define pow(x, y) oracle_pow((x), (y))
#define exp(x) exponential_approx(x)
#define log(x) logarithmic_approx(x)
static void model_rhs_a(const double *STATES, const double *P_, double VOI, double *RATES, double *ALGEBRAIC)
{{
  (void)VOI; (void)P_;
{chr(10).join(body)}
}}
#undef exp
#undef log
static void model_rhs_e(const double *STATES, const double *P_, double VOI, double *RATES, double *ALGEBRAIC)
{{
  (void)VOI; (void)P_;
{chr(10).join(body)}
}}
#undef pow

#ifdef FMSO_MODEL_RHS_VW
/* the same lines over four nodes per operation: states, parameters, rates and
   algebraics are four-lane vectors, CONSTANTS stay scalars (broadcast, as Vc
   broadcasts them), and each function picks its scalar or lane form by type */
static fmso_vw ipow_vw_(fmso_vw x, int n)
{{
  if (n == 1) return x;
  if (n == 2) return x * x;
  int e0 = n / 2, e1 = n - n / 2;
  if (e0 == 1) return x * (ipow_vw_(x, e1));
  return ipow_vw_(ipow_vw_(x, e1), e0);
}}
static fmso_vw oracle_pow_vw(fmso_vw x, double y)
{{
  if (y == (double)(int)y && y != 0) {{
    int n = (int)y;
    return n > 0 ? ipow_vw_(x, n) : 1. / (ipow_vw_(x, -n));
  }}
  fmso_vw r;
  for (int l = 0; l < FMSO_W; l++) r[l] = pow(x[l], y);
  return r;
}}
#define pow(x, y) _Generic((x), fmso_vw: oracle_pow_vw, default: oracle_pow)((x), (y))
#define exp(x) _Generic((x), fmso_vw: exponential_approx_vw, default: exponential_approx)(x)
#define log(x) _Generic((x), fmso_vw: logarithmic_approx_vw, default: logarithmic_approx)(x)
static void model_rhs_a_vw(const fmso_vw *STATES, const fmso_vw *P_, double VOI, fmso_vw *RATES, fmso_vw *ALGEBRAIC)
{{
  (void)VOI; (void)P_;
{chr(10).join(lane_selects(b) for b in body)}
}}
#undef pow
#undef exp
#undef log
#endif

void fmso_default_params(double *p)
{{
{chr(10).join(f"  p[{j}] = 0.0;" for j in range(len(p_alg)))}
{chr(10).join(f"  p[{len(p_alg) + j}] = CONSTANTS[{k}];" for j, k in enumerate(p_const))}
}}
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
    pa = [int(x) for x in sys.argv[3].split(",") if x] if len(sys.argv) > 3 else []
    pc = [int(x) for x in sys.argv[4].split(",") if x] if len(sys.argv) > 4 else []
    march = sys.argv[5] if len(sys.argv) > 5 else "x86-64-v3"
    scalar = sys.argv[6] if len(sys.argv) > 6 else ""
    print(main(sys.argv[1], sys.argv[2], pa, pc, march, scalar))
