"""HIP code generation for a parsed OpenCOR model.

The generated header defines ``struct <Name>`` with the interface of
``opendihu_amd/csrc/model_hh.h`` (NS, NA, NP, y0, p0, templated rhs<T, E>), so
the same fibre kernels run it.

Translation keeps each expression's structure (operators, parentheses and
their order) as written in the model file and only replaces atoms, because the
reference compiles exactly that text with g++ (core/src/cellml/
source_code_generator/03_generator_vc.cpp:937-1016); with -ffp-contract=off on
both sides the FP64 results agree bit for bit. Rewrites, as the reference's
preprocessCode (03_generator_vc.cpp:13-302) applies them:

* ``exp(`` -> ``E::exp(``, ``log(`` -> ``E::log(`` (approximation or device libm);
  ``log_accurate(`` (models converted by vc_kernel.py) -> the portable accurate log
* ``pow(x, n)`` with an integer literal n -> ``pow<n>(x)`` / ``powReciprocal<|n|>(x)``
  (recursive products, 03_generator_vc.cpp:541-611); other pow -> ``::pow``
* ``CONSTANTS[k]`` -> the constant, or a parameter when k is in
  parametersUsedAsConstant; ``ALGEBRAIC[k]`` for k in parametersUsedAsAlgebraic
  -> a parameter, and the line assigning it is dropped
  (00_parse_source_code_file.cpp:284-333). Parameter order: the
  parametersUsedAsAlgebraic entries first, then parametersUsedAsConstant.
* numeric literals become ``T(literal)`` so the same source compiles for FP64
  and for the FP32 gate arithmetic of the mixed-precision form.
"""
from __future__ import annotations

import re

from .opencor import OpenCorModel, parse_opencor_c

_TOKEN_RE = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|"
                       r"([A-Za-z_]\w*)|(&&|\|\||<=|>=|==|!=|[-+*/()?:<>,\[\]!]))")

_FUNCS = {"exp": "E::exp", "log": "E::log", "fabs": "fabs", "sqrt": "sqrt",
          # accurate log of models converted from hand-written kernels (vc_kernel.py)
          "log_accurate": "log_accurate"}


def _tokenize(expr: str) -> list[str]:
    pos, out = 0, []
    expr = expr.strip()
    while pos < len(expr):
        m = _TOKEN_RE.match(expr, pos)
        if not m or m.end() == pos:
            raise ValueError(f"cannot tokenize model expression at: {expr[pos:pos + 30]!r}")
        out.append(m.group(1) or m.group(2) or m.group(3))
        pos = m.end()
    return out


class _Translator:
    def __init__(self, model: OpenCorModel, p_alg: list[int], p_const: list[int]):
        self.model = model
        self.p_alg = {k: j for j, k in enumerate(p_alg)}
        self.p_const = {k: len(p_alg) + j for j, k in enumerate(p_const)}
        self.helpers: set[str] = set()

    def translate(self, expr: str, in_constants: bool = False) -> str:
        toks = _tokenize(expr)
        out, _ = self._seq(toks, 0, in_constants, stop=())
        return "".join(out)

    def _seq(self, toks, i, in_constants, stop):
        out = []
        while i < len(toks):
            t = toks[i]
            if t in stop:
                return out, i
            if t in ("STATES", "RATES", "ALGEBRAIC", "CONSTANTS"):
                if toks[i + 1] != "[" or toks[i + 3] != "]":
                    raise ValueError(f"malformed array access near {toks[i:i + 4]}")
                k = int(toks[i + 2])
                out.append(self._array(t, k, in_constants))
                i += 4
                continue
            if t == "VOI":
                out.append("T(t)")
                i += 1
                continue
            if t == "pow" and not in_constants:
                code, i = self._pow(toks, i, in_constants)
                out.append(code)
                continue
            if t in _FUNCS:
                out.append(_FUNCS[t] if not in_constants else t)
                i += 1
                continue
            if t == "(":
                inner, j = self._seq(toks, i + 1, in_constants, stop=(")",))
                out.append("(" + "".join(inner) + ")")
                i = j + 1
                continue
            if re.match(r"^[\d.]", t):
                out.append(t if in_constants else f"T({t})")
                i += 1
                continue
            if re.match(r"^[A-Za-z_]", t):
                raise ValueError(f"unsupported identifier {t!r} in model expression")
            out.append(f" {t} " if t in ("?", ":", "&&", "||", "<", ">", "<=", ">=", "==", "!=") else t)
            i += 1
        return out, i

    def _array(self, kind, k, in_constants):
        if kind == "STATES":
            return f"y[{k}]"
        if kind == "RATES":
            return f"r[{k}]"
        if kind == "ALGEBRAIC":
            if k in self.p_alg:
                return f"p[{self.p_alg[k]}]"
            return f"a[{k}]"
        # CONSTANTS
        if in_constants:
            return f"c{k}"
        if k in self.p_const:
            return f"p[{self.p_const[k]}]"
        return f"T(c{k})"

    def _pow(self, toks, i, in_constants):
        # pow ( base , exponent )
        if toks[i + 1] != "(":
            raise ValueError("pow without argument list")
        base, j = self._seq(toks, i + 2, in_constants, stop=(",",))
        if toks[j] != ",":
            raise ValueError("pow with one argument")
        # exponent tokens up to the matching ')'
        depth, k = 0, j + 1
        while True:
            if toks[k] == "(":
                depth += 1
            elif toks[k] == ")":
                if depth == 0:
                    break
                depth -= 1
            k += 1
        exp_toks = toks[j + 1:k]
        exp_txt = "".join(exp_toks)
        base_code = "".join(base)
        if re.fullmatch(r"[-+]?\d+(\.0*)?", exp_txt) and int(float(exp_txt)) != 0:
            n = int(float(exp_txt))
            name = f"pow{n}" if n > 0 else f"powReciprocal{-n}"
            if n == 1:
                code = f"({base_code})"
            else:
                self.helpers.add(name)
                code = f"{name}({base_code})"
        else:
            exp_code, _ = self._seq(exp_toks, 0, in_constants, stop=())
            code = f"pow({base_code}, {''.join(exp_code)})"
        return code, k + 1


def count_ops(model: OpenCorModel, parameters_used_as_algebraic=()) -> dict:
    """Algorithmic operation count of one right-hand-side evaluation: each binary
    + - * / is one op, each exp/log call one op, pow(x,n) with integer n its
    n-1 multiplications (a reciprocal power adds one division)."""
    p_alg = set(parameters_used_as_algebraic)
    c = dict(add=0, mul=0, div=0, exp=0, log=0, cmp=0)
    for kind, k, expr in model.lines:
        if kind == "ALGEBRAIC" and k in p_alg:
            continue
        toks = _tokenize(expr)
        prev = None
        i = 0
        while i < len(toks):
            t = toks[i]
            if t in ("+", "-"):
                if prev is not None and (prev in (")", "]") or re.match(r"^[\w.]", prev)):
                    c["add"] += 1
            elif t == "*":
                c["mul"] += 1
            elif t == "/":
                c["div"] += 1
            elif t == "exp":
                c["exp"] += 1
            elif t in ("log", "log_accurate"):
                c["log"] += 1
            elif t in ("<", ">", "<=", ">=", "==", "!="):
                c["cmp"] += 1
            elif t == "pow":
                # pow ( base , n ): count by exponent; base tokens are visited as usual
                depth, j = 0, i + 1
                while j < len(toks):
                    if toks[j] == "(":
                        depth += 1
                    elif toks[j] == ")":
                        depth -= 1
                        if depth == 0:
                            break
                    elif toks[j] == "," and depth == 1:
                        comma = j
                    j += 1
                e = "".join(toks[comma + 1:j])
                if re.fullmatch(r"[-+]?\d+(\.0*)?", e):
                    n = int(float(e))
                    c["mul"] += max(abs(n) - 1, 0)
                    c["div"] += 1 if n < 0 else 0
                    toks = toks[:comma] + toks[j:]
                else:
                    c["exp"] += 1
            prev = t
            i += 1
    c["total"] = c["add"] + c["mul"] + c["div"] + c["exp"] + c["log"]
    return c


def _cxx_double(v: float) -> str:
    r = repr(float(v))
    return r if ("." in r or "e" in r or "inf" in r or "nan" in r) else r + ".0"


def generate_hip_model(text: str, name: str, parameters_used_as_algebraic=(), parameters_used_as_constant=(),
                       struct_name: str | None = None) -> str:
    model = parse_opencor_c(text) if isinstance(text, str) else text
    p_alg = list(parameters_used_as_algebraic)
    p_const = list(parameters_used_as_constant)
    tr = _Translator(model, p_alg, p_const)
    struct = struct_name or f"Model_{re.sub(r'[^A-Za-z0-9_]', '_', name)}"
    cvals = model.constant_values()
    n_params = len(p_alg) + len(p_const)
    p0 = [0.0] * len(p_alg) + [cvals[k] for k in p_const]

    # constants are folded on the host in IEEE double with the C library (the
    # values the reference's g++ build computes); the device sees literals
    const_code = []
    for k, expr in model.constant_lines:
        const_code.append(f"    const double c{k} = {_cxx_double(cvals[k])};  // {expr}")
    used_consts = set()
    body = []
    for kind, k, expr in model.lines:
        if kind == "ALGEBRAIC" and k in tr.p_alg:
            body.append(f"    // ALGEBRAIC[{k}] is parameter {tr.p_alg[k]}")
            continue
        code = tr.translate(expr)
        used_consts.update(int(m) for m in re.findall(r"T\(c(\d+)\)", code))
        target = f"a[{k}]" if kind == "ALGEBRAIC" else f"r[{k}]"
        body.append(f"    {target} = {code};")
    # keep the constants the expressions use
    needed = set(used_consts)
    const_code = [c for (k, _), c in zip(model.constant_lines, const_code) if k in needed]

    ops = count_ops(model, p_alg)
    y0 = ", ".join(_cxx_double(v) for v in model.state_initial)
    p0s = ", ".join(_cxx_double(v) for v in p0) if p0 else "0.0"
    alg_is_param = " || ".join(f"k == {k}" for k in p_alg) if p_alg else "(void)k, false"
    guard = re.sub(r"[^A-Za-z0-9_]", "_", struct).upper()
    src = f"""// Generated by opendihu_amd.cellml from the OpenCOR model "{name}".
// parametersUsedAsAlgebraic = {p_alg}, parametersUsedAsConstant = {p_const}
#pragma once
#include "fms_math.h"

namespace fms {{

struct {struct} {{
  static constexpr int NS = {model.n_states};
  static constexpr int NA = {model.n_algebraics};
  static constexpr int NP = {n_params};
  static constexpr int STIM_PARAM = 0;
  // algorithmic ops of one rhs evaluation: {ops}
  static constexpr double OPS_RHS = {ops["total"]};
  static constexpr double DIVS_RHS = {ops["div"]};
  static constexpr double EXPS_RHS = {ops["exp"] + ops["log"]};
  static constexpr const char *name() {{ return "{name}"; }}
  static FMS_HD double y0(int k) {{
    const double v[NS] = {{{y0}}};
    return v[k];
  }}
  static FMS_HD double p0(int k) {{
    const double v[NP > 0 ? NP : 1] = {{{p0s}}};
    return v[k];
  }}
  // algebraics replaced by parameters (parametersUsedAsAlgebraic): not
  // computed, and not stored as algebraicsForTransfer
  static FMS_HD constexpr bool alg_is_param(int k) {{ return {alg_is_param}; }}

  template <class T, class E>
  static FMS_HD void rhs(const T *y, const T *p, double t, T *r, T *a) {{
    (void)t; (void)p;
{chr(10).join(const_code)}
{chr(10).join(body)}
  }}
}};

}}  // namespace fms
"""
    del guard
    return src
