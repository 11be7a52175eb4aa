"""Hand-written Vc Heun kernels as OpenCOR-form model text.

The reference ships the 57-state TK2014 muscle model only as a hand-written Vc
kernel (core/src/specialized_solver/fast_monodomain_solver/tk2014.cpp.disabled),
not as a CellML export: ``compute0DInstance`` declares the constants
(``const double constantK = ...``, lines 56-165), evaluates the right-hand side
once (``const double_v algebraicK / rateK = ...``, lines 167-297), takes the
Heun predictor, clamps, evaluates it again on the intermediate states and
averages (lines 298-566), and ``initializeStates`` sets the initial values
(lines 800-858). That Heun step is the one the fibre kernels already run, so
only the first evaluation and the initial values are needed. This module
reads them and writes the OpenCOR C subset that ``parse_opencor_c`` accepts,
so the same code generator compiles the model.

Translation (token by token, keeping every operator and parenthesis):

* ``constantK`` -> ``CONSTANTS[K]``, ``states[k]`` -> ``STATES[k]``,
  ``algebraicK`` -> ``ALGEBRAIC[K]``, ``rateK`` -> ``RATES[K]``;
* ``exponential(x)`` -> ``exp(x)`` (the kernel's own ``exponential`` is the
  generator's (1 + x/1024)^1024, tk2014.cpp.disabled:23-34);
* ``sqr(x)``, ``pow3(x)``, ``pow4(x)`` -> ``pow(x, 2|3|4)``: the kernel's
  x*x, x*x*x and sqr(sqr(x)) (lines 5-18) give the same bits as the
  generator's pow2, x*pow2(x) and pow2(pow2(x));
* ``pow(constantK, n)`` is ``std::pow`` on doubles in the kernel; it is folded
  into a new constant evaluated with the C library's pow (for constant57 = 0.2,
  pow(0.2, 4) differs from (0.2^2)^2 in the last bit);
* ``log(x)`` is ``Vc::log``, an accurate logarithm outside the
  approximateExponentialFunction policy: ``log_accurate(x)``;
* ``Vc::iif(c, a, b)`` -> ``(c ? a : b)``, ``double_v(Vc::One)`` -> 1,
  ``double_v(Vc::Zero)`` -> 0.
"""
from __future__ import annotations

import re

_TOK = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|"
                  r"([A-Za-z_][\w]*(?:::[A-Za-z_]\w*)*)|(&&|\|\||<=|>=|==|!=|[-+*/()?:<>,\[\]!]))")


def _tokens(expr: str) -> list[str]:
    out, pos = [], 0
    expr = expr.strip()
    while pos < len(expr):
        m = _TOK.match(expr, pos)
        if not m or m.end() == pos:
            raise ValueError(f"cannot tokenize kernel expression at {expr[pos:pos + 30]!r}")
        out.append(m.group(m.lastindex))
        pos = m.end()
    return out


class _Converter:
    def __init__(self, n_constants: int):
        self.folded: dict[tuple[int, str], int] = {}   # (constant, exponent) -> new constant index
        self.next_constant = n_constants

    def _args(self, toks, i):
        """toks[i] == '(': the comma-separated argument token lists and the index after ')'."""
        args, cur, depth, j = [], [], 0, i + 1
        while j < len(toks):
            t = toks[j]
            if t == "(":
                depth += 1
            elif t == ")":
                if depth == 0:
                    args.append(cur)
                    return args, j + 1
                depth -= 1
            elif t == "," and depth == 0:
                args.append(cur)
                cur = []
                j += 1
                continue
            cur.append(t)
            j += 1
        raise ValueError("unbalanced parentheses in kernel expression")

    def expr(self, toks) -> str:
        out, i = [], 0
        while i < len(toks):
            t = toks[i]
            m = re.fullmatch(r"(constant|algebraic|rate)(\d+)", t)
            if m:
                kind = {"constant": "CONSTANTS", "algebraic": "ALGEBRAIC", "rate": "RATES"}[m.group(1)]
                out.append(f"{kind}[{m.group(2)}]")
                i += 1
                continue
            if t == "states":
                if toks[i + 1] != "[" or toks[i + 3] != "]":
                    raise ValueError(f"malformed states access near {toks[i:i + 4]}")
                out.append(f"STATES[{toks[i + 2]}]")
                i += 4
                continue
            if t in ("exponential", "log", "sqr", "pow3", "pow4", "pow", "Vc::iif", "double_v"):
                if toks[i + 1] != "(":
                    raise ValueError(f"{t} without argument list")
                args, j = self._args(toks, i + 1)
                out.append(self._call(t, args))
                i = j
                continue
            if re.match(r"^[A-Za-z_]", t):
                raise ValueError(f"unsupported identifier {t!r} in kernel expression")
            out.append(t)
            i += 1
        return " ".join(out)

    def _call(self, name, args) -> str:
        if name == "double_v":
            if args == [["Vc::One"]]:
                return "1.00000"
            if args == [["Vc::Zero"]]:
                return "0.00000"
            return f"({self.expr(args[0])})"
        if name == "Vc::iif":
            c, a, b = (self.expr(x) for x in args)
            return f"(({c}) ? ({a}) : ({b}))"
        if name == "exponential":
            return f"exp({self.expr(args[0])})"
        if name == "log":
            return f"log_accurate({self.expr(args[0])})"
        if name in ("sqr", "pow3", "pow4"):
            n = {"sqr": 2, "pow3": 3, "pow4": 4}[name]
            return f"pow({self.expr(args[0])}, {n}.00000)"
        # pow(base, exponent)
        base, ex = args
        mb = len(base) == 1 and re.fullmatch(r"constant(\d+)", base[0])
        if mb and len(ex) == 1:
            key = (int(mb.group(1)), ex[0])
            if key not in self.folded:
                self.folded[key] = self.next_constant
                self.next_constant += 1
            return f"CONSTANTS[{self.folded[key]}]"
        return f"pow({self.expr(base)}, {self.expr(ex)})"


def convert_vc_kernel(text: str, name: str = "tk2014") -> str:
    """OpenCOR-form C text of a hand-written Vc Heun kernel (see module doc)."""
    k0 = text.index("compute0DInstance")
    k1 = text.index("intermediateState0", k0)
    kernel = text[k0:k1]
    consts = [(int(k), e) for k, e in re.findall(r"^\s*const double constant(\d+)\s*=\s*(.*?);", kernel, re.M)]
    lines = [(kind, int(k), e) for kind, k, e in
             re.findall(r"^\s*const double_v (algebraic|rate)(\d+)\s*=\s*(.*?);\s*$", kernel, re.M)]
    i0 = text.index("initializeStates")
    states = [(int(k), e) for k, e in re.findall(r"^\s*states\[(\d+)\]\s*=\s*(.*?);", text[i0:], re.M)]
    if not consts or not lines or not states:
        raise ValueError("not a hand-written Vc Heun kernel (constants, rates or initial states missing)")
    n_states = max(k for k, _ in states) + 1
    n_alg = max(k for kind, k, _ in lines if kind == "algebraic") + 1
    rates = sorted(k for kind, k, _ in lines if kind == "rate")
    if rates != list(range(n_states)):
        raise ValueError(f"rates assigned {len(rates)} of {n_states} states")
    conv = _Converter(max(k for k, _ in consts) + 1)
    const_lines = [f"CONSTANTS[{k}] = {conv.expr(_tokens(e))};" for k, e in consts]
    rate_lines = []
    for kind, k, e in lines:
        target = "ALGEBRAIC" if kind == "algebraic" else "RATES"
        rate_lines.append(f"{target}[{k}] = {conv.expr(_tokens(e))};")
    for (c, ex), k in sorted(conv.folded.items(), key=lambda kv: kv[1]):
        const_lines.append(f"CONSTANTS[{k}] = pow(CONSTANTS[{c}], {ex});")
    state_lines = [f"STATES[{k}] = {e};" for k, e in states]
    n_const = conv.next_constant
    return f"""/*
   Model "{name}" converted from a hand-written Vc kernel.
   There are a total of {n_alg} entries in the algebraic variable array.
   There are a total of {n_states} entries in each of the rate and state variable arrays.
   There are a total of {n_const} entries in the constant variable array.
*/
void initConsts(double* CONSTANTS, double* RATES, double *STATES)
{{
{chr(10).join(state_lines)}
{chr(10).join(const_lines)}
}}
void computeRates(double VOI, double* CONSTANTS, double* RATES, double* STATES, double* ALGEBRAIC)
{{
{chr(10).join(rate_lines)}
}}
"""
