"""Parser for the C files OpenCOR exports from CellML models.

Accepted subset (the one the reference's parser accepts,
core/src/cellml/source_code_generator/00_parse_source_code_file.cpp:159-357):

* a header comment ``There are a total of N entries in the algebraic variable
  array`` / ``... in each of the rate and state variable arrays`` / ``... in the
  constant variable array`` and ``STATES[i] is <name> in component <c> (<unit>)``
  style lines;
* ``initConsts``: ``STATES[i] = <number>;`` and ``CONSTANTS[i] = <expr>;``
  (constant expressions may use earlier constants);
* ``computeRates``: one assignment ``ALGEBRAIC[i] = <expr>;`` or
  ``RATES[i] = <expr>;`` per line, in evaluation order;
* ``computeVariables``: further ``ALGEBRAIC[i]`` lines; those already assigned
  in computeRates are skipped, the others appended (as the reference does).

Expressions use + - * /, parentheses, the ternary ``c ? a : b`` with
comparisons and ``&&``/``||``, numeric literals, ``VOI`` and the functions
exp, log, pow, fabs, sqrt.
"""
from __future__ import annotations

import ast
import math
import operator
import re
from dataclasses import dataclass, field

_TOTAL_RE = {
    "algebraics": re.compile(r"There are a total of (\d+) entries in the algebraic variable array"),
    "states": re.compile(r"There are a total of (\d+) entries in each of the rate and state variable arrays"),
    "constants": re.compile(r"There are a total of (\d+) entries in the constant variable array"),
}
_NAME_RE = re.compile(r"\*\s*(STATES|CONSTANTS|ALGEBRAIC)\[(\d+)\] is (\S+) in component (\S+) \(([^)]*)\)")
_ASSIGN_RE = re.compile(r"^\s*(STATES|CONSTANTS|ALGEBRAIC|RATES)\[(\d+)\]\s*=\s*(.*?);\s*$")


@dataclass
class OpenCorModel:
    n_states: int
    n_algebraics: int
    n_constants: int
    state_initial: list[float]
    constant_lines: list[tuple[int, str]]          # (index, expression) in initConsts order
    lines: list[tuple[str, int, str]]              # (ALGEBRAIC|RATES, index, expression)
    names: dict[str, dict[int, str]] = field(default_factory=dict)

    def state_name(self, i: int) -> str:
        return self.names.get("STATES", {}).get(i, f"state{i}")

    def constant_values(self) -> list[float]:
        """Initial constant values, evaluated in IEEE double like the C code."""
        vals: dict[int, float] = {}
        for idx, expr in self.constant_lines:
            vals[idx] = _eval_const(expr, vals)
        return [vals.get(i, 0.0) for i in range(self.n_constants)]


def parse_opencor_c(text: str) -> OpenCorModel:
    totals = {}
    for k, rx in _TOTAL_RE.items():
        m = rx.search(text)
        if not m:
            raise ValueError(f"model file lacks the 'There are a total of ... {k}' header")
        totals[k] = int(m.group(1))
    names: dict[str, dict[int, str]] = {}
    for m in _NAME_RE.finditer(text):
        names.setdefault(m.group(1), {})[int(m.group(2))] = f"{m.group(4)}/{m.group(3)}"
    if "OC_CellML_RHS_routine" in text:
        raise ValueError("OpenCMISS-generated model files are not supported (use an OpenCOR export)")

    section = None
    state_initial = [0.0] * totals["states"]
    constant_lines: list[tuple[int, str]] = []
    lines: list[tuple[str, int, str]] = []
    assigned_alg: set[int] = set()
    for raw in text.splitlines():
        line = raw.strip()
        if line.startswith("initConsts") or re.match(r"^void\s+initConsts", line):
            section = "init"
            continue
        if line.startswith("computeRates") or re.match(r"^void\s+computeRates", line):
            section = "rates"
            continue
        if line.startswith("computeVariables") or re.match(r"^void\s+computeVariables", line):
            section = "variables"
            continue
        m = _ASSIGN_RE.match(line)
        if not m or section is None:
            continue
        kind, idx, expr = m.group(1), int(m.group(2)), m.group(3)
        if section == "init":
            if kind == "STATES":
                state_initial[idx] = float(expr)
            elif kind == "CONSTANTS":
                constant_lines.append((idx, expr))
        elif section in ("rates", "variables"):
            if kind == "ALGEBRAIC":
                if idx in assigned_alg:
                    continue
                assigned_alg.add(idx)
                lines.append((kind, idx, expr))
            elif kind == "RATES" and section == "rates":
                lines.append((kind, idx, expr))
    return OpenCorModel(totals["states"], totals["algebraics"], totals["constants"], state_initial,
                        constant_lines, lines, names)


# ---------------------------------------------------------------------------
# constant folding in IEEE double (Python floats), for parameter defaults
# ---------------------------------------------------------------------------
_BIN = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv}
# constants are evaluated with the host C library, as the reference's generated
# "const double constantK = ..." lines are (03_generator_vc.cpp:937-944)
_FN = {"pow": math.pow, "exp": math.exp, "log": math.log, "fabs": math.fabs, "sqrt": math.sqrt}


def _eval_const(expr: str, consts: dict[int, float]) -> float:
    py = re.sub(r"CONSTANTS\[(\d+)\]", lambda m: f"C{m.group(1)}", expr)
    tree = ast.parse(py, mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant):
            return float(n.value)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub):
            return -ev(n.operand)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.UAdd):
            return ev(n.operand)
        if isinstance(n, ast.BinOp) and type(n.op) in _BIN:
            return _BIN[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.Name) and n.id.startswith("C"):
            return consts[int(n.id[1:])]
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FN:
            return _FN[n.func.id](*[ev(a) for a in n.args])
        raise ValueError(f"unsupported constant expression: {expr}")

    return ev(tree)
