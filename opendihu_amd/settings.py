"""Input formats and settings helpers of the fast fibre solver.

* firing times file: rows = sampling steps, columns = motor units, entries 0/1
  separated by spaces or tabs (parsed as the reference parses it,
  core/src/specialized_solver/fast_monodomain_solver/
  fast_monodomain_solver_initialization.tpp:151-206: each row is read number
  after number with atoi, separators are a space or tab followed by any
  non-digit characters)
* fibre distribution file: whitespace-separated motor-unit numbers, one per
  fibre, used cyclically (initialization.tpp:208-223, 324)
* 1D fibre meshes: element lengths are the distances between consecutive node
  positions (fast_monodomain_solver_communication.tpp:41-47)
"""
from __future__ import annotations

import math

import numpy as np


def _atoi(s: str) -> int:
    """C atoi: optional whitespace and sign, then digits; 0 if none"""
    s = s.lstrip(" \t\n\r\f\v")
    i = 0
    if i < len(s) and s[i] in "+-":
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j > i else 0


def parse_firing_times(text: str) -> np.ndarray:
    """firingTimesFile contents -> uint8 table [rows, motor units]"""
    rows = []
    n_cols = 0
    while text:
        end = text.find("\n")
        if end < 0:
            break
        line, text = text[:end], text[end + 1:]
        row = []
        while line:
            row.append(1 if _atoi(line) else 0)
            pos = min([p for p in (line.find("\t"), line.find(" ")) if p >= 0], default=-1)
            if pos < 0:
                break
            line = line[pos + 1:]
            while line and not line[0].isdigit():
                line = line[1:]
        if n_cols == 0:
            n_cols = len(row)
        rows.append(row)
    if not rows:
        raise ValueError("could not parse firing times")
    table = np.zeros((len(rows), n_cols), np.uint8)
    for r, row in enumerate(rows):
        k = min(n_cols, len(row))
        table[r, :k] = row[:k]
    return table


def parse_fiber_distribution(text: str) -> np.ndarray:
    """fiberDistributionFile contents -> motor-unit number of each listed fibre"""
    out = []
    while text:
        out.append(_atoi(text))
        pos = min([p for p in (text.find("\t"), text.find(" ")) if p >= 0], default=-1)
        if pos < 0:
            break
        text = text[pos + 1:]
    if not out:
        raise ValueError("could not parse motor units")
    return np.array(out, np.int32)


def node_positions(mesh: dict) -> np.ndarray:
    """node positions (n, 3) of a 1D structured mesh from its settings"""
    if "nodePositions" in mesh:
        pos = np.asarray(mesh["nodePositions"], np.float64)
        if pos.ndim == 1:
            pos = np.stack([pos, np.zeros_like(pos), np.zeros_like(pos)], 1)
        return pos
    n_el = mesh["nElements"]
    n_el = int(n_el[0] if isinstance(n_el, (list, tuple)) else n_el)
    ext = mesh.get("physicalExtent", 1.0)
    ext = float(ext[0] if isinstance(ext, (list, tuple)) else ext)
    off = mesh.get("physicalOffset", 0.0)
    off = float(off[0] if isinstance(off, (list, tuple)) else off)
    x = off + np.arange(n_el + 1) * (ext / n_el)
    return np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1)


def element_lengths(pos: np.ndarray) -> np.ndarray:
    """distance between consecutive nodes, summed component by component as the
    reference's MathUtility::distance does"""
    out = np.empty(len(pos) - 1)
    for e in range(len(pos) - 1):
        s = 0.0
        for c in range(pos.shape[1]):
            d = pos[e + 1, c] - pos[e, c]
            s += d * d
        out[e] = math.sqrt(s)
    return out


def vc_stimulation_range(offset: int, nmj: int, n: int, width: int = 4) -> tuple[int, int]:
    """nodes the reference's Vc path clamps: lanes 0..min(3,W)-1 of the W-aligned
    point buffer holding the neuromuscular junction (compute.tpp:161-166,
    03_generator_vc.cpp:1032-1038), clipped to the fibre"""
    base = width * ((offset + nmj) // width) - offset
    return max(0, min(n, base)), max(0, min(n, base + min(3, width)))


def gpu_stimulation_range(nmj: int, n: int) -> tuple[int, int]:
    """nodes the reference's gpu path clamps: NMJ-1..NMJ+1 (gpu.tpp:540)"""
    return max(0, nmj - 1), min(n, nmj + 2)


def write_stimulation_log(path: str, times, fibers, motor_units) -> None:
    """the stimulation log file format of StimulationLogging::writeLogFile
    (core/src/control/diagnostic_tool/stimulation_logging.cpp:126-166): entries
    sorted by motor unit, fibre and time; one line per (motor unit, fibre)"""
    entries = sorted(zip(motor_units, fibers, times))
    with open(path, "w") as f:
        f.write("# motor unit no; fiber no; stimulation times in ms\n")
        cur = None
        first = True
        for mu, fib, t in entries:
            if cur != (fib, mu):
                cur = (fib, mu)
                if not first:
                    f.write("\n")
                f.write(f"{mu};{fib};")
                first = False
            f.write(f"{t:g};")
