"""Fibre output in the reference's PythonFile format and its comparison metric.

PythonFile (core/src/output_writer/python_file/, python_structured_deformable.tpp:
17-93) writes one dict per output step: the 1D fibre mesh as a
"StructuredDeformable" mesh of one rank, and its field variables ("geometry"
with components x, y, z, and "solution" with the membrane voltage) as lists of
nodal values. With "binary": False it is JSON text, otherwise a pickle of the
same dict; testing/unit_testing/src/py_reader.py reads either. File names follow
the output writer's numbering (output_writer/generic.tpp:74-95):
"<filename>_<NNNNNNN>.py", NNNNNNN the output count ("incremental") or the
write-call count ("timeStepIndex"); here each fibre gets its own file base
"<filename>_<fibre>" when there is more than one fibre.

compare_directories is the metric of the reference's fast/slow comparison
(examples/electrophysiology/monodomain/hodgkin_huxley_fast/build_release/out/cmp.py,
testing/unit_testing/src/1_rank/cellml.cpp:532-559): for the files of two
directories in sorted order, ||v1 - v2||_2 / N of the "solution" field's first
component, averaged over the files.
"""
from __future__ import annotations

import json
import os
import pickle

import numpy as np

VERSION = "opendihu_amd fibre path"


def fibre_dict(positions: np.ndarray, vm: np.ndarray, step: int, time: float, field_name: str = "solution",
               component_name: str = "membrane/V") -> dict:
    """the PythonFile dict of one fibre: positions (n, 3), vm (n,)"""
    n = len(vm)
    pos = np.asarray(positions, np.float64).reshape(n, 3)
    return {
        "version": VERSION,
        "meta": "",
        "meshType": "StructuredDeformable",
        "dimension": 1,
        "nElementsGlobal": [n - 1],
        "nElementsLocal": [n - 1],
        "beginNodeGlobalNatural": [0],
        "hasFullNumberOfNodes": [True],
        "basisFunction": "Lagrange",
        "basisOrder": 1,
        "onlyNodalValues": True,
        "nRanks": 1,
        "ownRankNo": 0,
        "data": [
            {"name": "geometry", "components": [{"name": c, "values": pos[:, j].tolist()}
                                                for j, c in enumerate("xyz")]},
            {"name": field_name, "components": [{"name": component_name, "values": np.asarray(vm, np.float64).tolist()}]},
        ],
        "timeStepNo": int(step),
        "currentTime": float(time),
    }


def write_file(path: str, d: dict, binary: bool) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    if binary:
        with open(path, "wb") as f:
            pickle.dump(d, f, protocol=pickle.HIGHEST_PROTOCOL)
    else:
        with open(path, "w") as f:
            json.dump(d, f)


def read_file(path: str) -> dict:
    """JSON text or a pickle this module wrote (only files of this writer: a
    pickle runs code when loaded)"""
    try:
        with open(path) as f:
            return json.load(f)
    except (UnicodeDecodeError, json.JSONDecodeError):
        with open(path, "rb") as f:
            return pickle.load(f)


def solution_values(d: dict, field_name: str = "solution") -> np.ndarray:
    """py_reader.get_values(d, field_name, "0"): the field's first component"""
    for fv in d["data"]:
        if fv["name"] == field_name:
            return np.asarray(fv["components"][0]["values"], np.float64)
    raise KeyError(field_name)


def compare_directories(dir1: str, dir2: str) -> tuple[float, list[float]]:
    """mean over file pairs of ||v1 - v2||_2 / N (cmp.py); (mean, per file)"""
    f1 = sorted(os.path.join(dir1, f) for f in os.listdir(dir1) if f.endswith(".py"))
    f2 = sorted(os.path.join(dir2, f) for f in os.listdir(dir2) if f.endswith(".py"))
    n = min(len(f1), len(f2))
    if n == 0:
        raise ValueError("no .py output files to compare")
    errors = []
    for a, b in zip(f1[:n], f2[:n]):
        v1, v2 = solution_values(read_file(a)), solution_values(read_file(b))
        errors.append(float(np.linalg.norm(v1 - v2) / np.size(v1)))
    return float(np.mean(errors)), errors


class PythonFileWriter:
    """one "PythonFile" OutputWriter entry of the StrangSplitting settings"""

    def __init__(self, filename: str, interval: int, binary: bool = True, numbering: str = "incremental"):
        if interval < 1:
            raise ValueError("outputInterval must be >= 1 split step")
        if numbering not in ("incremental", "timeStepIndex"):
            raise ValueError(f"fileNumbering {numbering!r}: 'incremental' or 'timeStepIndex'")
        self.filename = filename
        self.interval = int(interval)
        self.binary = bool(binary)
        self.numbering = numbering
        self.n_written = 0
        self.files: list[str] = []

    @classmethod
    def from_settings(cls, w: dict, resolve=lambda p: p) -> "PythonFileWriter":
        return cls(resolve(w.get("filename", "out/fibers")), int(round(float(w.get("outputInterval", 1)))),
                   bool(w.get("binary", True)), w.get("fileNumbering", "incremental"))

    def write(self, solver, step: int, time: float) -> None:
        """all fibres of a FastMonodomainSolver after `step` split steps"""
        count = self.n_written if self.numbering == "incremental" else step // self.interval
        vms = solver.vm()
        positions = solver.node_positions
        for f, (pos, vm) in enumerate(zip(positions, vms)):
            base = self.filename if len(vms) == 1 else f"{self.filename}_{f:04d}"
            path = f"{base}_{count:07d}.py"
            write_file(path, fibre_dict(pos, vm, step, time), self.binary)
            self.files.append(path)
        self.n_written += 1
