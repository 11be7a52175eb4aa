"""Writes reference_kats.json: the known-answer values the reference's own unit
tests hold for the fibre path (run in a container that has the reference
checkout; the JSON is committed as a fixture).

Sources (paths relative to the reference checkout):
  testing/unit_testing/src/1_rank/cellml.cpp:14-64     HodgkinHuxley* (ExplicitEuler, dt=1e-5,
        y0=(-20, .05, .6, .325), I_stim=400, output at timeStepNo 90001)
  testing/unit_testing/src/1_rank/cellml.cpp:169-215   ShortenOpenCOR (ExplicitEuler, dt=1e-5,
        parameters [I_Stim=1000, l_hs=1], output at timeStepNo 90001)
  testing/unit_testing/src/1_rank/diffusion.cpp:118-150  ImplicitEuler1D (lumped mass, 5 elements on
        [0,4], u0=[2,2,4,5,2,2], 5 steps to t=0.1)
  testing/unit_testing/src/1_rank/diffusion.cpp:275-314  CrankNicolson1D (same setup)
  testing/unit_testing/src/1_rank/cellml.cpp:559        fast-vs-nested tolerance 0.1
The reference's test harness does not check these numbers (its parseNumber
returns 0, testing/unit_testing/src/utility.cpp:33-47), so they are candidates
that the oracle tests confirm or reject individually.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
UT = os.path.join(REF, "testing/unit_testing/src/1_rank")


def values_in(segment):
    return [float(v) for v in re.findall(r'\\"values\\": \[([-0-9.e+, ]+)\]', segment) for v in [v] ]


def solution_values(segment):
    sol = segment[segment.index('\\"solution\\"'):]
    out = []
    for m in re.finditer(r'\\"name\\": \\"([^\\]+)\\", \\"values\\": \[([-0-9.e+, ]+)\]', sol):
        out.append((m.group(1), [float(x) for x in m.group(2).split(",")]))
    return out


def test_segment(text, name):
    i = text.index(f"TEST({name}")
    j = text.index("assertFileMatchesContent", i)
    k = text.index(";", j)
    return text[i:k]


cellml = open(os.path.join(UT, "cellml.cpp")).read()
diffusion = open(os.path.join(UT, "diffusion.cpp")).read()
hh = solution_values(test_segment(cellml, "CellMLTest, HodgkinHuxleyVc"))
sh = solution_values(test_segment(cellml, "CellMLTest, ShortenOpenCOR"))
ie = solution_values(test_segment(diffusion, "DiffusionTest, ImplicitEuler1D"))
cn = solution_values(test_segment(diffusion, "DiffusionTest, CrankNicolson1D"))
geo = re.search(r'\\"name\\": \\"x\\", \\"values\\": \[([-0-9.e+, ]+)\]', test_segment(diffusion, "DiffusionTest, ImplicitEuler1D"))
out = {
    "hodgkin_huxley_explicit_euler": {
        "source": "testing/unit_testing/src/1_rank/cellml.cpp:14-64",
        "dt": 1e-5, "n_steps": 90001, "initial_states": [-20, 0.05, 0.6, 0.325], "i_stim": 400.0,
        "states": [v[0] for _, v in hh]},
    "shorten_explicit_euler": {
        "source": "testing/unit_testing/src/1_rank/cellml.cpp:169-215",
        "dt": 1e-5, "n_steps": 90001, "parameters": [1000.0, 1.0],
        "states": [v[0] for _, v in sh], "names": [n for n, _ in sh]},
    "diffusion_implicit_euler_lumped": {
        "source": "testing/unit_testing/src/1_rank/diffusion.cpp:118-150",
        "x": [float(x) for x in geo.group(1).split(",")], "u0": [2, 2, 4, 5, 2, 2], "dt": 0.02, "n_steps": 5,
        "u": ie[0][1]},
    "diffusion_crank_nicolson_lumped": {
        "source": "testing/unit_testing/src/1_rank/diffusion.cpp:275-314",
        "x": [float(x) for x in geo.group(1).split(",")], "u0": [2, 2, 4, 5, 2, 2], "dt": 0.02, "n_steps": 5,
        "u": cn[0][1]},
    "fast_vs_nested_tolerance": {"source": "testing/unit_testing/src/1_rank/cellml.cpp:559", "value": 0.1},
}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
with open(path, "w") as f:
    json.dump(out, f, indent=1)
print("wrote", path, {k: len(v.get("states", v.get("u", []))) for k, v in out.items()})
