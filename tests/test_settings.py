"""Settings front end and input formats (CPU)."""
import os

import numpy as np

from opendihu_amd.settings import (element_lengths, node_positions, parse_fiber_distribution, parse_firing_times,
                                   vc_stimulation_range, write_stimulation_log)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_parse_reference_firing_times_file():
    table = parse_firing_times(open(os.path.join(GOLDEN, "MU_firing_times_always.txt")).read())
    assert table.shape == (2, 100) and table.all()


def test_parse_firing_times_separators():
    table = parse_firing_times("0 1\t0\n1  0 1\n0 0 1 \n")
    np.testing.assert_array_equal(table, [[0, 1, 0], [1, 0, 1], [0, 0, 1]])


def test_parse_reference_fibre_distribution_file():
    mus = parse_fiber_distribution(open(os.path.join(GOLDEN, "MU_fibre_distribution_10MUs.txt")).read())
    np.testing.assert_array_equal(mus[:10], [8, 4, 8, 9, 10, 10, 10, 6, 10, 8])
    assert mus.min() == 1 and mus.max() == 10


def test_mesh_element_lengths():
    pos = node_positions({"nElements": [100], "physicalExtent": [1.0], "physicalOffset": [0]})
    h = element_lengths(pos)
    assert len(h) == 100 and np.allclose(h, 0.01, rtol=1e-12)
    h3 = element_lengths(np.array([[0, 0, 0], [3.0, 4.0, 0], [3.0, 4.0, 12.0]]))
    np.testing.assert_array_equal(h3, [5.0, 12.0])


def test_vc_stimulation_range():
    assert vc_stimulation_range(0, 50, 101, 4) == (48, 51)
    assert vc_stimulation_range(0, 500, 1001, 8) == (496, 499)


def test_stimulation_log_format(tmp_path):
    p = tmp_path / "stim.log"
    write_stimulation_log(str(p), [1.5, 0.25, 0.5], [3, 1, 3], [2, 2, 2])
    assert p.read_text() == "# motor unit no; fiber no; stimulation times in ms\n2;1;0.25;\n2;3;0.5;1.5;"
