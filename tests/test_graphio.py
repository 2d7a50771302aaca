"""METIS .graph files (README.md:208-218, DEVELOPMENT_GUIDE.md:846-874)
through rmetis_read_graph / rmetis_write_graph.  Host code only: no GPU.

The reference names the format (header "nvtxs nedges [fmt] [ncon]", 1-based
neighbour lists, '%' comments, fmt digits for vertex sizes / vertex weights
/ edge weights) but ships no reader, so the cases below are written from
that description: the guide's own 4-vertex example, every fmt, comments and
isolated vertices, round trips of generated graphs, and malformed files."""
import os

import numpy as np
import pytest

import rmetis_amd as R
from rmetis_amd import synth
from graphs import cycle4


def _write(tmp_path, text, name="g.graph"):
    p = os.path.join(tmp_path, name)
    with open(p, "w") as fh:
        fh.write(text)
    return p


def test_guide_example(tmp_path):
    # DEVELOPMENT_GUIDE.md:859-865: 4 vertices, 4 edges, unweighted
    p = _write(tmp_path, "% comment line\n4 4\n2 4\n1 3\n2 4\n1 3\n")
    g = R.read_graph(p)
    x, a = cycle4()
    assert g.nvtxs == 4 and g.ncon == 1 and g.nedges == 4
    assert np.array_equal(g.xadj, x) and np.array_equal(g.adjncy, a)
    assert g.vwgt is None and g.adjwgt is None and g.vsize is None


def test_every_format_digit(tmp_path):
    # path 1-2-3 with vertex sizes, two vertex weights and edge weights
    text = ("3 2 111 2\n"
            "5 1 2 2 7\n"
            "6 3 4 1 7 3 9\n"
            "% between vertices\n"
            "   7 5 6 2 9\n")
    g = R.read_graph(_write(tmp_path, text))
    assert g.nvtxs == 3 and g.ncon == 2
    assert list(g.xadj) == [0, 1, 3, 4]
    assert list(g.adjncy) == [1, 0, 2, 1]
    assert list(g.adjwgt) == [7, 7, 9, 9]
    assert list(g.vwgt) == [1, 2, 3, 4, 5, 6]
    assert list(g.vsize) == [5, 6, 7]


def test_edge_weights_and_isolated_vertex(tmp_path):
    # fmt 1 (edge weights only) and an empty line for an isolated vertex
    text = "4 2 1\n2 3\n1 3 3 4\n2 4\n\n"
    g = R.read_graph(_write(tmp_path, text))
    assert list(g.xadj) == [0, 1, 3, 4, 4]
    assert list(g.adjncy) == [1, 0, 2, 1]
    assert list(g.adjwgt) == [3, 3, 4, 4]
    assert g.vwgt is None


@pytest.mark.parametrize("make", [
    lambda: synth.grid(17, 23),
    lambda: synth.rgg(3000, 5, vwgt_max=4),
    lambda: synth.rgg(2000, 9, vwgt_max=8, ncon=2),
])
def test_round_trip(tmp_path, make):
    s = make()
    g = R.Graph(s.n, s.ncon, s.xadj, s.adjncy, s.vwgt, s.adjwgt)
    p = os.path.join(tmp_path, "rt.graph")
    R.write_graph(p, g)
    h = R.read_graph(p)
    assert h.nvtxs == g.nvtxs and h.ncon == g.ncon
    for a, b in ((h.xadj, g.xadj), (h.adjncy, g.adjncy)):
        assert np.array_equal(a, b)
    for a, b in ((h.vwgt, g.vwgt), (h.adjwgt, g.adjwgt)):
        assert (a is None and b is None) or np.array_equal(a, b)


@pytest.mark.parametrize("text", [
    "",                                   # no header
    "3\n",                                # header without nedges
    "2 1\n2\n",                           # a vertex line missing
    "2 1\n2\n3\n",                        # neighbour out of range
    "2 1\n2\n0\n",                        # 0 is not a 1-based id
    "2 2\n2\n1\n",                        # fewer entries than 2 * nedges
    "2 1\n2 1\n1\n",                      # more entries than 2 * nedges
    "2 1 1\n2\n1 5\n",                    # an edge weight missing
    "2 1 2\n2\n1\n",                      # fmt digit not 0/1
    "2 1\n2\n1\n5\n",                     # data after the last vertex
    "2 1\n2x\n1\n",                       # malformed number
])
def test_malformed_files_are_input_errors(tmp_path, text):
    with pytest.raises(R.InputError):
        R.read_graph(_write(tmp_path, text))


def test_missing_file_is_an_input_error(tmp_path):
    with pytest.raises(R.InputError):
        R.read_graph(os.path.join(tmp_path, "absent.graph"))
