"""AdjacencyGraph files through the C ABI, host only (no device): the
reference's own input files read as the oracle reads them, text and -b
binary round trips, the binary layout of gbbs/graph_io.cc:86-97, and
malformed files rejected with BYO_E_INVAL."""
import os

import numpy as np
import pytest

import oracle as O

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


@pytest.fixture(scope="module")
def B():
    import byo_amd
    return byo_amd


@pytest.mark.parametrize("name", ["rMatGraph_J_5_100", "star.txt", "triangles.txt"])
def test_reference_inputs(B, name):
    path = os.path.join(INPUTS, name)
    off, ids = B.read_adjacency(path)
    g = O.read_adjacency_graph(path)
    assert np.array_equal(off, g.offsets)
    assert np.array_equal(ids, g.ids)


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip(B, tmp_path, binary):
    g = O.rmat_graph(12)
    p = str(tmp_path / ("g.bin" if binary else "g.adj"))
    B.write_adjacency(p, g.offsets, g.ids, binary)
    off, ids = B.read_adjacency(p, binary)
    assert np.array_equal(off, g.offsets)
    assert np.array_equal(ids, g.ids)


def test_text_matches_writer_layout(B, tmp_path):
    """One number per line after the header, as write_graph_to_file writes
    (gbbs/graph_io.h:656-669); the oracle's reader agrees."""
    off = np.array([0, 2, 3, 3, 4], np.uint64)
    ids = np.array([1, 3, 0, 0], np.uint32)
    p = str(tmp_path / "t.adj")
    B.write_adjacency(p, off, ids)
    assert open(p).read() == "AdjacencyGraph\n4\n4\n0\n2\n3\n3\n1\n3\n0\n0\n"
    g = O.read_adjacency_graph(p)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.ids, ids)


def test_binary_layout(B, tmp_path):
    g = O.rmat_graph(10)
    p = str(tmp_path / "g.bin")
    B.write_adjacency(p, g.offsets, g.ids, binary=True)
    raw = open(p, "rb").read()
    n, m = g.n, g.m
    hdr = np.frombuffer(raw[:24], np.int64)
    assert list(hdr) == [n, m, 4 * m]
    assert np.array_equal(np.frombuffer(raw[24:24 + 8 * (n + 1)], np.uint64), g.offsets)
    assert np.array_equal(np.frombuffer(raw[24 + 8 * (n + 1):], np.uint32), g.ids)
    assert len(raw) == 24 + 8 * (n + 1) + 4 * m


def test_text_to_binary_to_text(B, tmp_path):
    src = os.path.join(INPUTS, "rMatGraph_J_5_100")
    off, ids = B.read_adjacency(src)
    b = str(tmp_path / "x.bin")
    t = str(tmp_path / "x.adj")
    B.write_adjacency(b, off, ids, binary=True)
    B.write_adjacency(t, *B.read_adjacency(b, True))
    off2, ids2 = B.read_adjacency(t)
    assert np.array_equal(off, off2) and np.array_equal(ids, ids2)


def test_empty_graph(B, tmp_path):
    for binary in (False, True):
        p = str(tmp_path / ("e%d" % binary))
        B.write_adjacency(p, np.zeros(4, np.uint64), np.zeros(0, np.uint32), binary)
        off, ids = B.read_adjacency(p, binary)
        assert list(off) == [0, 0, 0, 0] and len(ids) == 0


@pytest.mark.parametrize("text", [
    "WeightedAdjacencyGraph\n2\n2\n0\n1\n1\n0\n",   # weighted header
    "AdjacencyGraph\n2\n2\n0\n1\n1\n",              # an id missing
    "AdjacencyGraph\n2\n2\n0\n1\n1\n0\n7\n",        # a token too many
    "AdjacencyGraph\n2\n2\n0\n1\n1\nx\n",           # not a number
    "AdjacencyGraph\n2\n2\n0\n1\n1\n2\n",           # id >= n
    "AdjacencyGraph\n2\n2\n0\n3\n1\n0\n",           # offsets beyond m
    "AdjacencyGraph\n3\n2\n0\n2\n1\n1\n0\n",        # offsets not monotone
    "AdjacencyGraph\n2\n2\n-1\n1\n1\n0\n",          # negative
])
def test_malformed_text(B, tmp_path, text):
    p = str(tmp_path / "bad.adj")
    with open(p, "w") as f:
        f.write(text)
    with pytest.raises(B.ByoError) as e:
        B.read_adjacency(p)
    assert e.value.code == B.E_INVAL


def test_malformed_binary(B, tmp_path):
    g = O.rmat_graph(8)
    p = str(tmp_path / "g.bin")
    B.write_adjacency(p, g.offsets, g.ids, binary=True)
    raw = open(p, "rb").read()
    for cut in (10, 24 + 8, len(raw) - 3):
        q = str(tmp_path / ("cut%d.bin" % cut))
        with open(q, "wb") as f:
            f.write(raw[:cut])
        with pytest.raises(B.ByoError):
            B.read_adjacency(q, True)
    bad = bytearray(raw)
    bad[-4:] = np.array([g.n], np.uint32).tobytes()  # last id out of range
    q = str(tmp_path / "badid.bin")
    with open(q, "wb") as f:
        f.write(bytes(bad))
    with pytest.raises(B.ByoError):
        B.read_adjacency(q, True)


def test_missing_file_and_bad_writes(B, tmp_path):
    with pytest.raises(B.ByoError):
        B.read_adjacency(str(tmp_path / "nope.adj"))
    with pytest.raises(B.ByoError):  # offsets[n] != m
        B.write_adjacency(str(tmp_path / "w.adj"), np.array([0, 1], np.uint64), np.zeros(0, np.uint32))
    with pytest.raises(B.ByoError):
        B.write_adjacency(str(tmp_path / "no_dir" / "w.adj"), np.zeros(2, np.uint64), np.zeros(0, np.uint32))
