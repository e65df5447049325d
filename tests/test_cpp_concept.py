"""The host C++ concept classes (include/byo_graph.hpp) compile against the
C ABI with g++ (CPU) and, on the GPU, reproduce the oracle's answers when
driven like the reference harness drives a container."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "concept_test.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "concept_test")


def build_exe():
    lib = os.path.join(ROOT, "byo_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", EXE,
           "-L", lib, "-lbyo", "-Wl,-rpath," + lib]
    subprocess.check_call(cmd)
    return EXE


def test_concept_classes_compile_and_link():
    assert os.path.exists(build_exe())


@pytest.mark.gpu
def test_concept_classes_match_oracle():
    exe = EXE if os.path.exists(EXE) else build_exe()
    path = os.path.join(ROOT, "tests", "golden", "inputs", "rMatGraph_J_5_100")
    out = json.loads(subprocess.check_output([exe, path], timeout=120).decode().strip().splitlines()[-1])
    g = O.read_adjacency_graph(path)
    assert out["n"] == g.n and out["m"] == g.m and out["deleted"] == 1
    assert out["degsum"] == g.m and out["dstsum"] == O.map_sum(g) == out["map_sum"]
    assert out["first_stop"] == 1
    assert out["reached"] == int((O.bfs_levels(g, 0) != O.NONE).sum())
    assert out["components"] == len(np.unique(O.cc(g)))
    assert out["triangles"] == O.tc(g)
    keys = set(O.graph_keys(g).tolist())
    added = {(0 << 32) | 5, (5 << 32) | 0, (1 << 32) | 2, (2 << 32) | 1} - keys
    assert out["m_after_insert"] == g.m + len(added)
    assert out["m_after_remove"] == g.m - (4 - len(added))
