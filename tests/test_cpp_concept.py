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


ABL_SRC = os.path.join(ROOT, "tests", "cpp", "api_ablation.cpp")
ABL_EXE = os.path.join(ROOT, "tests", "cpp", "api_ablation")


def build_ablation():
    lib = os.path.join(ROOT, "byo_amd")
    cmd = ["g++", "-std=c++20", "-O2", "-I", os.path.join(ROOT, "include"), ABL_SRC, "-o", ABL_EXE,
           "-L", lib, "-lbyo", "-Wl,-rpath," + lib]
    subprocess.check_call(cmd)
    return ABL_EXE


def test_api_ablation_harness_compiles():
    """The Graph_API_use variants compile against the concept classes."""
    assert os.path.exists(build_ablation())


@pytest.mark.gpu
@pytest.mark.parametrize("name,src", [("rMatGraph_J_5_100", 10), ("star.txt", 0), ("triangles.txt", 1)])
def test_api_ablations_agree(name, src):
    """full_api, minimal_api, no_early_exit, no_parallel_map, no_degree and
    no_M give the same degrees, M and BFS levels, equal to the oracle's."""
    exe = ABL_EXE if os.path.exists(ABL_EXE) else build_ablation()
    path = os.path.join(ROOT, "tests", "golden", "inputs", name)
    rows = [json.loads(l) for l in subprocess.check_output([exe, path, str(src)], timeout=120).decode().splitlines()
            if l.startswith("{")]
    assert [r["api"] for r in rows] == ["full_api", "minimal_api", "no_early_exit", "no_parallel_map", "no_degree",
                                        "no_M"]
    g = O.read_adjacency_graph(path)
    lv = O.bfs_levels(g, src)
    reached = lv != O.NONE
    for r in rows:
        assert (r["n"], r["m"], r["degsum"]) == (g.n, g.m, g.m), r
        assert r["reached"] == int(reached.sum()) and r["level_sum"] == int(lv[reached].sum()), r
    assert len({r["pull_levels"] for r in rows}) == 1
