"""The C++ drop-in toolkits (endgraph_amd/toolkits) on a binary edge file,
checked against the oracle array by array: each toolkit writes its result
(parents, labels, distances, ranks) with dump_vertex_array when EG_DUMP names
a file, and its printed figures are checked too."""
import os
import re
import subprocess

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "endgraph_amd", "toolkits", "_build")


def run(tool, *args, dump=None):
    exe = os.path.join(BIN, tool)
    if not os.path.exists(exe):
        pytest.fail(f"toolkit {tool} not built (make -C endgraph_amd/toolkits)")
    print(f"run {tool} {' '.join(map(str, args))}", flush=True)
    env = dict(os.environ)
    if dump:
        env["EG_DUMP"] = str(dump)
    p = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, p.stderr + p.stdout
    return p.stdout


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    rng = np.random.default_rng(21)
    V = 5000
    edges = H.random_edges(rng, V, 40000, weighted=False)
    wedges = H.random_edges(rng, V, 40000, weighted=True)
    d = tmp_path_factory.mktemp("g")
    f8, f12 = d / "g.bin", d / "gw.bin"
    edges.tofile(f8)
    wedges.tofile(f12)
    return V, edges, wedges, str(f8), str(f12)


@pytest.mark.parametrize("P,S,policy", [(1, 1, "reference"), (3, 2, "auto")])
def test_bfs_toolkit(oracle_lib, graph_file, tmp_path, P, S, policy):
    V, edges, _, f8, _ = graph_file
    og = oracle_lib.OracleGraph(edges, V, P=P, S=S, eta=1, m=1)
    root = int(np.argmax(og.out_degree()))
    out = run("bfs", f8, V, root, 1, P, S, policy, dump=tmp_path / "parent.bin")
    found = [int(x) for x in re.findall(r"found_vertices = (\d+)", out)]
    parent, depth, _ = og.bfs(root)
    assert found and all(f == int((parent < V).sum()) for f in found)
    assert np.array_equal(np.fromfile(tmp_path / "parent.bin", np.uint32), parent)


@pytest.mark.parametrize("P,S", [(1, 1), (4, 1)])
def test_cc_toolkit(oracle_lib, graph_file, tmp_path, P, S):
    V, edges, _, f8, _ = graph_file
    og = oracle_lib.OracleGraph(edges[:3000], V, P=P, S=S, eta=1, m=1, symmetrize=True)
    small = os.path.join(os.path.dirname(f8), f"small{P}.bin")
    edges[:3000].tofile(small)
    out = run("cc", small, V, 0, 1, P, S, dump=tmp_path / "label.bin")
    comps = [int(x) for x in re.findall(r"components = (\d+)", out)]
    label, _ = og.cc()
    assert comps and all(c == len(np.unique(label)) for c in comps)
    assert np.array_equal(np.fromfile(tmp_path / "label.bin", np.uint32), label)


def test_sssp_toolkit(oracle_lib, graph_file, tmp_path):
    V, _, wedges, _, f12 = graph_file
    og = oracle_lib.OracleGraph(wedges, V, P=2, S=1, eta=1, m=1)
    root = int(np.argmax(og.out_degree()))
    out = run("sssp", f12, V, root, 1, 2, 1, dump=tmp_path / "dist.bin")
    dist, _ = og.sssp(root)
    got = np.fromfile(tmp_path / "dist.bin", np.float32)
    assert np.array_equal(got.view(np.uint32), dist.view(np.uint32))
    reach = dist < 1e9
    vmax = int(np.flatnonzero(reach & (dist == dist[reach].max()))[0])
    m = re.findall(r"distance\[(\d+)\]=([0-9.]+)", out)
    assert m and all(int(v) == vmax and float(x) == float(f"{dist[vmax]:f}") for v, x in m)


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_pagerank_toolkit(oracle_lib, graph_file, tmp_path, prec):
    V, edges, _, f8, _ = graph_file
    og = oracle_lib.OracleGraph(edges, V, P=2, S=1, eta=1, m=10)
    pr, deltas = og.pagerank(10)
    out = run("pagerank", f8, V, 10, 1, 2, 1, prec, dump=tmp_path / "pr.bin")
    got = np.fromfile(tmp_path / "pr.bin", np.float64)
    assert np.abs(got - pr).sum() / np.abs(pr).sum() <= (1e-6 if prec == "fp32c" else 1e-12)
    sums = [float(x) for x in re.findall(r"pr_sum=([0-9.]+)", out)]
    assert sums and all(abs(s - pr.sum()) <= 1e-6 * pr.sum() + 1e-6 for s in sums)
    d = [float(x) for x in re.findall(r"delta\(\d+\)=([0-9.]+)", out)[:10]]
    assert np.allclose(d, deltas[:10], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("P,S,policy", [(1, 1, "reference"), (2, 1, "auto"), (3, 2, "pull"), (1, 1, "push")])
def test_widest_toolkit_on_the_generic_engine(oracle_lib, graph_file, tmp_path, P, S, policy):
    """A fifth program written only against the facade's generic device
    engine (toolkits/widest.hip: the four signal/slot functors, no library
    kernel of its own): widths bit-exact against the numpy relaxation, in
    every mode (reference and push: sparse steps only; pull: dense steps
    only; auto: both)."""
    V, _, wedges, _, f12 = graph_file
    deg = np.bincount(wedges["src"], minlength=V)
    root = int(np.argmax(deg))
    out = run("widest", f12, V, root, 1, P, S, policy, dump=tmp_path / "width.bin")
    exp = H.widest_paths(V, wedges["src"].astype(np.int64), wedges["dst"].astype(np.int64), wedges["w"], root)
    got = np.fromfile(tmp_path / "width.bin", np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    reached = [int(x) for x in re.findall(r"reached = (\d+)", out)]
    assert reached and all(r == int((exp > 0).sum()) for r in reached)
    if policy == "auto":
        steps = re.findall(r"active\(\d+\)", out)
        assert len(steps) > 2
