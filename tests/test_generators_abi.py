"""CPU tests of the host half of the library: the synthetic generators pinned by
their numpy restatements and by R-MAT statistics, and the C ABI surface."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import vertexsurge_amd as vs
from vertexsurge_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "vertexsurge.h").read_text()
    declared = set(re.findall(r"^(?:vs_status|const char \*)\s*(vs_\w+)\(", header, re.M))
    assert len(declared) >= 40
    lib = C.CDLL(str(_lib.LIB_PATH))
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)


@pytest.mark.parametrize("scale,ef,seed", [(10, 16, 1), (12, 8, 7), (14, 4, 99)])
def test_rmat_matches_numpy_restatement(orc, scale, ef, seed):
    g = vs.HostCSR.rmat(scale, ef, seed=seed)
    u, v = orc.rmat_pairs(scale, ef, 0.57, 0.19, 0.19, seed)
    off, adj = orc.csr_from_pairs(1 << scale, u, v)
    assert np.array_equal(g.offsets, off)
    assert np.array_equal(g.adj, adj)


def test_rmat_invariants_and_quadrants(orc):
    g = vs.HostCSR.rmat(14, 16, seed=5)
    off, adj = g.offsets, g.adj
    V = g.V
    src = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64))
    assert np.all(src != adj)                                   # no self-loops
    key = src.astype(np.uint64) * V + adj
    assert np.all(np.diff(key.astype(np.int64)) > 0)            # sorted, no duplicates
    rev = np.sort(adj.astype(np.uint64) * V + src)
    assert np.array_equal(np.sort(key), rev)                    # symmetric
    # top-bit quadrant frequencies of the generated pairs follow a, b, c, d
    u, v = orc.rmat_pairs(14, 16, 0.57, 0.19, 0.19, 5)
    hu, hv = u >> 13, v >> 13
    n = len(u)
    for (bu, bv), p in {(0, 0): 0.57, (0, 1): 0.19, (1, 0): 0.19, (1, 1): 0.05}.items():
        f = np.count_nonzero((hu == bu) & (hv == bv)) / n
        assert abs(f - p) < 0.005, (bu, bv, f, p)


def test_er_graph():
    g = vs.HostCSR.erdos_renyi(5000, 40000, seed=3)
    assert g.V == 5000 and 0.95 * 80000 < g.E <= 80000
    d = g.degrees()
    assert abs(d.mean() - g.E / g.V) < 1e-9 and d.max() < 60


def test_sample_sources_matches_restatement(orc):
    for V, S, seed in [(1000, 10, 1), (1 << 16, 512, 12345), (100, 90, 2), (5, 5, 0)]:
        a = vs.sample_sources(V, S, seed)
        b = orc.sample_sources(V, S, seed)
        assert np.array_equal(a, b)
        assert len(np.unique(a)) == S and a.max() < V and np.all(np.diff(a.astype(np.int64)) > 0)


def test_sample_sources_rejects_too_many():
    with pytest.raises(vs.VSError) as e:
        vs.sample_sources(10, 11, 1)
    assert e.value.code == _lib.VS_ERR_INVALID


def test_sample_paths3_are_paths():
    g = vs.HostCSR.rmat(12, 8, seed=2)
    t = vs.sample_paths3(g.offsets, g.adj, 2000, seed=4)
    deg = g.degrees()
    for a, b, c in t[:500]:
        assert deg[a] > 0
        assert b in g.adj[g.offsets[a]:g.offsets[a + 1]]
        assert c in g.adj[g.offsets[b]:g.offsets[b + 1]]
    t2 = vs.sample_paths3(g.offsets, g.adj, 2000, seed=4)
    assert np.array_equal(t, t2)


def test_sample_paths3_rejects_malformed_csr():
    """The sampler reads the caller's CSR without an edge count: decreasing
    offsets or a neighbour id >= V are reported instead of read past."""
    off = np.array([0, 2, 1, 3], dtype=np.uint64)  # vertex 1's list runs backwards
    adj = np.array([1, 2, 0], dtype=np.uint32)
    with pytest.raises(vs.VSError) as e:
        vs.sample_paths3(off, adj, 10, seed=1)
    assert e.value.code == _lib.VS_ERR_INVALID and "decrease" in str(e.value)
    off = np.array([0, 1, 2, 2], dtype=np.uint64)
    adj = np.array([1, 7], dtype=np.uint32)  # 7 >= V = 3
    with pytest.raises(vs.VSError) as e:
        vs.sample_paths3(off, adj, 50, seed=1)
    assert e.value.code == _lib.VS_ERR_INVALID and ">= V" in str(e.value)


def test_csr_from_edges_checks_range():
    with pytest.raises(vs.VSError):
        vs.HostCSR.from_edges(4, [0, 1], [1, 9])
    g = vs.HostCSR.from_edges(4, [0, 1, 1, 3], [1, 2, 2, 3], symmetrize=False,
                              drop_self_loops=False)
    assert g.offsets.tolist() == [0, 1, 2, 2, 3] and g.adj.tolist() == [1, 2, 3]


def test_graph_create_without_gpu_is_an_error():
    if vs.device_count() > 0:
        pytest.skip("a GPU is visible")
    g = vs.HostCSR.rmat(8, 4, seed=1)
    with pytest.raises(vs.VSError) as e:
        vs.Graph.from_csr(g)
    assert e.value.code == _lib.VS_ERR_NODEVICE


def test_graph_create_rejects_malformed_csr():
    """Checked on the host before anything reaches the device (so also here,
    without a GPU): decreasing offsets and neighbour ids >= V are INVALID."""
    cases = [
        (np.array([0, 2, 1, 3], dtype=np.uint64), np.array([1, 2, 0], dtype=np.uint32),
         "decrease"),
        (np.array([0, 1, 2, 2], dtype=np.uint64), np.array([1, 3], dtype=np.uint32), ">= V"),
    ]
    for off, adj, msg in cases:
        with pytest.raises(vs.VSError) as e:
            vs.Graph(off, adj, symmetric=False)
        assert e.value.code == _lib.VS_ERR_INVALID and msg in str(e.value)
    # a large valid CSR passes the check (then needs a device)
    g = vs.HostCSR.rmat(14, 16, seed=3)
    try:
        G = vs.Graph.from_csr(g)
        G.close()
    except vs.VSError as e:
        assert e.code == _lib.VS_ERR_NODEVICE
