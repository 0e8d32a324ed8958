"""GPU parity of the multi-way intersection against the CPU oracles."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from test_oracle import _mintersect_kat_deps

pytestmark = pytest.mark.gpu


def test_reference_kat():
    (off, adj), base = _mintersect_kat_deps()
    dep = vs.Graph(off, adj, symmetric=False)
    counts, offsets, lists, st = vs.mintersect([dep, dep], base)
    assert counts.tolist() == [1, 1, 1, 0, 0, 0, 0, 0]
    assert lists.tolist() == [1, 1, 1] and st["total"] == 3
    assert offsets.tolist() == [0, 1, 2, 3, 3, 3, 3, 3, 3]


@pytest.mark.parametrize("scale,T", [(12, 5000), (16, 40000)])
def test_rmat_paths(orc, scale, T):
    g = vs.HostCSR.rmat(scale, 16, seed=scale)
    G = vs.Graph.from_csr(g)
    tup = vs.sample_paths3(g.offsets, g.adj, T, seed=3)
    counts, offsets, lists, st = vs.mintersect([G, G, G], tup)
    c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj)] * 3, tup)
    assert np.array_equal(counts, c2)
    assert np.array_equal(offsets, o2)
    assert np.array_equal(lists, l2)
    # invariants: ascending, duplicate-free, subset of each list
    for t in range(0, T, max(1, T // 50)):
        seg = lists[offsets[t]:offsets[t + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        for x in tup[t]:
            assert np.isin(seg, g.adj[g.offsets[x]:g.offsets[x + 1]]).all()


def test_arities_and_bad_ids(orc):
    g = vs.HostCSR.rmat(11, 16, seed=4)
    G = vs.Graph.from_csr(g)
    rng = np.random.default_rng(0)
    hubs = np.argsort(g.degrees())[-20:]
    for arity in (1, 2, 4, 8):
        tup = rng.choice(hubs, size=(500, arity)).astype(np.uint32)
        tup[::50, 0] = g.V + 5  # an id outside the dependency: empty set
        counts, offsets, lists, _ = vs.mintersect([G] * arity, tup)
        c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj)] * arity, tup)
        assert np.array_equal(counts, c2) and np.array_equal(lists, l2)
        assert np.all(counts[::50] == 0)


def test_counts_only_and_empty():
    g = vs.HostCSR.rmat(10, 8, seed=1)
    G = vs.Graph.from_csr(g)
    tup = vs.sample_paths3(g.offsets, g.adj, 100, seed=1)
    c, o, l, st = vs.mintersect([G, G, G], tup, with_lists=False)
    assert l is None and st["total"] == int(c.sum())
    c, o, l, st = vs.mintersect([G, G, G], np.zeros((0, 3), dtype=np.uint32))
    assert len(c) == 0 and o.tolist() == [0]


def test_lists_must_be_sets():
    off = np.array([0, 3, 5, 6], dtype=np.uint64)
    tup = np.array([[0, 1]], dtype=np.uint32)
    # a repeated id, then a misordered list, in list 0
    for adj in ([1, 1, 2, 0, 2, 0], [2, 1, 0, 0, 2, 0]):
        G = vs.Graph(off, np.array(adj, dtype=np.uint32), symmetric=False)
        for _ in range(2):  # the check is remembered per CSR
            with pytest.raises(vs.VSError) as e:
                vs.mintersect([G, G], tup)
            assert e.value.code == vs._lib.VS_ERR_INVALID
        G.close()
    G = vs.Graph(off, np.array([0, 1, 2, 0, 2, 0], dtype=np.uint32), symmetric=False)
    c, o, l, _ = vs.mintersect([G, G], tup)
    assert c.tolist() == [2] and l.tolist() == [0, 2]
    G.close()


@pytest.mark.parametrize("stage_mb", [None, "0", "1"])
def test_staged_and_two_pass_agree(orc, monkeypatch, stage_mb):
    """Lists come from one staged pass by default; VS_MI_STAGE_MAX_MB=0 forces
    count -> scan -> fill, and a 1 MB limit (below this workload's 4-byte
    per driver entry staging) falls back the same way.  All match the oracle."""
    if stage_mb is not None:
        monkeypatch.setenv("VS_MI_STAGE_MAX_MB", stage_mb)
    g = vs.HostCSR.rmat(16, 16, seed=21)
    G = vs.Graph.from_csr(g)
    tup = vs.sample_paths3(g.offsets, g.adj, 40000, seed=8)
    b, d = g.offsets[tup], g.offsets[tup + 1]
    assert (d - b).min(axis=1).sum() * 4 > (1 << 20)  # the 1 MB case cannot stage
    counts, offsets, lists, st = vs.mintersect([G, G, G], tup)
    c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj)] * 3, tup)
    assert np.array_equal(counts, c2) and np.array_equal(offsets, o2)
    assert np.array_equal(lists, l2) and st["total"] == len(l2)
    G.close()


def test_device_capacity_error_keeps_counts(orc):
    """On HBM buffers, a list capacity one short of the total fails with
    VS_ERR_CAPACITY after the staged pass, with the counts, offsets and total
    already written; the same call with enough room gives the lists."""
    import ctypes as C
    lib = vs._lib
    g = vs.HostCSR.rmat(12, 16, seed=12)
    G = vs.Graph.from_csr(g)
    tup = vs.sample_paths3(g.offsets, g.adj, 3000, seed=5)
    T = len(tup)
    c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj)] * 3, tup)
    need = int(o2[-1])
    assert need > 1
    d_t, d_c, d_o, d_l = (C.c_void_p() for _ in range(4))
    lib.call("vs_dev_alloc", tup.nbytes, C.byref(d_t))
    lib.call("vs_dev_alloc", T * 4, C.byref(d_c))
    lib.call("vs_dev_alloc", (T + 1) * 8, C.byref(d_o))
    lib.call("vs_dev_alloc", need * 4, C.byref(d_l))
    try:
        lib.call("vs_memcpy_h2d", d_t, tup.ctypes.data_as(C.c_void_p), tup.nbytes)
        deps = (C.c_void_p * 3)(G.handle, G.handle, G.handle)
        total = C.c_uint64()
        st = lib.MintersectStats()
        with pytest.raises(vs.VSError) as e:
            lib.call("vs_mintersect_device", deps, 3, d_t, T, d_c, d_o, d_l, need - 1,
                     C.byref(total), C.byref(st))
        assert e.value.code == lib.VS_ERR_CAPACITY and total.value == need
        counts = np.zeros(T, dtype=np.uint32)
        offs = np.zeros(T + 1, dtype=np.uint64)
        lib.call("vs_memcpy_d2h", counts.ctypes.data_as(C.c_void_p), d_c, T * 4)
        lib.call("vs_memcpy_d2h", offs.ctypes.data_as(C.c_void_p), d_o, (T + 1) * 8)
        assert np.array_equal(counts, c2) and np.array_equal(offs, o2)
        lib.call("vs_mintersect_device", deps, 3, d_t, T, d_c, d_o, d_l, need,
                 C.byref(total), C.byref(st))
        lists = np.zeros(need, dtype=np.uint32)
        lib.call("vs_memcpy_d2h", lists.ctypes.data_as(C.c_void_p), d_l, need * 4)
        assert total.value == need and np.array_equal(lists, l2)
    finally:
        for p in (d_t, d_c, d_o, d_l):
            lib.call("vs_dev_free", p)
        G.close()


def giants_graph(seed=5):
    """Three vertices whose lists are longer than the 8192 entries above which
    a tuple's shortest list gets a wave of its own, two whose lists exceed the
    2048 entries of the next class, among short lists."""
    rng = np.random.default_rng(seed)
    V = 40000
    lists = []
    for v in range(V):
        if v < 5:
            k = 9000 + 3000 * v if v < 3 else 3000 + 2000 * (v - 3)
            nb = rng.choice(np.arange(5, V), size=k, replace=False)
        else:
            nb = rng.choice(V, size=int(rng.integers(0, 12)), replace=False)
        lists.append(np.sort(nb).astype(np.uint32))
    off = np.zeros(V + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in lists])
    return off, np.concatenate(lists)


def giant_tuples(T, V, seed, hi=5):
    """min(T, 40) tuples over the first hi vertices (the long lists), the rest
    random."""
    rng = np.random.default_rng(seed)
    ng = min(T, 40)
    tup = np.concatenate([rng.integers(0, hi, size=(ng, 3)),
                          rng.integers(0, V, size=(T - ng, 3))]).astype(np.uint32)
    rng.shuffle(tup)
    return tup


# T = 20: as many giants as waves (they are strided); T = 30: one wave
# each with waves to spare; T = 5000: giants and the next class among short
# and long tuples
@pytest.mark.parametrize("T", [20, 30, 5000])
def test_giant_tuples(orc, T):
    off, adj = giants_graph()
    G = vs.Graph(off, adj, symmetric=False)
    tup = giant_tuples(T, len(off) - 1, seed=T, hi=3 if T <= 30 else 5)
    counts, offsets, lists, st = vs.mintersect([G, G, G], tup)
    c2, o2, l2 = orc.mintersect_sorted([(off, adj)] * 3, tup)
    assert np.array_equal(counts, c2)
    assert np.array_equal(offsets, o2)
    assert np.array_equal(lists, l2)
    assert int(c2.max()) > 1000  # the giants' intersections are large
