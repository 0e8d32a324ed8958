"""Pins of the table-operator oracle (oracle/tables.py) on the CPU."""
import numpy as np
import pytest

from oracle import tables as ot


def _flat_mat_mat_deps():
    """MintersectFlatMatMat (mintersect.cpp:411-512): base (a, b); deps are
    10x10 ColMain matrices (rows c, columns a resp. b through a column map)."""
    a = [1, 2, 3, 4, 5, 6, 7, 8]
    b = [1, 1, 1, 2, 2, 3, 3, 4]
    c = [1, 1, 1, 2, 2, 3, 3, 4]
    deps = []
    for keys in (a, a):  # dep (c, a) and dep (c, b): both built over keys 1..8
        kmap = {}
        for x in keys:
            kmap.setdefault(x, len(kmap))
        M = np.zeros((10, 10), dtype=bool)
        for x, y in zip(keys, c):
            M[y, kmap[x]] = True
        deps.append((M, None, np.array(list(kmap), dtype=np.uint32)))
    return deps, np.stack([a, b], axis=1).astype(np.uint32)


def test_reference_flat_mat_mat_fixture():
    deps, base = _flat_mat_mat_deps()
    counts, offsets, lists = ot.mintersect_colmain(deps, base)
    # the reference asserts 3 result tuples: (1,1,1), (2,1,1), (3,1,1)
    assert int(counts.sum()) == 3
    assert counts.tolist() == [1, 1, 1, 0, 0, 0, 0, 0]
    assert lists.tolist() == [1, 1, 1]


def test_missing_column_gives_empty_set():
    deps, base = _flat_mat_mat_deps()
    base = base.copy()
    base[0, 0] = 99  # id absent from the column map
    counts, _, _ = ot.mintersect_colmain(deps, base)
    assert counts[0] == 0


def test_colmain_word_layout_by_hand():
    M = np.zeros((70, 2), dtype=bool)
    M[0, 0] = M[65, 0] = M[69, 1] = True
    w = ot.colmain_words(M)
    assert w.shape == (2, 8)  # align8(ceil(70/64)) = 8
    expect = np.zeros((2, 8), dtype=np.uint64)
    expect[0, 0] = 1
    expect[0, 1] = 1 << 1
    expect[1, 1] = 1 << 5
    assert np.array_equal(w, expect)


def test_table_convert_properties():
    rng = np.random.default_rng(3)
    M = rng.random((37, 90)) < 0.2
    src = np.arange(37, dtype=np.uint32) * 3
    c, rm, cm = ot.table_convert(M, src, None, transpose=False)
    assert np.array_equal(c, M) and rm is src and cm is None
    c, rm, cm = ot.table_convert(M, src, None, transpose=True)
    assert all(c[j, i] == M[i, j] for i in range(37) for j in range(90))
    assert rm is None and cm is src


def test_column_count_pair_order_with_ties():
    M = np.zeros((3, 4), dtype=bool)
    M[:2, 0] = True        # count 2
    M[0, 1] = True         # count 1
    M[1:, 2] = True        # count 2
    cmap = np.array([10, 11, 12, 13], dtype=np.uint32)
    ids, cnt = ot.column_count(M, cmap)
    assert ids.tolist() == [10, 11, 12, 13] and cnt.tolist() == [2, 1, 2, 0]
    ids, cnt = ot.column_count(M, cmap, "asc")
    assert list(zip(cnt.tolist(), ids.tolist())) == [(0, 13), (1, 11), (2, 10), (2, 12)]
    ids, cnt = ot.column_count(M, cmap, "desc")
    assert list(zip(cnt.tolist(), ids.tolist())) == [(2, 12), (2, 10), (1, 11), (0, 13)]
    assert ot.row_or_sum(M, cmap).tolist() == [10, 11, 12]
    assert ot.row_or_sum(M, None).tolist() == [0, 1, 2]


def test_adjacency_colmain_matches_sorted_merge(orc):
    u, v = orc.rmat_pairs(8, 8, 0.57, 0.19, 0.19, seed=5)
    off, adj = orc.csr_from_pairs(256, u, v)
    M = ot.adjacency_colmain(off, adj, 256)
    rng = np.random.default_rng(1)
    tup = rng.integers(0, 256, size=(300, 3)).astype(np.uint32)
    c1, o1, l1 = ot.mintersect_colmain([(M, None, None)] * 3, tup)
    c2, o2, l2 = orc.mintersect_sorted([(off, adj)] * 3, tup)
    assert np.array_equal(c1, c2) and np.array_equal(o1, o2) and np.array_equal(l1, l2)


def test_planted_mistake_is_caught(orc):
    """A transposed operand (AND of rows instead of columns) must fail the
    sorted-merge check."""
    u, v = orc.rmat_pairs(7, 8, 0.57, 0.19, 0.19, seed=2)
    off, adj = orc.csr_from_pairs(128, u, v)
    M = ot.adjacency_colmain(off, adj, 128)
    tup = np.random.default_rng(4).integers(0, 128, size=(200, 2)).astype(np.uint32)
    # symmetric adjacency makes M == M.T; break symmetry to expose the swap
    Mt = M.copy()
    Mt[0, :] = True
    c_ok, _, _ = ot.mintersect_colmain([(M, None, None)] * 2, tup)
    c_bad, _, _ = ot.mintersect_colmain([(Mt.T.copy(), None, None)] * 2, tup)
    c_ref, _, _ = orc.mintersect_sorted([(off, adj)] * 2, tup)
    assert np.array_equal(c_ok, c_ref)
    assert not np.array_equal(c_bad, c_ref)


def test_varrow_triangle_by_hand():
    """Path 0-1-2 plus edge 0-2 (a triangle) and a pendant 2-3.  Base: rows
    are sources {0, 2} (row map), bit (r, c) = c adjacent to the source.
    Intersecting N(start) and N(target) lists the third corner of each
    triangle through the base edge."""
    off = np.array([0, 2, 4, 7, 8])
    adj = np.array([1, 2, 0, 2, 0, 1, 3, 2], dtype=np.uint32)
    M = np.zeros((2, 4), dtype=bool)
    M[0, [1, 2]] = True  # N(0)
    M[1, [0, 1, 3]] = True  # N(2)
    s, t, w = ot.mintersect_varrow(M, np.array([0, 2]), None, [0, 1],
                                   adjacency=[(off, adj)] * 2)
    got = list(zip(s.tolist(), t.tolist(), w.tolist()))
    assert got == [(0, 1, 2), (0, 2, 1), (2, 0, 1), (2, 1, 0)]
    # the same with ColMain adjacency dependencies
    A = ot.adjacency_colmain(off, adj, 4)
    s2, t2, w2 = ot.mintersect_varrow(M, np.array([0, 2]), [(A, None, None)] * 2, [0, 1])
    assert np.array_equal(s, s2) and np.array_equal(t, t2) and np.array_equal(w, w2)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_varrow_brute_force(orc, seed):
    """Every (start, target, w) with bit (start, target) set and w in the sets
    the var_idx choice names, by a triple loop over all vertices."""
    import vertexsurge_amd as vs
    g = vs.HostCSR.rmat(6, 4, seed=seed)
    off, adj = np.asarray(g.offsets), np.asarray(g.adj)
    rng = np.random.default_rng(seed)
    src = rng.choice(g.V, size=9, replace=False).astype(np.uint32)
    M = rng.random((9, g.V)) < 0.2
    A = ot.adjacency_colmain(off, adj, g.V)
    for var_idx in ([0, 1], [1, 1, 0], [1]):
        exp = []
        for r in range(9):
            for c in range(g.V):
                if not M[r, c]:
                    continue
                bound = [(int(src[r]), c)[v] for v in var_idx]
                for w in range(g.V):
                    if all(A[w, b] for b in bound):
                        exp.append((int(src[r]), c, w))
        for kw in (dict(deps=[(A, None, None)] * len(var_idx)),
                   dict(deps=None, adjacency=[(off, adj)] * len(var_idx))):
            s, t, w = ot.mintersect_varrow(M, src, var_idx=var_idx, **kw)
            assert list(zip(s.tolist(), t.tolist(), w.tolist())) == exp
