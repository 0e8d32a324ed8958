"""GPU parity of the table operators (TableConvert, MatrixTableSelect,
MatrixColumnCount, MatrixRowOrSum, Mintersect over ColMain columns) against
the oracle in oracle/tables.py."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from oracle import tables as ot
from test_tables_oracle import _flat_mat_mat_deps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g12():
    g = vs.HostCSR.rmat(12, 16, seed=8)
    return g, vs.Graph.from_csr(g)


def _layer(orc, g, src, k_min, k_max, path="shortest"):
    mats, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, k_min, k_max, path)
    bits = np.unpackbits(mats[0].view(np.uint8), axis=1, bitorder="little")
    return bits[:, :g.V].astype(bool)


@pytest.mark.parametrize("tile_words", [4, 64])
@pytest.mark.parametrize("transpose", [False, True])
def test_table_convert(orc, g12, tile_words, transpose):
    g, G = g12
    src = vs.sample_sources(g.V, 700, 11)
    r, _ = vs.vexpand(G, src, 1, 3, tile_words=tile_words)
    M = _layer(orc, g, src, 1, 3)
    cm = r.to_colmain(transpose=transpose)
    E, rm, cmap = ot.table_convert(M, src, None, transpose)
    assert (cm.rows, cm.cols) == E.shape
    assert np.array_equal(cm.export(), ot.colmain_words(E))
    got_rm, got_cm = cm.maps()
    assert (got_rm is None) == (rm is None) and (got_cm is None) == (cmap is None)
    if rm is not None:
        assert np.array_equal(got_rm, rm)
    if cmap is not None:
        assert np.array_equal(got_cm, cmap)
    cm.close()
    r.close()


def test_select_cols_and_reductions(orc, g12):
    g, G = g12
    src = vs.sample_sources(g.V, 300, 12)
    r, _ = vs.vexpand(G, src, 0, 2, tile_words=4)
    M = _layer(orc, g, src, 0, 2)
    rng = np.random.default_rng(2)
    ids = rng.integers(0, g.V, size=500).astype(np.uint32)
    ids[7] = ids[3]  # a repeated column
    s = r.select_cols(ids)
    E, _, cmap = ot.select_cols(M, src, ids)
    assert (s.rows, s.cols) == E.shape
    got = s.export()
    assert np.array_equal(np.unpackbits(got.view(np.uint8), axis=1, bitorder="little")
                          [:, :len(ids)].astype(bool), E)
    assert np.array_equal(s.col_map(), ids)
    assert np.array_equal(s.row_counts(), E.sum(axis=1).astype(np.uint64))
    for order in (None, "asc", "desc"):
        gi, gc = r.col_counts(order=order)
        ei, ec = ot.column_count(M, None, order)
        assert np.array_equal(gi, ei) and np.array_equal(gc, ec), order
        gi, gc = s.col_counts(order=order)
        ei, ec = ot.column_count(E, cmap, order)
        assert np.array_equal(gi, ei) and np.array_equal(gc, ec), order
    assert np.array_equal(r.row_or(), ot.row_or_sum(M, None))
    assert np.array_equal(s.row_or(), ot.row_or_sum(E, cmap))
    cm = r.to_colmain(transpose=True)
    Et, _, cmt = ot.table_convert(M, src, None, True)
    for order in (None, "desc"):
        gi, gc = cm.col_counts(order=order)
        ei, ec = ot.column_count(Et, cmt, order)
        assert np.array_equal(gi, ei) and np.array_equal(gc, ec)
    assert np.array_equal(cm.row_or(), ot.row_or_sum(Et, cmt))
    for h in (cm, s, r):
        h.close()


def test_mintersect_colmat_reference_fixture(g12):
    _, G = g12
    deps, base = _flat_mat_mat_deps()
    mats = [vs.ColMat.from_bool(G, M, col_map=cmap) for M, _, cmap in deps]
    counts, offsets, lists, st = vs.mintersect_colmat(mats, base)
    assert counts.tolist() == [1, 1, 1, 0, 0, 0, 0, 0]
    assert lists.tolist() == [1, 1, 1] and st["total"] == 3
    # 8 of 10 columns mapped: the last two report their index
    M, _, cmap = deps[0]
    for order in (None, "asc", "desc"):
        gi, gc = mats[0].col_counts(order)
        ei, ec = ot.column_count(M, cmap, order)
        assert np.array_equal(gi, ei) and np.array_equal(gc, ec)
    assert np.array_equal(mats[0].row_or(), ot.row_or_sum(M, cmap))
    assert np.array_equal(mats[0].maps()[1], cmap)


def test_mintersect_colmat_adjacency_vs_sorted(orc):
    g = vs.HostCSR.rmat(10, 16, seed=3)
    G = vs.Graph.from_csr(g)
    M = ot.adjacency_colmain(g.offsets, g.adj, g.V)
    cm = vs.ColMat.from_bool(G, M)
    tup = vs.sample_paths3(g.offsets, g.adj, 4000, seed=9)
    tup[5, 1] = g.V + 3  # a missing column
    counts, offsets, lists, _ = vs.mintersect_colmat([cm, cm, cm], tup)
    c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj)] * 3, tup)
    assert np.array_equal(counts, c2) and np.array_equal(offsets, o2)
    assert np.array_equal(lists, l2)
    c3, _, l3, _ = vs.mintersect([G, G, G], tup)
    assert np.array_equal(counts, c3) and np.array_equal(lists, l3)


def test_pattern_vexpand_convert_intersect(orc, g12):
    """a -[1..2]-> c <-[1..2]- b: VExpand, TableConvert (transpose), then
    Mintersect over the two reach columns of every (a, b) pair."""
    g, G = g12
    src = vs.sample_sources(g.V, 200, 21)
    r, _ = vs.vexpand(G, src, 1, 2, path="any", tile_words=2)
    M = _layer(orc, g, src, 1, 2, "any")
    cm = r.to_colmain(transpose=True)
    rng = np.random.default_rng(5)
    base = src[rng.integers(0, len(src), size=(600, 2))]
    counts, offsets, lists, st = vs.mintersect_colmat([cm, cm], base)
    E, rm, cmap = ot.table_convert(M, src, None, True)
    c2, o2, l2 = ot.mintersect_colmain([(E, rm, cmap)] * 2, base)
    assert np.array_equal(counts, c2) and np.array_equal(offsets, o2)
    assert np.array_equal(lists, l2)
    assert st["bytes_alg"] > 0
    cm.close()
    r.close()


@pytest.mark.parametrize("rows", [63, 64 * 4096 + 70])
def test_mintersect_colmat_chunks_and_row_map(rows, g12):
    """Columns longer than one wave chunk (4096 words) and a row map."""
    _, G = g12
    rng = np.random.default_rng(rows)
    cols = 5
    M1 = rng.random((rows, cols)) < 0.6
    M2 = rng.random((rows, cols)) < 0.6
    rmap = rng.permutation(rows * 2)[:rows].astype(np.uint32)
    cmap = np.array([9, 4, 9, 1, 7], dtype=np.uint32)  # 9 twice: the first wins
    a = vs.ColMat.from_bool(G, M1, row_map=rmap, col_map=cmap)
    b = vs.ColMat.from_bool(G, M2)
    base = rng.integers(0, 11, size=(40, 2)).astype(np.uint32)
    counts, offsets, lists, _ = vs.mintersect_colmat([a, b], base)
    c2, o2, l2 = ot.mintersect_colmain([(M1, rmap, cmap), (M2, None, None)], base)
    assert np.array_equal(counts, c2) and np.array_equal(offsets, o2)
    assert np.array_equal(lists, l2)
    assert np.array_equal(a.to_bool(), M1)


def test_colmat_upload_clears_padding(g12):
    _, G = g12
    w = np.full((3, 8), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    m = vs.ColMat.create(G, 70, 3, w)
    out = m.export()
    assert np.all(out[:, 0] == np.uint64(0xFFFFFFFFFFFFFFFF))
    assert np.all(out[:, 1] == np.uint64(0x3F)) and np.all(out[:, 2:] == 0)
    assert m.get(69, 2) and m.get(0, 0)


@pytest.fixture(scope="module")
def adj12(g12):
    g, G = g12
    A = ot.adjacency_colmain(g.offsets, g.adj, g.V)
    cm = vs.ColMat.from_bool(G, A)
    yield A, cm
    cm.close()


@pytest.mark.parametrize("tile_words,keep_order", [(4, False), (4, True), (64, False)])
def test_mintersect_varrow_vs_oracle(orc, g12, adj12, tile_words, keep_order):
    """Mintersect with a VarRow base: every set bit of a 1-hop VExpand result
    is a (start, target) edge; ColMain adjacency and edge-relation
    dependencies over the start or target vertex, against the oracle (rows
    packed or in caller order, several tiles)."""
    g, G = g12
    A, cm = adj12
    src = vs.sample_sources(g.V, 700, 31).astype(np.uint32)
    src[::40] = src[2]  # repeated sources
    r, _ = vs.vexpand(G, src, 1, 1, tile_words=tile_words, keep_order=keep_order)
    M = _layer(orc, g, src, 1, 1)
    for var_idx in ([0, 1], [1], [1, 0, 1]):
        nd = len(var_idx)
        es, et, ew = ot.mintersect_varrow(M, src, [(A, None, None)] * nd, var_idx)
        for deps in ([cm] * nd, [G] * nd):
            s, t, w = vs.mintersect_varrow(r, deps, var_idx)
            assert len(s) == len(es) > 0, (var_idx, type(deps[0]))
            assert np.array_equal(s, es) and np.array_equal(t, et), var_idx
            assert np.array_equal(w, ew), var_idx
            assert vs.mintersect_varrow(r, deps, var_idx, with_lists=False) == len(es)
    r.close()


def test_mintersect_varrow_row_and_col_maps(orc, g12):
    """ColMain dependencies with a row map (new vertices go through it) and a
    column map (a bound vertex with no column gives the empty set): the
    transposed TableConvert of another VExpand."""
    g, G = g12
    src = vs.sample_sources(g.V, 250, 41)
    r, _ = vs.vexpand(G, src, 1, 1, tile_words=2)
    M = _layer(orc, g, src, 1, 1)
    src2 = vs.sample_sources(g.V, 900, 42)
    r2, _ = vs.vexpand(G, src2, 1, 2, path="any", tile_words=8)
    M2 = _layer(orc, g, src2, 1, 2, "any")
    cm = r2.to_colmain(transpose=True)  # rows: reached vertices, columns: src2
    E, rm, cmap = ot.table_convert(M2, src2, None, True)
    for var_idx in ([0], [1, 0]):
        nd = len(var_idx)
        es, et, ew = ot.mintersect_varrow(M, src, [(E, rm, cmap)] * nd, var_idx)
        s, t, w = vs.mintersect_varrow(r, [cm] * nd, var_idx)
        assert np.array_equal(s, es) and np.array_equal(t, et) and np.array_equal(w, ew)
    cm.close()
    r2.close()
    r.close()


def test_mintersect_varrow_capacity_and_errors(g12, adj12):
    import ctypes as C
    from vertexsurge_amd import _lib
    g, G = g12
    _, cm = adj12
    r, _ = vs.vexpand(G, vs.sample_sources(g.V, 64, 5), 1, 1)
    n = vs.mintersect_varrow(r, [G, G], [0, 1], with_lists=False)
    assert n > 1
    out = [np.zeros(n, dtype=np.uint32) for _ in range(3)]
    got = C.c_uint64()
    vi = np.array([0, 1], dtype=np.uint32)
    arr = (C.c_void_p * 2)(G.handle, G.handle)
    rc = _lib.load().vs_mintersect_varrow(r._h, 0, None, arr, 2,
                                         vi.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         *[o.ctypes.data_as(C.POINTER(C.c_uint32)) for o in out],
                                         n - 1, C.byref(got), None)
    assert rc == _lib.VS_ERR_CAPACITY and got.value == n
    with pytest.raises(ValueError):
        vs.mintersect_varrow(r, [cm, G], [0, 1])
    with pytest.raises(RuntimeError):
        vs.mintersect_varrow(r, [G], [2])
    r.close()
