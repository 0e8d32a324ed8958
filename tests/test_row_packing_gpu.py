"""Row packing of vs_vexpand: a call expands its rows sorted (within each tile)
by the source's 2-hop volume and puts the result matrices back into the
caller's order.  Checked against a call that keeps the caller's order
(VS_VX_KEEP_ORDER) and against the oracle: identical matrices in every
layout, row counts and digests, flat tables, per-hop layers and FindShortest
distances, for ragged row counts, several tiles, duplicate and isolated
sources."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from vertexsurge_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat13():
    g = vs.HostCSR.rmat(13, 16, seed=9)
    return g, vs.Graph.from_csr(g)


def _sources(g, S, seed):
    src = vs.sample_sources(g.V, S, seed=seed).astype(np.uint32)
    rng = np.random.default_rng(seed)
    src[rng.integers(0, S, size=S // 10)] = src[0]  # duplicates
    deg = np.diff(g.offsets.astype(np.int64))
    iso = np.flatnonzero(deg == 0)
    if len(iso):
        src[1::37] = iso[0]  # isolated sources: empty rows
    return rng.permutation(src)


def _same(a, b):
    assert a.num_matrices == b.num_matrices and a.k_actual_max == b.k_actual_max
    for i in range(a.num_matrices):
        assert a.distance(i) == b.distance(i)
        assert np.array_equal(a.export(i), b.export(i))
        assert np.array_equal(a.export(i, _lib.LAYOUT_DVSM), b.export(i, _lib.LAYOUT_DVSM))
        assert np.array_equal(a.row_counts(i), b.row_counts(i))
        assert np.array_equal(a.row_digests(i), b.row_digests(i))
        assert a.tuple_count(i) == b.tuple_count(i)


@pytest.mark.parametrize("path", ["shortest", "any"])
@pytest.mark.parametrize("tw,S", [(64, 3000), (16, 2500), (4, 700)])
def test_packed_equals_caller_order(orc, rmat13, path, tw, S):
    g, G = rmat13
    src = _sources(g, S, seed=S + tw)
    for k_min, k_max in ((1, 3), (0, 3), (2, 4), (1, 2)):  # k_max 2 expands unpacked
        a, sa = vs.vexpand(G, src, k_min, k_max, path, tile_words=tw)
        b, sb = vs.vexpand(G, src, k_min, k_max, path, tile_words=tw, keep_order=True)
        assert sa["te"] == sb["te"]
        _same(a, b)
        if (k_min, k_max) == (1, 3):
            mats, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, k_min, k_max, path)
            bits = np.unpackbits(mats[0].view(np.uint8), axis=1, bitorder="little")[:, :g.V]
            assert np.array_equal(a.row_counts(0), bits.sum(axis=1))
            sa_, ta_ = a.to_flat()
            sb_, tb_ = b.to_flat()
            assert np.array_equal(sa_, sb_) and np.array_equal(ta_, tb_)
        a.close()
        b.close()


@pytest.mark.parametrize("path", ["shortest", "any"])
def test_packed_layers_with_distance(rmat13, path):
    g, G = rmat13
    src = _sources(g, 1500, seed=5)
    a, _ = vs.vexpand(G, src, 1, 4, path, with_distance=True, tile_words=16)
    b, _ = vs.vexpand(G, src, 1, 4, path, with_distance=True, tile_words=16, keep_order=True)
    _same(a, b)
    fa = a.to_flat(with_distance=True)
    fb = b.to_flat(with_distance=True)
    assert all(np.array_equal(x, y) for x, y in zip(fa, fb))
    a.close()
    b.close()


def test_packed_find_shortest(orc, rmat13):
    g, G = rmat13
    src = _sources(g, 2000, seed=7)
    tgt = vs.sample_sources(g.V, 2000, seed=8).astype(np.uint32)
    d = vs.find_shortest(G, src, tgt, k_max=6)
    # the oracle's Shortest layers give every (source, target) distance up to 6
    mats, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, 1, 6, "shortest",
                                     with_distance=True)
    bits = np.unpackbits(mats.view(np.uint8), axis=2, bitorder="little")[:, :, :g.V]
    rows = np.arange(len(src))
    hit = bits[:, rows, tgt.astype(np.int64)]  # [layer, row]
    exp = np.where(hit.any(axis=0), hit.argmax(axis=0) + 1, -1)
    exp[src == tgt] = 0
    assert np.array_equal(d, exp)


@pytest.mark.parametrize("tw,S", [(64, 3000), (8, 1500)])
def test_packed_table_operators(rmat13, tw, S):
    # a packed result keeps its rows in expansion order and applies the
    # stored-row map when read: every table operator must see caller order
    g, G = rmat13
    src = _sources(g, S, seed=11 * tw)
    a, _ = vs.vexpand(G, src, 1, 3, "any", tile_words=tw)
    b, _ = vs.vexpand(G, src, 1, 3, "any", tile_words=tw, keep_order=True)
    for transpose in (False, True):
        ca, cb = a.to_colmain(transpose=transpose), b.to_colmain(transpose=transpose)
        assert np.array_equal(ca.export(), cb.export()), transpose
        assert [None if x is None else x.tolist() for x in ca.maps()] == \
               [None if x is None else x.tolist() for x in cb.maps()]
        ca.close()
        cb.close()
    cols = np.unique(src[::7])[:50]
    sa, sb = a.select_cols(cols), b.select_cols(cols)
    _same(sa, sb)
    for o in (None, "asc", "desc"):
        x, y = a.col_counts(order=o), b.col_counts(order=o)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
    assert np.array_equal(a.row_or(), b.row_or())
    rng = np.random.default_rng(S)
    for row, col in zip(rng.integers(0, S, 40), rng.integers(0, g.V, 40)):
        assert a.get(int(row), int(col)) == b.get(int(row), int(col))
    bits = a.export()
    r0 = int(np.flatnonzero(bits.any(axis=1))[0])
    c0 = int(np.flatnonzero(np.unpackbits(bits[r0].view(np.uint8), bitorder="little"))[0])
    assert a.get(r0, c0) and b.get(r0, c0)
    for x in (a, b, sa, sb):
        x.close()
