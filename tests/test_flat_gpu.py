"""Flat export of VExpand results (the CSR strategy's (start, target[, distance])
table, merge_to_Flat, vexpand.cpp:380-446), checked against the oracle's
matrices and against closed forms."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from conftest import load_golden

pytestmark = pytest.mark.gpu


def expected_flat(mats, sources, V, k_min=None):
    """Tuples of the union of the row-major matrices, by row then column; the
    distance of a pair is k_min + the first matrix holding it."""
    S = mats.shape[1]
    bits = np.unpackbits(mats.view(np.uint8), axis=2, bitorder="little")[:, :, :V].astype(bool)
    union = bits.any(axis=0)
    rows, cols = np.nonzero(union)  # row-major order: by row, then column
    start = np.asarray(sources, dtype=np.uint32)[rows]
    dist = None
    if k_min is not None:
        first = np.argmax(bits[:, rows, cols], axis=0)
        dist = (k_min + first).astype(np.int32)
    assert len(rows) == union.sum() and S == len(sources)
    return start, cols.astype(np.uint32), dist


@pytest.fixture(scope="module")
def rmat14():
    g = vs.HostCSR.rmat(14, 16, seed=5)
    return g, vs.Graph.from_csr(g)


@pytest.mark.parametrize("path", ["shortest", "any"])
@pytest.mark.parametrize("tw", [16, 64])
def test_flat_merged(orc, rmat14, path, tw):
    g, G = rmat14
    src = vs.sample_sources(g.V, 700, seed=3)  # ragged words and tiles
    res, st = vs.vexpand(G, src, 1, 3, path, tile_words=tw)
    mats, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, 1, 3, path)
    s, t = res.to_flat()
    es, et, _ = expected_flat(mats, src, g.V)
    assert np.array_equal(s, es) and np.array_equal(t, et)
    assert len(s) == res.tuple_count()
    with pytest.raises(vs.VSError) as e:
        res.to_flat(with_distance=True)
    assert e.value.code == vs._lib.VS_ERR_INVALID


@pytest.mark.parametrize("path", ["shortest", "any"])
def test_flat_with_distance(orc, rmat14, path):
    g, G = rmat14
    src = vs.sample_sources(g.V, 300, seed=4)
    k_min, k_max = 1, 4
    res, st = vs.vexpand(G, src, k_min, k_max, path, with_distance=True)
    mats, ka, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, k_min, k_max, path,
                                      with_distance=True)
    s, t, d = res.to_flat(with_distance=True)
    es, et, ed = expected_flat(mats, src, g.V, k_min=k_min)
    assert np.array_equal(s, es) and np.array_equal(t, et) and np.array_equal(d, ed)
    if path == "shortest":
        # disjoint layers: as many tuples as the layers hold together
        assert len(s) == sum(res.tuple_count(i) for i in range(res.num_matrices))


def test_flat_path_graph_closed_form():
    """Path 0-1-...-69: from source 10, hops 1..3 reach 7..13 minus 10 with
    distance |v - 10|; Any with k_min = 2 also reaches 10 (back and forth)."""
    gd = load_golden("path70")
    G = vs.Graph(gd["offsets"], gd["adj"], symmetric=True)
    res, _ = vs.vexpand(G, [10], 1, 3, "shortest", with_distance=True)
    s, t, d = res.to_flat(with_distance=True)
    assert t.tolist() == [7, 8, 9, 11, 12, 13]
    assert d.tolist() == [3, 2, 1, 1, 2, 3] and set(s.tolist()) == {10}
    res, _ = vs.vexpand(G, [10], 2, 3, "any", with_distance=True)
    s, t, d = res.to_flat(with_distance=True)
    # walks of length 2: 8, 10, 12; length 3 adds 7, 9, 11, 13
    assert t.tolist() == [7, 8, 9, 10, 11, 12, 13]
    assert d.tolist() == [3, 2, 3, 2, 3, 2, 3]


def test_flat_selected_columns(orc, rmat14):
    g, G = rmat14
    src = vs.sample_sources(g.V, 200, seed=6)
    res, _ = vs.vexpand(G, src, 1, 2, "shortest")
    cols = np.array([5, 77, 1000, 3, 77], dtype=np.uint32)
    sel = res.select_cols(cols)
    s, t = sel.to_flat()
    mats, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, 1, 2, "shortest")
    bits = np.unpackbits(mats[0].view(np.uint8), axis=1, bitorder="little")[:, :g.V]
    rows, k = np.nonzero(bits[:, cols.astype(np.int64)])
    assert np.array_equal(s, src[rows]) and np.array_equal(t, cols[k])


def test_flat_empty():
    gd = load_golden("two_comp")
    G = vs.Graph(gd["offsets"], gd["adj"], symmetric=True)
    res, _ = vs.vexpand(G, np.zeros(0, dtype=np.uint32), 1, 2, "shortest")
    s, t = res.to_flat()
    assert len(s) == 0 and len(t) == 0
