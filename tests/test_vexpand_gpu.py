"""GPU parity of VExpand (through the C ABI) against the CPU oracle.

Bit-exact on every output: merged result matrix (row-major export), per-hop
layers (with_distance), per-source counts, digests, reference DVSM layout
(checked against oracle B, the reference-algorithm restatement), k_actual_max,
traversed-edge count and the compulsory-byte counter.
"""
import numpy as np
import pytest

import vertexsurge_amd as vs
from conftest import golden_names, load_golden

pytestmark = pytest.mark.gpu


def reverse_csr(off, adj):
    V = len(off) - 1
    src = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64))
    order = np.lexsort((src, adj))
    radj = src[order]
    roff = np.zeros(V + 1, dtype=np.uint64)
    np.add.at(roff, adj.astype(np.int64) + 1, 1)
    return np.cumsum(roff, dtype=np.uint64), radj


def union_csr(off, adj, roff, radj):
    V = len(off) - 1
    s1 = np.repeat(np.arange(V, dtype=np.uint64), np.diff(off).astype(np.int64))
    s2 = np.repeat(np.arange(V, dtype=np.uint64), np.diff(roff).astype(np.int64))
    key = np.unique(np.concatenate([s1 * V + adj, s2 * V + radj]))
    uo = np.zeros(V + 1, dtype=np.uint64)
    np.add.at(uo, (key // V).astype(np.int64) + 1, 1)
    return np.cumsum(uo, dtype=np.uint64), (key % V).astype(np.uint32)


def check(orc, graph, off, adj, src, k_min, k_max, path, with_distance=False,
          sched="heuristic", tile_words=0, check_bytes=True):
    res, st = vs.vexpand(graph, src, k_min, k_max, path, with_distance=with_distance,
                         sched=sched, tile_words=tile_words)
    mats, ka, te, pop = orc.vexpand_rows(off, adj, src, k_min, k_max, path, with_distance)
    assert res.k_actual_max == ka == st["k_actual_max"]
    assert res.num_matrices == (len(mats) if with_distance else 1)
    V = len(off) - 1
    for i in range(res.num_matrices):
        got = res.export(i, vs.LAYOUT_ROWMAJOR)
        assert np.array_equal(got, mats[i]), f"matrix {i} differs"
        assert np.array_equal(res.row_counts(i), orc.row_counts(mats[i]))
        assert res.tuple_count(i) == int(orc.row_counts(mats[i]).sum())
        if with_distance:
            assert res.distance(i) == k_min + i
    assert st["te"] == te
    if check_bytes and not with_distance:
        # the compulsory-byte counter counts nonzero words of the layout the
        # call expands; the oracle's is the caller's row order, so compare a
        # call that keeps that order (its results are the same)
        lay, ka2, _, _ = orc.vexpand_rows(off, adj, src, 0, k_max, path, with_distance=True)
        assert ka2 == ka
        kres, kst = vs.vexpand(graph, src, k_min, k_max, path, sched=sched,
                               tile_words=tile_words, keep_order=True)
        assert np.array_equal(kres.export(0, vs.LAYOUT_ROWMAJOR), mats[0])
        kres.close()
        exp_b = orc.bytes_alg(off, len(src), lay, ka, k_min, path, mats[0],
                              kst["tile_words"] * 64)
        assert kst["bytes_alg"] == exp_b and kst["te"] == te
    return res, st, mats


@pytest.fixture(scope="module")
def rmat16():
    g = vs.HostCSR.rmat(16, 16, seed=1)
    G = vs.Graph.from_csr(g)
    return g, G


@pytest.mark.parametrize("name", golden_names())
@pytest.mark.parametrize("path", ["shortest", "any"])
def test_golden_graphs(orc, name, path):
    gd = load_golden(name)
    sym = bool(gd["symmetric"][0])
    G = vs.Graph(gd["offsets"], gd["adj"], symmetric=sym)
    for k_min, k_max in [(0, 3), (1, 4), (2, 2), (3, 1)]:
        for sched in ("push", "pull", "heuristic"):
            check(orc, G, gd["offsets"], gd["adj"], gd["sources"], k_min, k_max, path,
                  sched=sched)


def test_directed_directions(orc):
    gd = load_golden("dchain")
    off, adj = gd["offsets"], gd["adj"]
    G = vs.Graph(off, adj, symmetric=False)
    roff, radj = reverse_csr(off, adj)
    uoff, uadj = union_csr(off, adj, roff, radj)
    src = gd["sources"]
    for direction, (o, a) in {"forward": (off, adj), "backward": (roff, radj),
                              "both": (uoff, uadj)}.items():
        for sched in ("push", "pull"):
            res, st = vs.vexpand(G, src, 1, 4, "shortest", direction=direction, sched=sched)
            mats, ka, te, _ = orc.vexpand_rows(o, a, src, 1, 4, "shortest")
            assert np.array_equal(res.export(0, vs.LAYOUT_ROWMAJOR), mats[0]), direction
            assert st["te"] == te


@pytest.mark.parametrize("path", ["shortest", "any"])
@pytest.mark.parametrize("sched", ["heuristic", "push", "pull"])
def test_rmat16_config1(orc, rmat16, path, sched):
    """Plain VExpand, the CPU-runnable workload: R-MAT 16, EF 16, 512 sources, hops 1..3."""
    g, G = rmat16
    src = vs.sample_sources(g.V, 512, seed=2024)
    res, st, mats = check(orc, G, g.offsets, g.adj, src, 1, 3, path, sched=sched)
    # reference layout, against oracle B (the reference algorithm restatement)
    blocks = orc.EdgeBlocks(g.offsets, g.adj, nblocks=200)
    dv, ka, _, _ = blocks.vexpand(src, 1, 3, path)
    assert np.array_equal(res.export(0, vs.LAYOUT_DVSM), dv)
    assert np.array_equal(res.row_digests(0), orc.row_digests(mats[0], g.V))


@pytest.mark.parametrize("tw", [1, 2, 4, 8, 16, 32, 64])
def test_tile_widths(orc, rmat16, tw):
    g, G = rmat16
    src = vs.sample_sources(g.V, 1000, seed=tw)  # ragged last word and tile
    res, st, mats = check(orc, G, g.offsets, g.adj, src, 0, 3, "shortest", tile_words=tw)
    assert st["tile_words"] == min(tw, 16)
    native = res.export(0, vs.LAYOUT_NATIVE)
    assert np.array_equal(native, orc.rowmajor_to_native(mats[0], g.V, st["tile_words"]))


def test_with_distance_layers(orc, rmat16):
    g, G = rmat16
    src = vs.sample_sources(g.V, 300, seed=77)
    for path in ("shortest", "any"):
        for k_min, k_max in [(0, 3), (2, 4), (5, 3)]:
            check(orc, G, g.offsets, g.adj, src, k_min, k_max, path, with_distance=True)


def test_edge_cases(orc, rmat16):
    g, G = rmat16
    deg = g.degrees()
    iso = np.nonzero(deg == 0)[0][:3]
    hub = int(np.argmax(deg))
    # duplicates, isolated vertices, the biggest hub, S = 1, S = 65
    cases = [np.array([hub, hub, 5, 5], dtype=np.uint32), iso.astype(np.uint32),
             np.array([hub], dtype=np.uint32),
             vs.sample_sources(g.V, 65, seed=3)]
    for src in cases:
        for path in ("shortest", "any"):
            for sched in ("push", "pull"):
                check(orc, G, g.offsets, g.adj, src, 0, 3, path, sched=sched)
    # empty source list: empty result, status OK (the reference warns and skips)
    res, st = vs.vexpand(G, np.zeros(0, dtype=np.uint32), 1, 3)
    assert res.rows == 0 and res.k_actual_max == 0 and st["te"] == 0
    # k_max beyond the eccentricity: Shortest stops early, same bits
    src = vs.sample_sources(g.V, 64, seed=8)
    r1, s1, m1 = check(orc, G, g.offsets, g.adj, src, 0, 30, "shortest")
    assert s1["k_actual_max"] < 30


def test_invalid_arguments(rmat16):
    g, G = rmat16
    with pytest.raises(vs.VSError) as e:
        vs.vexpand(G, np.array([g.V], dtype=np.uint32), 1, 2)
    assert e.value.code == 1
    with pytest.raises(vs.VSError):
        vs.vexpand(G, np.array([1], dtype=np.uint32), 1, 2, tile_words=3)
    with pytest.raises(vs.VSError):
        vs.vexpand(G, np.array([1], dtype=np.uint32), 1, 64)


def test_find_shortest(orc, rmat16):
    g, G = rmat16
    src = vs.sample_sources(g.V, 200, seed=5)
    rng = np.random.default_rng(1)
    tgt = rng.integers(0, g.V, size=200).astype(np.uint32)
    tgt[:5] = src[:5]  # distance 0
    dist = vs.find_shortest(G, src, tgt, k_max=50)
    for r in range(200):
        d = orc.bfs_dist(g.offsets, g.adj, int(src[r]))
        assert dist[r] == d[tgt[r]], r
    # a small k_max leaves far targets at -1
    d3 = vs.find_shortest(G, src, tgt, k_max=2)
    assert np.array_equal(d3, np.where((dist >= 0) & (dist <= 2), dist, -1))


def test_repeat_calls_reuse_pool(orc, rmat16):
    g, G = rmat16
    src = vs.sample_sources(g.V, 512, seed=2024)
    outs = []
    for sched in ("push", "pull", "push", "heuristic"):
        res, st = vs.vexpand(G, src, 1, 3, "shortest", sched=sched)
        outs.append(res.export(0, vs.LAYOUT_ROWMAJOR))
        res.close()
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("pml", [1000, 200, 2000])
@pytest.mark.parametrize("path", ["shortest", "any"])
@pytest.mark.parametrize("tw", [16, 64])
def test_word_push(orc, rmat16, pml, path, tw):
    """Word push on the pull hops that gather whole rows (the cheapest words
    scattered into N, the rest pulled, a word-push epilogue after the
    chunked-vertex one): bit-exact against the oracle with budgets that push
    most words (1000 per mille of E), some (200) and none (2000 = off), over
    R-MAT 16 with chunked in-lists, a ragged last word, k_min 0..2."""
    g, G = rmat16
    saved = G.get_tuning()
    try:
        G.set_tuning(**{**saved, "wpush_pml": pml})
        src = vs.sample_sources(g.V, 1000, seed=11)
        pushed = 0
        for k_min, k_max in [(1, 4), (2, 4), (0, 3)]:
            res, st, _ = check(orc, G, g.offsets, g.adj, src, k_min, k_max, path,
                               sched="pull", tile_words=tw)
            pushed += sum(st["wpush_words"])
            res.close()
        if pml > 1000:
            assert pushed == 0
        if pml == 1000:
            assert pushed > 0
    finally:
        G.set_tuning(**saved)


def test_word_push_layers_and_find_shortest(orc, rmat16):
    """with_distance layers and FindShortest distances with the word push on."""
    g, G = rmat16
    saved = G.get_tuning()
    try:
        G.set_tuning(**{**saved, "wpush_pml": 1000})
        src = vs.sample_sources(g.V, 300, seed=78)
        for path in ("shortest", "any"):
            check(orc, G, g.offsets, g.adj, src, 1, 4, path, with_distance=True, sched="pull")
        rng = np.random.default_rng(3)
        tgt = rng.integers(0, g.V, size=300).astype(np.uint32)
        dist = vs.find_shortest(G, src, tgt, k_max=50)
        for r in range(0, 300, 7):
            d = orc.bfs_dist(g.offsets, g.adj, int(src[r]))
            assert dist[r] == d[tgt[r]], r
    finally:
        G.set_tuning(**saved)


def test_word_push_directed_and_multi_tile(orc, rmat16):
    """Word push where the push CSR is not the pull CSR (a directed graph, all
    three directions) and over several tiles of one call (3000 rows at tile
    width 16), with a budget that pushes most words."""
    rng = np.random.default_rng(21)
    V = 3000
    e = np.unique(rng.integers(0, V, size=(60000, 2)).astype(np.uint64) @
                  np.array([V, 1], dtype=np.uint64))
    e = e[e // V != e % V]
    h = vs.HostCSR.from_edges(V, (e // V).astype(np.uint32), (e % V).astype(np.uint32),
                              symmetrize=False)
    off, adj = h.offsets.copy(), h.adj.copy()
    G = vs.Graph(off, adj, symmetric=False)
    roff, radj = reverse_csr(off, adj)
    uoff, uadj = union_csr(off, adj, roff, radj)
    src = vs.sample_sources(V, 700, seed=4)
    G.set_tuning(wpush_pml=1000)
    for direction, (o, a) in {"forward": (off, adj), "backward": (roff, radj),
                              "both": (uoff, uadj)}.items():
        res, st = vs.vexpand(G, src, 1, 5, "shortest", direction=direction, sched="pull")
        mats, ka, te, _ = orc.vexpand_rows(o, a, src, 1, 5, "shortest")
        assert np.array_equal(res.export(0, vs.LAYOUT_ROWMAJOR), mats[0]), direction
        assert st["te"] == te and res.k_actual_max == ka
    g, G16 = rmat16
    saved = G16.get_tuning()
    try:
        G16.set_tuning(**{**saved, "wpush_pml": 1000})
        src = vs.sample_sources(g.V, 3000, seed=12)
        for path in ("shortest", "any"):
            res, st, _ = check(orc, G16, g.offsets, g.adj, src, 1, 4, path, sched="pull",
                               tile_words=16)
            assert st["num_tiles"] == 3
            res.close()
    finally:
        G16.set_tuning(**saved)
