"""Batched VExpand (shared tiles, persistent kernel, replayed hipGraph) must give
the same per-row counts and digests as running every query alone, and as the
CPU oracle."""
import numpy as np
import pytest

import vertexsurge_amd as vs

pytestmark = pytest.mark.gpu


def make_queries(V, n, seed, smax=700, kmax=(1, 5), kmins=(0, 1)):
    rng = np.random.default_rng(seed)
    qs = []
    for q in range(n):
        S = int(np.exp(rng.uniform(np.log(1), np.log(smax))))
        k = int(rng.integers(kmax[0], kmax[1] + 1))
        kmin = int(rng.choice(kmins))
        qs.append(vs.BatchQuery(vs.sample_sources(V, S, seed * 1000 + q), kmin, k))
    return qs


@pytest.fixture(scope="module")
def g14():
    g = vs.HostCSR.rmat(14, 16, seed=4)
    return g, vs.Graph.from_csr(g)


def expected(orc, g, qs):
    counts, digests = [], []
    for q in qs:
        mats, ka, te, _ = orc.vexpand_rows(g.offsets, g.adj, q.sources, q.k_min, q.k_max, q.path)
        counts.append(orc.row_counts(mats[0]))
        digests.append(orc.row_digests(mats[0], g.V))
    return np.concatenate(counts), np.concatenate(digests)


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("sched", ["heuristic", "push", "pull"])
def test_batch_matches_oracle(orc, g14, use_graph, sched):
    g, G = g14
    qs = make_queries(g.V, 12, seed=5)
    c, d, st = vs.vexpand_batch(G, qs, sched=sched, use_graph=use_graph, tile_words=8)
    ec, ed = expected(orc, g, qs)
    assert np.array_equal(c, ec)
    assert np.array_equal(d, ed)
    assert st["num_tiles"] >= 1 and st["launches"] > 0


def test_batch_many_waves_and_general_kmin(orc, g14):
    g, G = g14
    qs = make_queries(g.V, 10, seed=9, kmins=(0, 1, 2, 3))
    # a tiny memory budget forces one slot per wave: many graph replays
    c, d, st = vs.vexpand_batch(G, qs, tile_words=4, mem_budget=1, use_graph=True)
    ec, ed = expected(orc, g, qs)
    assert st["num_waves"] == st["num_tiles"] and st["num_waves"] > 1
    assert np.array_equal(c, ec) and np.array_equal(d, ed)


def test_batch_any_and_mixed(orc, g14):
    g, G = g14
    qs = make_queries(g.V, 6, seed=3)
    for q in qs[::2]:
        q.path = "any"
    c, d, st = vs.vexpand_batch(G, qs, tile_words=16)
    ec, ed = expected(orc, g, qs)
    assert np.array_equal(c, ec) and np.array_equal(d, ed)


def test_batch_te_equals_sum_of_plain(g14):
    g, G = g14
    qs = make_queries(g.V, 8, seed=11)
    te = 0
    for q in qs:
        _, st = vs.vexpand(G, q.sources, q.k_min, q.k_max, q.path)
        te += st["te"]
    c, d, bst = vs.vexpand_batch(G, qs)
    assert bst["te"] == te


def test_batch_replays_are_identical(g14):
    g, G = g14
    qs = make_queries(g.V, 8, seed=12)
    c1, d1, _ = vs.vexpand_batch(G, qs, replays=1)
    c3, d3, st = vs.vexpand_batch(G, qs, replays=3)
    assert np.array_equal(c1, c3) and np.array_equal(d1, d3)


def test_batch_learned_scheduler_same_bits(orc, g14):
    g, G = g14
    s = vs.Scheduler(seed=5)
    G.set_scheduler(s)
    try:
        qs = make_queries(g.V, 8, seed=13)
        c, d, st = vs.vexpand_batch(G, qs, sched="learned")
        ec, ed = expected(orc, g, qs)
        assert np.array_equal(c, ec) and np.array_equal(d, ed)
        assert st["push_decisions"] + st["pull_decisions"] > 0
    finally:
        G.set_scheduler(None)


@pytest.mark.parametrize("keep_order", [True, False])
def test_batch_row_packing_order(orc, g14, keep_order):
    """Packing rows by k_max and source degree must not change any output:
    results come back in query order either way."""
    g, G = g14
    qs = make_queries(g.V, 9, seed=21, kmax=(2, 6), kmins=(0, 1, 2))
    for q in qs[1::3]:
        q.path = "any"
    c, d, st = vs.vexpand_batch(G, qs, tile_words=4, keep_order=keep_order)
    ec, ed = expected(orc, g, qs)
    assert np.array_equal(c, ec) and np.array_equal(d, ed)


def test_scheduler_handle_freed_while_attached(orc, g14):
    """The graph keeps its own reference to an attached scheduler: freeing the
    caller's handle must leave the learned path working (it used to leave the
    graph with a dangling pointer)."""
    g, G = g14
    s = vs.Scheduler(seed=7)
    G.set_scheduler(s)
    s.close()
    try:
        qs = make_queries(g.V, 6, seed=17)
        c, d, st = vs.vexpand_batch(G, qs, sched="learned")
        ec, ed = expected(orc, g, qs)
        assert np.array_equal(c, ec) and np.array_equal(d, ed)
        r, _ = vs.vexpand(G, qs[0].sources, qs[0].k_min, qs[0].k_max, sched="learned")
        assert np.array_equal(r.row_counts(0), ec[:len(qs[0].sources)])
        r.close()
    finally:
        G.set_scheduler(None)


def _same_result(a, b, V):
    """Two VExpand results hold the same rows, bit for bit, in every output."""
    assert (a.rows, a.cols, a.num_matrices) == (b.rows, b.cols, b.num_matrices)
    assert a.k_actual_max == b.k_actual_max
    if a.rows == 0:
        return
    assert np.array_equal(a.export(), b.export())
    assert np.array_equal(a.export(0, vs.LAYOUT_DVSM), b.export(0, vs.LAYOUT_DVSM))
    assert np.array_equal(a.row_counts(), b.row_counts())
    assert np.array_equal(a.row_digests(), b.row_digests())
    assert a.tuple_count() == b.tuple_count()
    sa, ta = a.to_flat()
    sb, tb = b.to_flat()
    assert np.array_equal(sa, sb) and np.array_equal(ta, tb)


@pytest.mark.parametrize("mode", ["one_query", "many_queries", "many_waves"])
@pytest.mark.parametrize("path,kmins", [("shortest", (0, 1)), ("shortest", (0, 1, 2, 3)),
                                        ("any", (0, 1, 2))])
def test_batch_results_equal_plain(orc, g14, mode, path, kmins):
    g, G = g14
    if mode == "one_query":
        src = vs.sample_sources(g.V, 1500, seed=3).astype(np.uint32)
        src[::50] = src[0]  # duplicate rows
        qs = [vs.BatchQuery(src, kmins[-1], 4, path)]
    else:
        qs = make_queries(g.V, 9, seed=21, smax=900, kmax=(1, 5), kmins=kmins)
        for q in qs:
            q.path = path
        qs.insert(3, vs.BatchQuery(np.zeros(0, dtype=np.uint32), 1, 3, path))  # empty query
    kw = dict(tile_words=8, mem_budget=1) if mode == "many_waves" else dict(tile_words=16)
    res, st = vs.vexpand_batch_results(G, qs, **kw)
    assert len(res) == len(qs) and st["num_tiles"] >= 1
    if mode == "many_waves":
        assert st["num_waves"] > 1
    counts, digests, _ = vs.vexpand_batch(G, qs, **kw)
    row = 0
    for q, r in zip(qs, res):
        p, _ = vs.vexpand(G, q.sources, q.k_min, q.k_max, q.path, tile_words=16)
        _same_result(r, p, g.V)
        n = len(q.sources)
        if n:
            assert np.array_equal(r.row_counts(), counts[row:row + n])
            assert np.array_equal(r.row_digests(), digests[row:row + n])
            mats, ka, _, _ = orc.vexpand_rows(g.offsets, g.adj, q.sources, q.k_min, q.k_max,
                                              q.path)
            assert r.k_actual_max == ka
            assert np.array_equal(r.export(), mats[0])
        row += n
        p.close()
        r.close()


def test_batch_results_table_operators(g14):
    g, G = g14
    qs = make_queries(g.V, 5, seed=77, smax=600, kmax=(2, 4), kmins=(1,))
    res, _ = vs.vexpand_batch_results(G, qs, tile_words=8)
    for q, r in zip(qs, res):
        p, _ = vs.vexpand(G, q.sources, q.k_min, q.k_max, q.path, tile_words=8,
                          keep_order=True)
        for transpose in (False, True):
            a, b = r.to_colmain(transpose=transpose), p.to_colmain(transpose=transpose)
            assert np.array_equal(a.export(), b.export())
            a.close()
            b.close()
        x, y = r.col_counts(order="desc"), p.col_counts(order="desc")
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
        assert np.array_equal(r.row_or(), p.row_or())
        r.close()
        p.close()


def test_batch_results_replayed_graph_and_learned(g14):
    # several replays of the captured graph rewrite the same result matrices
    g, G = g14
    qs = make_queries(g.V, 6, seed=13, smax=2000, kmax=(2, 5))
    r1, _ = vs.vexpand_batch_results(G, qs, replays=3, use_graph=True)
    r2, _ = vs.vexpand_batch_results(G, qs, use_graph=False, sched="pull")
    for a, b in zip(r1, r2):
        _same_result(a, b, g.V)
        a.close()
        b.close()


def test_autotune_batch_exact_and_no_slower(g14):
    g, G = g14
    saved = G.get_tuning()
    qs = make_queries(g.V, 8, seed=41, smax=3000, kmax=(2, 5))
    c0, d0, _ = vs.vexpand_batch(G, qs)
    tuning, best, start = vs.autotune_batch(G, qs)
    try:
        # the starting tuning is one of the candidates: never slower by the
        # tuner's own clock; the pick is stored in the graph
        assert 0 < best <= start
        assert G.get_tuning() == tuning
        assert tuning["sparse_pct"] in (35, 50, 65) and tuning["push_alpha"] in (4, 16, 64)
        assert tuning["wpush_pml"] in (40, 200, 2000)
        c1, d1, _ = vs.vexpand_batch(G, qs)
        assert np.array_equal(c0, c1) and np.array_equal(d0, d1)
    finally:
        G.set_tuning(**saved)


@pytest.mark.parametrize("pml", [1000, 200])
def test_batch_word_push(orc, g14, pml):
    """Batched engine with the word push (decided on the device per slot and
    hop): the result matrices equal the plain engine's without it, the counts
    and digests the oracle's; hops that are a tile's last never push."""
    g, G = g14
    saved = G.get_tuning()
    qs = make_queries(g.V, 10, seed=17, smax=3000, kmax=(2, 5), kmins=(0, 1, 2))
    try:
        G.set_tuning(**{**saved, "wpush_pml": 2000})
        plain = [vs.vexpand(G, q.sources, q.k_min, q.k_max, q.path, sched="pull")[0]
                 for q in qs]
        G.set_tuning(**{**saved, "wpush_pml": pml})
        for use_graph in (True, False):
            rs, st = vs.vexpand_batch_results(G, qs, sched="pull", use_graph=use_graph,
                                              tile_words=16)
            if pml == 1000:
                assert st["wpush_words"] > 0
            for a, b in zip(rs, plain):
                _same_result(a, b, g.V)
                a.close()
        c, d, _ = vs.vexpand_batch(G, qs, sched="pull", tile_words=16)
        ec, ed = expected(orc, g, qs)
        assert np.array_equal(c, ec) and np.array_equal(d, ed)
    finally:
        G.set_tuning(**saved)
        for b in plain:
            b.close()
