"""Parity at the benchmark's own sizes, on the benchmark's own inputs.

Graph: symmetrized R-MAT scale 24, edge factor 32 (a/b/c = .57/.19/.19, seed
as in bench.py): V = 16,777,216, E = 1,023,750,138 directed entries.

* VExpand, 4096 seeded sources (one full 4096-row tile at tile width 64),
  Shortest, hops 1..4: every row's full reachability, in the reference's
  DenseVarStackMatrix words, against oracle B (the reference's stacked,
  destination-partitioned CPU algorithm restated) -- for the default call
  (rows packed by 2-hop volume), the caller-order call, tile widths 16 and 32
  under forced pull / push, and the batched engine; per-row counts of all
  4096 rows against the oracle's (Expand<DVSM>::expand_dense,
  expand.cpp:263-311; merge_to_MT, vexpand.cpp:298-354).
* Mintersect: all 4,194,304 sampled 2-hop paths (a, b, c), counts and
  sorted lists against the set-intersection oracle, one-pass and two-pass
  (Mintersect::run_internal, mintersect.cpp:205-243).
* The 256-query batch: per-row counts and digests of the batched engine equal
  those of every query run alone, and full rows of a seeded subset of the
  queries equal oracle B's.
* One full tile at width 64 on R-MAT 19 under every schedule (the width at
  which an earlier unpacking kernel once corrupted bit 63 of a few rows).

The module holds one R-MAT 24 graph for all its tests (about 24 s to build
on 16 host threads) and oracle B's edge blocks (about 23 s).
"""
import time

import numpy as np
import pytest

import bench
import vertexsurge_amd as vs
from vertexsurge_amd import _lib

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[scale] {msg}", flush=True)


@pytest.fixture(scope="module")
def rmat24():
    t = time.time()
    g = vs.HostCSR.rmat(24, 32, 0.57, 0.19, 0.19, bench.GRAPH_SEED)
    G = vs.Graph.from_csr(g)
    log(f"R-MAT 24 V={g.V} E={g.E} built and uploaded in {time.time() - t:.1f}s")
    yield g, G
    G.close()
    g.close()


@pytest.fixture(scope="module")
def blocks24(orc, rmat24):
    g, _ = rmat24
    t = time.time()
    b = orc.EdgeBlocks(g.offsets, g.adj, nblocks=1000)
    log(f"oracle B edge blocks in {time.time() - t:.1f}s")
    yield b
    b.close()


def _first_diff(a, b):
    d = np.flatnonzero(a != b)
    return f"{len(d)} words differ, first at {d[0] if len(d) else -1}"


def test_c2_rmat24_s4096_every_row_vs_oracle_b(orc, rmat24, blocks24):
    g, G = rmat24
    S = 4096
    src = vs.sample_sources(g.V, S, bench.SOURCE_SEED)
    t = time.time()
    dv, ka, vc, _ = blocks24.vexpand(src, 1, 4, "shortest")
    counts = orc.dvsm_row_counts(dv, S, g.V)
    log(f"oracle B, 4096 rows, hops 1..4: {time.time() - t:.1f}s, k_actual {ka}, "
        f"{int(counts.sum())} pairs")
    assert ka == 4 and counts.sum() > 0
    variants = [dict(), dict(keep_order=True), dict(tile_words=16, sched="pull"),
                dict(tile_words=32, sched="push")]
    digests = te = None
    for kw in variants:
        r, st = vs.vexpand(G, src, 1, 4, "shortest", **kw)
        assert r.k_actual_max == st["k_actual_max"] == ka, kw
        assert np.array_equal(r.row_counts(0), counts), kw
        assert r.tuple_count(0) == int(counts.sum()), kw
        d = r.row_digests(0)
        if digests is None:
            digests, te = d, st["te"]
        assert np.array_equal(d, digests) and st["te"] == te, kw
        got = r.export(0, _lib.LAYOUT_DVSM)
        r.close()
        assert np.array_equal(got, dv), (kw, _first_diff(got, dv))
        del got
        log(f"vexpand {kw or 'default'}: tile width {st['tile_words']}, modes "
            f"{st['mode_hop']}, exact")
    for sched in ("heuristic", "pull"):
        c, dg, bst = vs.vexpand_batch(G, [vs.BatchQuery(src, 1, 4)], sched=sched)
        assert np.array_equal(c, counts) and np.array_equal(dg, digests), sched
        assert bst["te"] == te, sched
    log("batched engine: counts, digests and TE exact")
    # the batched engine's result matrix (expanded in place, rows packed)
    (r,), bst = vs.vexpand_batch_results(G, [vs.BatchQuery(src, 1, 4)])
    assert r.k_actual_max == ka and bst["te"] == te
    assert np.array_equal(r.row_counts(0), counts)
    assert np.array_equal(r.row_digests(0), digests)
    got = r.export(0, _lib.LAYOUT_DVSM)
    r.close()
    assert np.array_equal(got, dv), _first_diff(got, dv)
    del got
    log("batched engine: result matrix exact")


def test_c2_rmat24_any_kmin2_vs_oracle_b(orc, rmat24, blocks24):
    g, G = rmat24
    src = vs.sample_sources(g.V, 1024, bench.SOURCE_SEED + 1)
    dv, ka, _, _ = blocks24.vexpand(src, 2, 3, "any")
    r, st = vs.vexpand(G, src, 2, 3, "any")
    got = r.export(0, _lib.LAYOUT_DVSM)
    assert np.array_equal(r.row_counts(0), orc.dvsm_row_counts(dv, 1024, g.V))
    r.close()
    assert ka == st["k_actual_max"] == 3
    assert np.array_equal(got, dv), _first_diff(got, dv)


@pytest.mark.parametrize("stage_mb", [None, "0"])
def test_c3_all_4m_triples_vs_sorted_merge(orc, rmat24, monkeypatch, stage_mb):
    g, G = rmat24
    if stage_mb is not None:  # the two-pass path (count, scan, intersect again)
        monkeypatch.setenv("VS_MI_STAGE_MAX_MB", stage_mb)
    T = 4 << 20
    tup = vs.sample_paths3(g.offsets, g.adj, T, seed=bench.BATCH_SEED)
    c, o, lists, _ = vs.mintersect([G, G, G], tup)
    t = time.time()
    c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj)] * 3, tup)
    log(f"sorted-merge oracle over {T} triples: {time.time() - t:.1f}s, "
        f"{len(l2)} list entries")
    assert np.array_equal(c, c2)
    assert np.array_equal(o, o2)
    assert np.array_equal(lists, l2)


def test_c4_batch256_vs_plain_and_oracle_b(orc, rmat24, blocks24):
    g, G = rmat24
    qs = bench.batch_queries(vs, np, g.V, 256, bench.BATCH_SEED)
    t = time.time()
    bc, bd, bst = vs.vexpand_batch(G, qs, sched="heuristic")
    log(f"batched: {len(qs)} queries, {len(bc)} rows, {bst['num_tiles']} tiles, "
        f"{bst['num_waves']} waves, {time.time() - t:.1f}s")
    te, row = 0, 0
    rng = np.random.default_rng(bench.BATCH_SEED)
    small = [i for i, q in enumerate(qs) if len(q.sources) <= 1024]
    checked = set(rng.choice(small, size=3, replace=False).tolist())
    checked.add(max(small, key=lambda i: (qs[i].k_max, len(qs[i].sources))))
    t = time.time()
    for i, q in enumerate(qs):
        r, st = vs.vexpand(G, q.sources, q.k_min, q.k_max, q.path)
        n = len(q.sources)
        assert np.array_equal(r.row_counts(0), bc[row:row + n]), i
        assert np.array_equal(r.row_digests(0), bd[row:row + n]), i
        te += st["te"]
        if i in checked:
            dv, ka, _, _ = blocks24.vexpand(q.sources, q.k_min, q.k_max, q.path)
            got = r.export(0, _lib.LAYOUT_DVSM)
            assert ka == r.k_actual_max, i
            assert np.array_equal(got, dv), (i, _first_diff(got, dv))
            assert np.array_equal(orc.dvsm_row_counts(dv, n, g.V), bc[row:row + n]), i
            log(f"query {i} (S={n}, hops {q.k_min}..{q.k_max}) equals oracle B")
        r.close()
        row += n
        if i % 64 == 63:
            log(f"plain {i + 1}/{len(qs)} queries compared, {time.time() - t:.1f}s")
    assert te == bst["te"]


@pytest.fixture(scope="module")
def rmat19():
    g = vs.HostCSR.rmat(19, 32, 0.57, 0.19, 0.19, 5)
    return g, vs.Graph.from_csr(g)


@pytest.mark.parametrize("path", ["shortest", "any"])
def test_full_tile_width64_rmat19(orc, rmat19, path):
    g, G = rmat19
    S = 4096  # exactly one tile of 64 words
    src = vs.sample_sources(g.V, S, seed=19).astype(np.uint32)
    src[::97] = src[1]  # duplicate rows
    blocks = orc.EdgeBlocks(g.offsets, g.adj, nblocks=200)
    dv, ka, _, _ = blocks.vexpand(src, 1, 4, path)
    blocks.close()
    counts = orc.dvsm_row_counts(dv, S, g.V)
    digests = None
    for sched in ("heuristic", "pull", "push"):
        for keep in (False, True):
            r, st = vs.vexpand(G, src, 1, 4, path, sched=sched, tile_words=64,
                               keep_order=keep)
            assert st["tile_words"] == 64 and st["num_tiles"] == 1
            got = r.export(0, _lib.LAYOUT_DVSM)
            assert np.array_equal(got, dv), (sched, keep, _first_diff(got, dv))
            assert np.array_equal(r.row_counts(0), counts), (sched, keep)
            d = r.row_digests(0)
            digests = d if digests is None else digests
            assert np.array_equal(d, digests), (sched, keep)
            r.close()
        c, dg, _ = vs.vexpand_batch(G, [vs.BatchQuery(src, 1, 4, path)], sched=sched,
                                    tile_words=64)
        assert np.array_equal(c, counts) and np.array_equal(dg, digests), sched
