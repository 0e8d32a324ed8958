"""Learned push/pull scheduler: trained on the GPU from the plain method's own
traces, judged against the density heuristic on held-out graphs.  Scheduling
never changes a result bit."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from vertexsurge_amd import sched_train

pytestmark = pytest.mark.gpu

TRAIN = [{"kind": "rmat", "scale": 15, "ef": 16, "a": a, "seed": 10 + i}
         for i, a in enumerate((0.45, 0.52, 0.6))] + \
        [{"kind": "rmat", "scale": 16, "ef": 16, "a": 0.57, "seed": 3}]
HELD = [{"kind": "rmat", "scale": 17, "ef": 16, "a": 0.57, "seed": 99},
        {"kind": "er", "V": 1 << 17, "pairs": 16 << 17, "seed": 99}]


@pytest.fixture(scope="module")
def trained():
    X, Y, M = sched_train.collect(TRAIN, source_sizes=(64, 512, 2048), k_max=5, log=print,
                                  tw_sizes=(2048, 4096, 8192), tw_hops=(3, 5))
    assert len(X) >= 20 and (~sched_train.hop_rows(X)).sum() >= 8
    s = sched_train.train(X, Y, M, epochs=1500, lr=3e-3)
    assert s.curve[-1] < s.curve[0]
    return s, X, Y, M


def test_training_fits_the_traces(trained):
    s, X, Y, M = trained
    acc = sched_train.accuracy(s, X, Y)
    # on its own traces the MLP must pick the faster option most of the time
    assert acc["agree"] >= 0.7, acc
    assert acc["ms_mlp"] <= min(acc["ms_push"], acc["ms_pull"]) * 1.05, acc
    tw = sched_train.tw_accuracy(s, X, Y, M)
    assert tw["ms_mlp"] <= tw["ms_fallback"] * 1.0001, tw


def test_learned_gains_on_held_out_graphs(trained):
    """The learned decisions, judged on measured traces of graphs the MLP never
    saw: for each decision the same measured table prices the learned choice
    and the method without it (density heuristic for push/pull, the widest
    tile for the tile width), so no timing noise separates them.  The learned
    scheduler must not lose on any held-out graph."""
    s = trained[0]
    rows = sched_train.heldout_traces(s, HELD, source_sizes=(128, 1024, 4096), k_max=5,
                                      tw_sizes=(2048, 4096, 8192), tw_hops=(3, 5), log=print)
    for r in rows:
        gain = 1 - r["ms_learned"] / r["ms_heuristic"]
        tw_gain = 1 - r["tw_ms_learned"] / r["tw_ms_default"]
        print(f"{r['graph']}: push/pull gain {100 * gain:.2f}% ({r['changed']} of {r['hops']} "
              f"decisions changed), tile width gain {100 * tw_gain:.2f}% ({r['tw_changed']} of "
              f"{r['queries']} changed)")
        assert r["ms_learned"] <= r["ms_heuristic"] * 1.0001, r
        assert r["tw_ms_learned"] <= r["tw_ms_default"] * 1.0001, r


def test_learned_is_bit_exact(trained):
    s = trained[0]
    for spec in HELD:
        csr = sched_train.make_graph(spec)
        G = vs.Graph.from_csr(csr)
        G.set_scheduler(s)
        src = vs.sample_sources(csr.V, 4096, 7)
        r1, s1 = vs.vexpand(G, src, 1, 4, sched="learned")
        r2, s2 = vs.vexpand(G, src, 1, 4, sched="heuristic")
        assert np.array_equal(r1.export(0, vs.LAYOUT_DVSM), r2.export(0, vs.LAYOUT_DVSM))
        assert s1["te"] == s2["te"]
        print(f"{spec}: learned tile width {s1['tile_words']} (heuristic {s2['tile_words']})")
        b1, _ = vs.vexpand_batch_results(G, [vs.BatchQuery(src, 1, 4)], sched="learned")
        assert np.array_equal(b1[0].export(0, vs.LAYOUT_DVSM), r2.export(0, vs.LAYOUT_DVSM))
        for x in (r1, r2, b1[0]):
            x.close()
        G.set_scheduler(None)


def test_weights_roundtrip(tmp_path, trained):
    s, X, _, _ = trained
    p = tmp_path / "w.npz"
    sched_train.save(s, p, {"note": "test"})
    s2 = sched_train.load(p)
    assert np.array_equal(s.predict(X[:16]), s2.predict(X[:16]))


def test_domain_guard_returns_decisions_to_the_heuristic():
    """An MLP that always predicts push 2^10 times faster overrides every
    decision inside its domain, and none outside it -- on the host path and
    inside the replayed batched graph -- without changing a bit."""
    g = vs.HostCSR.rmat(15, 16, seed=12)
    G = vs.Graph.from_csr(g)
    s = vs.Scheduler(seed=3)
    p = np.zeros(s.num_params, dtype=np.float32)
    p[-8] = -5.0   # b3[push]: log2 ms predicted for push
    p[-7] = 5.0    # b3[pull]
    s.set_params(p)
    G.set_scheduler(s)
    src = vs.sample_sources(g.V, 1024, seed=4)
    ref, rst = vs.vexpand(G, src, 1, 4, sched="heuristic")
    assert 2 in rst["mode_hop"]  # the heuristic pulls somewhere
    want = ref.export(0, vs.LAYOUT_DVSM)
    for rows, radius, forced in ((np.zeros((0, 8)), 0.0, True),       # no domain
                                 (np.full((1, 8), 5.0), 0.01, False)):  # far from every decision
        s.set_domain(rows, radius)
        r, st = vs.vexpand(G, src, 1, 4, sched="learned")
        assert np.array_equal(r.export(0, vs.LAYOUT_DVSM), want)
        assert st["mode_hop"] == ([1] * len(st["mode_hop"]) if forced else rst["mode_hop"])
        (b,), bst = vs.vexpand_batch_results(G, [vs.BatchQuery(src, 1, 4)], sched="learned")
        assert np.array_equal(b.export(0, vs.LAYOUT_DVSM), want)
        assert (bst["pull_decisions"] == 0) == forced
        r.close()
        b.close()
    ref.close()
    G.set_scheduler(None)
