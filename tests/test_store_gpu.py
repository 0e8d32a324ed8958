"""Graphs opened from the reference's on-disk store run the same HIP path as
graphs built in memory: VExpand and Mintersect over a loaded store are
bit-exact against the oracle on the store's (normalised) CSR, and a store
written by the library loads back to identical results."""
import numpy as np
import pytest

import vertexsurge_amd as vs
from conftest import GOLDEN
from oracle import store as ost

pytestmark = pytest.mark.gpu

STORE = GOLDEN / "store"


def rowmajor_counts(res):
    return res.row_counts().tolist()


def test_knows_fixture_hand_values():
    g = vs.Graph.load(STORE / "knows")
    src = np.arange(5, dtype=np.uint32)
    # sets 0:{1,2} 1:{0,2} 2:{0,1} 3:{3} 4:{} (repeats and the doubled
    # self-loop of the stored lists dropped)
    r, st = vs.vexpand(g, src, 1, 3, "shortest")
    assert rowmajor_counts(r) == [2, 2, 2, 0, 0] and r.k_actual_max == 2
    r, _ = vs.vexpand(g, src, 1, 2, "any")
    assert rowmajor_counts(r) == [3, 3, 3, 1, 0]
    g.close()


def test_fin_fixture_directed():
    g = vs.Graph.load(STORE / "fin", "transfer")
    # 0 -> 1 -> 2 -> 0 once the three stored (0,1) multi-edges collapse
    src = np.arange(3, dtype=np.uint32)
    r, _ = vs.vexpand(g, src, 1, 1, "shortest", direction="forward")
    assert [r.get(i, (i + 1) % 3) for i in range(3)] == [True] * 3
    assert rowmajor_counts(r) == [1, 1, 1]
    r, _ = vs.vexpand(g, src, 1, 1, "shortest", direction="backward")
    assert [r.get(i, (i + 2) % 3) for i in range(3)] == [True] * 3
    r, _ = vs.vexpand(g, src, 1, 2, "shortest", direction="forward")
    assert rowmajor_counts(r) == [2, 2, 2]
    g.close()


def test_importer_store_parity_with_oracle(orc, tmp_path):
    """An import-style store (multi-edges, repeats the importer leaves in)
    gives the oracle's results on the store's sets."""
    rng = np.random.default_rng(11)
    V = 3000
    e = rng.integers(0, V, size=(20000, 2))
    e = np.concatenate([e, e[:3000], e[:1000]])
    fi, fd, _, _, n = ost.importer_csr(V, e.tolist(), directed=False, allow_multiple=False)
    d = tmp_path / "imp"
    d.mkdir()
    names = ost.store_names("Person", "knows", "Person", False)
    ost.write_diskarray(d / names["fwd_index"], fi)
    ost.write_diskarray(d / names["fwd_data"], fd)
    import json
    (d / "schema.json").write_text(json.dumps({
        "edge": [{"allow_multiple_edge": False, "directed": False, "dst": "Person",
                  "properties": {}, "src": "Person", "statistics": {"count": n},
                  "store_info": {"flats": [], "partition_counts": [], "partitioned_flats": []},
                  "type": "knows"}],
        "graph_name": "", "graph_type": "",
        "vertex": [{"properties": {}, "statistics": {"count": V}, "type": "Person"}]}))
    _, off, adj, info = ost.read_store(d)
    assert info["duplicates"] > 0
    g = vs.Graph.load(d)
    assert (g.V, g.E) == (V, len(adj))
    src = vs.sample_sources(V, 130, seed=5)
    for path, kmin, kmax in (("shortest", 1, 4), ("any", 0, 2)):
        res, st = vs.vexpand(g, src, kmin, kmax, path)
        mats, ka, te, _ = orc.vexpand_rows(off, adj, src, kmin, kmax, path)
        assert res.k_actual_max == ka and st["te"] == te
        assert np.array_equal(res.export(0, vs.LAYOUT_ROWMAJOR), mats[0])
    # intersections over the loaded relation
    tup = rng.integers(0, V, size=(400, 3)).astype(np.uint32)
    counts, offsets, lists, _ = vs.mintersect([g, g, g], tup)
    ec, eo, el = orc.mintersect_sorted([(off, adj)] * 3, tup)
    assert np.array_equal(counts, ec) and np.array_equal(offsets, eo)
    assert np.array_equal(lists, el)
    g.close()


@pytest.mark.parametrize("directed", [False, True])
def test_written_store_round_trip(tmp_path, directed):
    csr = vs.HostCSR.rmat(14, 16, seed=3)
    off, adj = csr.offsets.copy(), csr.adj.copy()
    d = tmp_path / "rt"
    d.mkdir()
    vs.store_write(d, off, adj, "Person", "knows", directed=directed)
    g0 = vs.Graph(off, adj, symmetric=not directed)
    g1 = vs.Graph.load(d)
    src = vs.sample_sources(csr.V, 500, seed=9)
    for direction in ("forward", "both") if directed else ("forward",):
        r0, s0 = vs.vexpand(g0, src, 1, 4, "shortest", direction=direction)
        r1, s1 = vs.vexpand(g1, src, 1, 4, "shortest", direction=direction)
        assert np.array_equal(r0.row_digests(), r1.row_digests())
        assert np.array_equal(r0.row_counts(), r1.row_counts())
        assert s0["te"] == s1["te"] and r0.k_actual_max == r1.k_actual_max
    g0.close()
    g1.close()
