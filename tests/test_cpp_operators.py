"""The C++ operator shims (VExpand / Mintersect with the reference's
constructor parameters) driven from a C++ program, checked against the oracle."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import vertexsurge_amd as vs

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent.parent / "vertexsurge_amd" / "_build" / "test_operators"


def test_cpp_operators_match_oracle(orc):
    store = Path(__file__).resolve().parent / "golden" / "store" / "knows"
    out = subprocess.run([str(DRIVER), str(store)], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    r = json.loads(out)
    g = vs.HostCSR.rmat(12, 16, seed=1)
    starts = vs.sample_sources(g.V, 200, seed=7)
    m, ka, te, _ = orc.vexpand_rows(g.offsets, g.adj, starts, 1, 3, "shortest")
    assert r["shortest_k_actual"] == ka
    assert r["shortest_tuple_count"] == int(orc.row_counts(m[0]).sum())
    assert r["shortest_row_counts"] == orc.row_counts(m[0]).tolist()
    assert r["shortest_get_0_first"] == 1
    assert r["csr_flat_tuples"] == r["shortest_tuple_count"]
    first = np.nonzero(np.unpackbits(m[0].view(np.uint8), axis=1, bitorder="little"))
    assert r["csr_flat_first"] == [int(starts[first[0][0]]), int(first[1][0])]
    m, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, starts, 0, 2, "any")
    assert r["any_row_counts"] == orc.row_counts(m[0]).tolist()
    lay, ka, _, _ = orc.vexpand_rows(g.offsets, g.adj, starts, 1, 3, "shortest",
                                     with_distance=True)
    assert r["layer_distances"] == list(range(1, ka + 1))
    assert r["layer_tuple_counts"] == [int(orc.row_counts(x).sum()) for x in lay]
    targets = starts[::-1]
    exp = []
    for s, t in zip(starts, targets):
        d = orc.bfs_dist(g.offsets, g.adj, int(s))[t]
        exp.append(int(d) + 1 if 0 <= d <= 50 else 0)
    assert r["find_shortest_dist_plus1"] == exp
    assert r["empty_tuple_count"] == 0
    assert r["unknown_schema_throws"] == 1
    # k_max 2 over Person-likes-Post throws, k_max 1 does not; Adaptive throws;
    # CSR has no matrix result, Dense no flat result
    assert r["schema_checks_throw"] == [1, 0, 1, 1, 1]
    assert r["mintersect_kat_tuples"] == 3 and r["mintersect_kat_count"] == 3
    assert r["mintersect_kat_flat"] == [1, 1, 1, 2, 1, 1, 3, 1, 1]
    assert r["mintersect_matmat_tuples"] == 3
    assert r["mintersect_matmat_flat"] == [1, 1, 1, 2, 1, 1, 3, 1, 1]
    # the VExpand -> TableConvert -> Mintersect pattern against the oracle
    from oracle import tables as ot
    # triangles through the edges of 40 sources: VarRow-base Mintersect
    few = starts[:40]
    m1, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, few, 1, 1, "any")
    M1 = np.unpackbits(m1[0].view(np.uint8), axis=1, bitorder="little")[:, :g.V].astype(bool)
    es, et, ew = ot.mintersect_varrow(M1, few, None, [0, 1],
                                      adjacency=[(g.offsets, g.adj)] * 2)
    assert len(ew) > 0
    assert r["varrow_triangles_count"] == len(ew)
    assert r["varrow_triangles_flat"] == np.stack([es, et, ew], axis=1).ravel().tolist()
    m, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, starts, 1, 2, "any")
    M = np.unpackbits(m[0].view(np.uint8), axis=1, bitorder="little")[:, :g.V].astype(bool)
    E, rm, cm = ot.table_convert(M, starts, None, True)
    assert (r["convert_rows"], r["convert_cols"]) == E.shape
    base = starts[: len(starts) // 2 * 2].reshape(-1, 2)
    c, _, _ = ot.mintersect_colmain([(E, rm, cm)] * 2, base)
    assert r["pattern_count"] == int(c.sum())
    ids, cnt = ot.column_count(M, None, "desc")
    assert r["colcount_desc_top5"] == [int(x) for p in zip(ids[:5], cnt[:5]) for x in p]
    assert r["row_or_size"] == len(ot.row_or_sum(M, None))
    sel, _, _ = ot.select_cols(M, starts, [starts[0], starts[1], starts[0]])
    assert r["select_tuple_count"] == int(sel.sum())
    # the hand-derived knows store: sets 0:{1,2} 1:{0,2} 2:{0,1} 3:{3} 4:{}
    assert r["store_shortest_row_counts"] == [2, 2, 2, 0, 0]
    assert r["store_any_row_counts"] == [3, 3, 3, 1, 0]
