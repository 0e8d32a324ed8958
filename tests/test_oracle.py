"""Pins the CPU oracle before anything is checked against it.

* golden vectors computed by scipy / networkx (tests/golden/make_golden.py);
* closed forms on hand-made graphs;
* the reference's Mintersect known-answer tests (mintersect.cpp:347-512);
* oracle A (per-row semantic BFS) == oracle B (the reference's stacked,
  dst-partitioned Hilbert-block algorithm) bit for bit on R-MAT graphs;
* a planted mistake in a copy of the inputs is caught.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden


def _rows_to_bits(m, V):
    return np.unpackbits(m.view(np.uint8).reshape(m.shape[0], -1), axis=1,
                         bitorder="little")[:, :V]


@pytest.mark.parametrize("name", golden_names())
@pytest.mark.parametrize("k_min,k_max", [(0, 3), (1, 3), (2, 5), (1, 1), (3, 2)])
def test_shortest_matches_scipy_bfs(orc, name, k_min, k_max):
    g = load_golden(name)
    V = int(g["V"][0])
    mats, ka, te, pop = orc.vexpand_rows(g["offsets"], g["adj"], g["sources"], k_min, k_max,
                                         "shortest")
    dist = g["dist"]
    # k_actual: first hop with no new vertex for any source, else k_max
    ka_exp = k_max
    for i in range(1, k_max + 1):
        if not (dist == i).any():
            ka_exp = i
            break
    assert ka == ka_exp
    exp = (dist >= k_min) & (dist <= ka_exp) & (dist >= 0)
    assert np.array_equal(_rows_to_bits(mats[0], V), exp.astype(np.uint8))


@pytest.mark.parametrize("name", golden_names())
def test_shortest_layers_match_scipy(orc, name):
    g = load_golden(name)
    V = int(g["V"][0])
    mats, ka, te, pop = orc.vexpand_rows(g["offsets"], g["adj"], g["sources"], 0, 4, "shortest",
                                         with_distance=True)
    for i in range(len(mats)):
        assert np.array_equal(_rows_to_bits(mats[i], V), (g["dist"] == i).astype(np.uint8))


@pytest.mark.parametrize("name", golden_names())
@pytest.mark.parametrize("k_min,k_max", [(0, 2), (1, 4), (3, 5)])
def test_any_matches_boolean_powers(orc, name, k_min, k_max):
    g = load_golden(name)
    V = int(g["V"][0])
    layers = np.unpackbits(g["any_layers"], axis=2, bitorder="little")[:, :, :V]
    mats, ka, te, pop = orc.vexpand_rows(g["offsets"], g["adj"], g["sources"], k_min, k_max,
                                         "any")
    assert ka == k_max
    exp = np.zeros_like(layers[0])
    for i in range(k_min, k_max + 1):
        exp |= layers[i]
    assert np.array_equal(_rows_to_bits(mats[0], V), exp)


def test_closed_forms(orc):
    # path 0-1-...-69: Shortest layer i of source s = {s-i, s+i}
    g = load_golden("path70")
    mats, ka, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [10, 35], 0, 3, "shortest",
                                      with_distance=True)
    bits = [_rows_to_bits(m, 70) for m in mats]
    for i in range(4):
        assert set(np.nonzero(bits[i][0])[0]) == ({10 - i, 10 + i})
        assert set(np.nonzero(bits[i][1])[0]) == ({35 - i, 35 + i})
    # star: hop 1 from the centre = all leaves; Any hop 2 from a leaf = all
    # leaves (incl. itself); Shortest hop 2 from a leaf = the other leaves
    g = load_golden("star33")
    m, _, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [0, 5], 1, 1, "shortest")
    b = _rows_to_bits(m[0], 33)
    assert b[0].sum() == 32 and b[0][0] == 0 and set(np.nonzero(b[1])[0]) == {0}
    m, _, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [5], 2, 2, "any")
    assert set(np.nonzero(_rows_to_bits(m[0], 33)[0])[0]) == set(range(1, 33))
    m, _, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [5], 2, 2, "shortest")
    assert set(np.nonzero(_rows_to_bits(m[0], 33)[0])[0]) == set(range(1, 33)) - {5}
    # clique K9: Any layer i >= 1 is every vertex (n >= 3)
    g = load_golden("clique9")
    m, _, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [4], 2, 2, "any")
    assert _rows_to_bits(m[0], 9).sum() == 9
    # Shortest stops after the first empty hop; k_max beyond it changes nothing
    m1, ka1, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [4], 0, 2, "shortest")
    m2, ka2, _, _ = orc.vexpand_rows(g["offsets"], g["adj"], [4], 0, 9, "shortest")
    assert ka1 == 2 and ka2 == 2 and np.array_equal(m1, m2)
    # isolated source: only itself with k_min = 0, nothing otherwise
    g = load_golden("two_comp")
    m, ka, te, _ = orc.vexpand_rows(g["offsets"], g["adj"], [120], 0, 3, "shortest")
    assert set(np.nonzero(_rows_to_bits(m[0], 130)[0])[0]) == {120} and ka == 1 and te == 0


def test_te_counts_degrees(orc):
    g = load_golden("star33")
    # source = centre, hops 1..2 Shortest: TE_1 = deg(0) = 32, TE_2 = 32 leaves * 1
    _, ka, te, _ = orc.vexpand_rows(g["offsets"], g["adj"], [0], 1, 2, "shortest")
    assert ka == 2 and te == 32 + 32


def _mintersect_kat_deps():
    # Inline data of Operator.MintersectFlatFlatFlat / MintersectFlatMatMat
    # (mintersect.cpp:347-512): base (a, b); dependency (a -> c) and (b -> c)
    # with the same pairs a/b = 1..8 -> c = 1,1,1,2,2,3,3,4.
    keys = [1, 2, 3, 4, 5, 6, 7, 8]
    vals = [1, 1, 1, 2, 2, 3, 3, 4]
    V = 10
    off = np.zeros(V + 1, dtype=np.uint64)
    for k in keys:
        off[k + 1] += 1
    off = np.cumsum(off, dtype=np.uint64)
    adj = np.array(vals, dtype=np.uint32)  # keys are ascending, one value each
    base = np.array([[1, 1], [2, 1], [3, 1], [4, 2], [5, 2], [6, 3], [7, 3], [8, 4]],
                    dtype=np.uint32)
    return (off, adj), base


def test_mintersect_reference_kat(orc):
    dep, base = _mintersect_kat_deps()
    for fn in (lambda d, t: orc.mintersect_sorted(d, t),
               lambda d, t: orc.mintersect_bitmap(d, t, 10)):
        counts, offsets, lists = fn([dep, dep], base)
        # the reference asserts 3 result tuples: (1,1,1), (2,1,1), (3,1,1)
        assert int(counts.sum()) == 3
        assert counts.tolist() == [1, 1, 1, 0, 0, 0, 0, 0]
        assert lists.tolist() == [1, 1, 1]


def test_mintersect_sorted_equals_bitmap(orc):
    import vertexsurge_amd as vs
    g = vs.HostCSR.rmat(10, 8, seed=11)
    tup = vs.sample_paths3(g.offsets, g.adj, 3000, seed=5)
    dep = (g.offsets, g.adj)
    c1, o1, l1 = orc.mintersect_sorted([dep] * 3, tup)
    c2, o2, l2 = orc.mintersect_bitmap([dep] * 3, tup, g.V)
    assert np.array_equal(c1, c2) and np.array_equal(l1, l2)
    # brute force on a few tuples
    for t in range(0, 3000, 97):
        sets = [set(g.adj[g.offsets[x]:g.offsets[x + 1]].tolist()) for x in tup[t]]
        exp = sorted(sets[0] & sets[1] & sets[2])
        assert l1[o1[t]:o1[t + 1]].tolist() == exp
    assert c1.sum() > 0


@pytest.mark.parametrize("path", ["shortest", "any"])
@pytest.mark.parametrize("k_min", [0, 1, 2])
def test_oracle_a_equals_oracle_b(orc, path, k_min):
    import vertexsurge_amd as vs
    g = vs.HostCSR.rmat(13, 8, seed=3)
    src = vs.sample_sources(g.V, 1100, seed=9)  # two 1024-row stacks, ragged
    mats, ka, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, k_min, 3, path)
    blocks = orc.EdgeBlocks(g.offsets, g.adj, nblocks=100)
    dv, ka2, _, _ = blocks.vexpand(src, k_min, 3, path)
    assert ka == ka2
    assert np.array_equal(dv, orc.rowmajor_to_dvsm(mats[0], g.V))


def test_oracle_b_catches_planted_mistake(orc):
    import vertexsurge_amd as vs
    g = vs.HostCSR.rmat(11, 8, seed=3)
    src = vs.sample_sources(g.V, 300, seed=1)
    mats, _, _, _ = orc.vexpand_rows(g.offsets, g.adj, src, 1, 2, "shortest")
    # drop one edge from the copy the CPU baseline sees
    off = g.offsets.copy()
    adj = g.adj.copy()
    u = int(src[0])
    k = int(off[u])
    adj = np.delete(adj, k)
    off[u + 1:] -= 1
    blocks = orc.EdgeBlocks(off, adj, nblocks=20)
    dv, _, _, _ = blocks.vexpand(src, 1, 2, "shortest")
    assert not np.array_equal(dv, orc.rowmajor_to_dvsm(mats[0], g.V))


def test_bytes_alg_formula_small(orc):
    # path 0-1-2-3 from source 0, Shortest hops 1..2, one tile:
    # L0={0} L1={1} L2={2}; nz: 1 word each; act(F1)={0} deg 1; act(F2)={1} deg 2
    off = np.array([0, 1, 3, 5, 6], dtype=np.uint64)
    adj = np.array([1, 0, 2, 1, 3, 2], dtype=np.uint32)
    lay, ka, te, _ = orc.vexpand_rows(off, adj, [0], 0, 2, "shortest", with_distance=True)
    R, _, _, _ = orc.vexpand_rows(off, adj, [0], 1, 2, "shortest")
    b = orc.bytes_alg(off, 1, lay, ka, 1, "shortest", R[0], 64)
    # hop1: 8+8 + vis 8*(1+2) + 4*1+16 ; hop2: 8+8 + 8*(2+3) + 4*2+16 ;
    # final R = {1, 2}: 8*nz(R) = 16
    assert b == (16 + 24 + 20) + (16 + 40 + 24) + 16
    assert te == 1 + 2


@pytest.mark.parametrize("name", golden_names())
def test_bfs_dist_matches_scipy(orc, name):
    # orc_bfs_dist is the checker of FindShortest; the scipy distances
    # (shortest_path, unweighted, directed: tests/golden/make_golden.py) pin it
    g = load_golden(name)
    for r, s in enumerate(g["sources"].tolist()):
        d = orc.bfs_dist(g["offsets"], g["adj"], s)
        assert np.array_equal(d, g["dist"][r]), (name, r, s)


def test_bfs_dist_closed_forms(orc):
    # path 0-1-...-69: dist(s, v) = |s - v|; two components: the other side is -1
    g = load_golden("path70")
    for s in (0, 10, 69):
        assert orc.bfs_dist(g["offsets"], g["adj"], s).tolist() == [abs(s - v) for v in range(70)]
    g = load_golden("star33")
    d = orc.bfs_dist(g["offsets"], g["adj"], 7)
    assert d[7] == 0 and d[0] == 1 and all(d[v] == 2 for v in range(1, 33) if v != 7)


def _dvsm_hand_case():
    # S = 1100 rows (two 1024-row stacks, the second ragged), V = 3 columns.
    # Every row holds one bit, at its source column: column 0 except rows
    # 1023 (column 2), 1024 (column 1) and 1099 (column 2).
    src = np.zeros(1100, dtype=np.uint32)
    src[1023], src[1024], src[1099] = 2, 1, 2
    # DenseVarStackMatrix (bitmap.hh:348-356, bitmap.cpp:910-920): K = 1024,
    # K / 64 = 16 words per column per stack, v_per_stack = align8(16 * 3) = 48,
    # 2 stacks -> align8(96) = 96 words.  Worked by hand:
    #   rows 0..1022, column 0  -> words 0..15 (row r in word r / 64, bit r % 64),
    #                              word 15 lacks bit 63 (row 1023)
    #   row 1023, column 2      -> word 0 + 2 * 16 + 15 = 47, bit 63
    #   rows 1025..1098, col 0  -> stack 1 base 48: word 48 (rows 1024..1087,
    #                              without bit 0 = row 1024), word 49 (rows
    #                              1088..1098 = bits 0..10)
    #   row 1024, column 1      -> 48 + 16 + 0 = 64, bit 0
    #   row 1099, column 2      -> 48 + 32 + 1 = 81, bit 1099 % 64 = 11
    exp = np.zeros(96, dtype=np.uint64)
    exp[0:15] = np.uint64(0xFFFFFFFFFFFFFFFF)
    exp[15] = np.uint64(0x7FFFFFFFFFFFFFFF)
    exp[47] = np.uint64(1 << 63)
    exp[48] = np.uint64(0xFFFFFFFFFFFFFFFE)
    exp[49] = np.uint64(0x7FF)
    exp[64] = np.uint64(1)
    exp[81] = np.uint64(1 << 11)
    return src, exp


def test_dvsm_layout_hand_worked(orc):
    src, exp = _dvsm_hand_case()
    assert int(orc.lib().orc_dvsm_words(1100, 3)) == 96
    rm = np.zeros((1100, 1), dtype=np.uint64)
    rm[np.arange(1100), 0] = np.left_shift(np.uint64(1), src.astype(np.uint64))
    assert np.array_equal(orc.rowmajor_to_dvsm(rm, 3), exp)
    # oracle B's own indexing: an edgeless graph, Shortest with k_min = 0 is
    # the visited matrix = L0 = the source bits
    off = np.zeros(4, dtype=np.uint64)
    adj = np.zeros(0, dtype=np.uint32)
    blocks = orc.EdgeBlocks(off, adj, nblocks=1)
    dv, ka, vc, _ = blocks.vexpand(src, 0, 2, "shortest")
    blocks.close()
    assert ka == 1 and vc == 1100
    assert np.array_equal(dv, exp)


def test_dvsm_row_counts(orc):
    # hand case: every row of the 1100-row matrix holds exactly one bit
    src, exp = _dvsm_hand_case()
    assert orc.dvsm_row_counts(exp, 1100, 3).tolist() == [1] * 1100
    # against the row-major popcounts of random matrices (ragged stacks and
    # columns), through the independently pinned rowmajor_to_dvsm
    rng = np.random.default_rng(4)
    for S, V in ((1, 1), (70, 130), (1100, 77), (2048, 64), (2050, 200)):
        rm = rng.integers(0, 2**63, size=(S, (V + 63) // 64), dtype=np.uint64)
        if V % 64:
            rm[:, -1] &= np.uint64((1 << (V % 64)) - 1)
        rm[S // 2] = 0
        dv = orc.rowmajor_to_dvsm(rm, V)
        assert np.array_equal(orc.dvsm_row_counts(dv, S, V), orc.row_counts(rm)), (S, V)
