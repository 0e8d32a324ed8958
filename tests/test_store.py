"""Graph store in the reference's on-disk format (host side, no GPU).

The hand-derived fixture under tests/golden/store/ pins both the oracle
(oracle/store.py) and the library's reader (vs_store_open); round trips
through the library's writer are read back by the oracle.
"""
import json
import os

import numpy as np
import pytest

import vertexsurge_amd as vs
from conftest import GOLDEN
from oracle import store as ost

STORE = GOLDEN / "store"

KNOWS_RAW = [(0, 1), (0, 1), (1, 2), (2, 0), (0, 1), (3, 3)]
FIN_RAW = [(0, 1), (0, 1), (2, 0), (1, 2), (0, 1)]


def lib_open(d, edge=None):
    with vs.GraphStore(d, edge) as st:
        o, a = st.csr()
        return o, a, dict(st.info)


def test_fixture_files_follow_diskarray_layout():
    for d in ("knows", "fin"):
        for name in os.listdir(STORE / d):
            if name == "schema.json":
                continue
            raw = (STORE / d / name).read_bytes()
            assert len(raw) == 4096, name
            count, size = np.frombuffer(raw[:16], dtype="<i8")
            assert size in (4, 8) and 64 + count * size <= 4096
            assert raw[16:64] == b"\0" * 48


def test_importer_restatement_matches_hand_derived_fixture():
    fi, fd, bi, bd, n = ost.importer_csr(5, KNOWS_RAW, directed=False, allow_multiple=False)
    assert n == 10 and bi is None
    _, _, idx = ost.read_diskarray(STORE / "knows" / "Person-knows-Person_csr_src_dst_index",
                                   np.int64)
    _, _, dat = ost.read_diskarray(STORE / "knows" / "Person-knows-Person_csr_src_dst",
                                   np.uint32)
    assert idx.tolist() == [0, 3, 6, 8, 10, 10] == fi.tolist()
    assert dat.tolist() == [1, 1, 2, 0, 0, 2, 0, 1, 3, 3] == fd.tolist()
    fi, fd, bi, bd, n = ost.importer_csr(3, FIN_RAW, directed=True, allow_multiple=True)
    assert n == 5
    assert fi.tolist() == [0, 3, 4, 5] and fd.tolist() == [1, 1, 1, 2, 0]
    assert bi.tolist() == [0, 1, 4, 5] and bd.tolist() == [2, 0, 0, 0, 1]
    _, _, b2 = ost.read_diskarray(STORE / "fin" / "Account<-transfer-Account_csr_dst_src",
                                  np.uint32)
    assert b2.tolist() == bd.tolist()


def test_oracle_reader_hand_values():
    V, o, a, info = ost.read_store(STORE / "knows")
    assert V == 5
    assert o.tolist() == [0, 2, 4, 6, 7, 7] and a.tolist() == [1, 2, 0, 2, 0, 1, 3]
    assert (info["duplicates"], info["self_loops"], info["unsorted_lists"]) == (3, 2, 0)
    V, o, a, info = ost.read_store(STORE / "fin")
    assert o.tolist() == [0, 1, 2, 3] and a.tolist() == [1, 2, 0]
    assert info["duplicates"] == 2 and info["directed"] == 1


@pytest.mark.parametrize("d,edge", [("knows", None), ("knows", "knows"),
                                    ("knows", "Person_knows_Person"), ("fin", None),
                                    ("fin", "transfer"), ("fin", "Account_transfer_Account")])
def test_library_reader_matches_oracle(d, edge):
    o, a, info = lib_open(STORE / d, edge)
    V, eo, ea, einfo = ost.read_store(STORE / d, edge)
    assert np.array_equal(o, eo) and np.array_equal(a, ea)
    for k, v in einfo.items():
        assert info[k] == v, k


def test_knows_info_hand_values():
    _, _, info = lib_open(STORE / "knows")
    assert info == {"num_vertices": 5, "stored_entries": 10, "num_edges": 7, "duplicates": 3,
                    "unsorted_lists": 0, "self_loops": 2, "directed": 0,
                    "allow_multiple_edge": 0}


def test_relation_between_two_types_is_unsupported():
    with pytest.raises(vs.VSError) as e:
        vs.GraphStore(STORE / "fin", "signIn")
    assert e.value.code == vs._lib.VS_ERR_UNSUPPORTED
    with pytest.raises(vs.VSError) as e:
        vs.GraphStore(STORE / "fin", "nosuch")
    assert e.value.code == vs._lib.VS_ERR_INVALID


def test_missing_directory_is_io_error(tmp_path):
    with pytest.raises(vs.VSError) as e:
        vs.GraphStore(tmp_path / "absent")
    assert e.value.code == vs._lib.VS_ERR_IO


def _copy_store(src, dst):
    dst.mkdir()
    for name in os.listdir(src):
        (dst / name).write_bytes((src / name).read_bytes())


def test_corrupt_files_are_rejected(tmp_path):
    base = STORE / "knows"
    idx_name = "Person-knows-Person_csr_src_dst_index"
    dat_name = "Person-knows-Person_csr_src_dst"
    cases = {}
    # element size in the header disagrees with int64 offsets
    raw = bytearray((base / idx_name).read_bytes())
    raw[8:16] = np.array([4], dtype="<i8").tobytes()
    cases["elem"] = (idx_name, bytes(raw))
    # element count disagrees with the schema's statistics.count
    raw = bytearray((base / dat_name).read_bytes())
    raw[0:8] = np.array([9], dtype="<i8").tobytes()
    cases["count"] = (dat_name, bytes(raw))
    # truncated data area
    cases["short"] = (dat_name, (base / dat_name).read_bytes()[:64 + 12])
    # a neighbour id >= V
    raw = bytearray((base / dat_name).read_bytes())
    raw[64:68] = np.array([5], dtype="<u4").tobytes()
    cases["id"] = (dat_name, bytes(raw))
    # last offset != E
    raw = bytearray((base / idx_name).read_bytes())
    raw[64 + 5 * 8:64 + 6 * 8] = np.array([9], dtype="<i8").tobytes()
    cases["off"] = (idx_name, bytes(raw))
    for k, (name, data) in cases.items():
        d = tmp_path / k
        _copy_store(base, d)
        (d / name).write_bytes(data)
        with pytest.raises(vs.VSError) as e:
            vs.GraphStore(d)
        assert e.value.code == vs._lib.VS_ERR_INVALID, k
    d = tmp_path / "json"
    _copy_store(base, d)
    (d / "schema.json").write_text('{"vertex": [], "edge": [')
    with pytest.raises(vs.VSError) as e:
        vs.GraphStore(d)
    assert e.value.code == vs._lib.VS_ERR_INVALID


def test_schema_string_escapes(tmp_path):
    """Type names go through JSON string escapes: a \\u surrogate pair names the
    same type as its UTF-8 text; an unpaired surrogate is a malformed schema."""
    sch = json.loads((STORE / "knows" / "schema.json").read_text())
    data = {n: (STORE / "knows" / n).read_bytes() for n in os.listdir(STORE / "knows")
            if n != "schema.json"}
    vt = "P\U0001F600"  # outside the BMP: written as a surrogate pair by json.dumps
    sch["vertex"][0]["type"] = vt
    sch["edge"][0]["src"] = sch["edge"][0]["dst"] = vt
    d = tmp_path / "esc"
    d.mkdir()
    for n, b in data.items():
        (d / n.replace("Person", vt)).write_bytes(b)
    text = json.dumps(sch)  # ensure_ascii writes the pair as two \\u escapes
    assert "\\ud83d\\ude00" in text
    (d / "schema.json").write_text(text)
    o, a, _ = lib_open(d)
    eo, ea, _ = lib_open(STORE / "knows")
    assert np.array_equal(o, eo) and np.array_equal(a, ea)
    for bad in ("\\ud83d", "\\ud83d\\u0041", "\\ude00"):
        (d / "schema.json").write_text(text.replace("\\ud83d\\ude00", bad, 1))
        with pytest.raises(vs.VSError) as e:
            vs.GraphStore(d)
        assert e.value.code == vs._lib.VS_ERR_INVALID and "surrogate" in str(e.value), bad


def test_unsorted_lists_are_sorted(tmp_path):
    d = tmp_path / "u"
    _copy_store(STORE / "knows", d)
    # list of vertex 0 stored as 2,1,1 and of vertex 2 as 1,0
    ost.write_diskarray(d / "Person-knows-Person_csr_src_dst",
                        np.array([2, 1, 1, 0, 0, 2, 1, 0, 3, 3], dtype=np.uint32))
    o, a, info = lib_open(d)
    assert o.tolist() == [0, 2, 4, 6, 7, 7] and a.tolist() == [1, 2, 0, 2, 0, 1, 3]
    assert info["unsorted_lists"] == 2 and info["duplicates"] == 3
    _, eo, ea, einfo = ost.read_store(d)
    assert np.array_equal(o, eo) and np.array_equal(a, ea) and einfo["unsorted_lists"] == 2


@pytest.mark.parametrize("directed", [False, True])
def test_writer_round_trip(tmp_path, directed):
    rng = np.random.default_rng(7 + directed)
    V = 300
    edges = rng.integers(0, V, size=(2000, 2))
    fi, fd, bi, bd, n = ost.importer_csr(V, edges.tolist(), directed, allow_multiple=False)
    # the library writes ascending sets; give it the normalised CSR
    off = [0]
    lists = []
    for u in range(V):
        s = np.unique(fd[fi[u]:fi[u + 1]])
        lists.append(s)
        off.append(off[-1] + len(s))
    off = np.array(off, dtype=np.uint64)
    adj = np.concatenate(lists).astype(np.uint32)
    d = tmp_path / "w"
    d.mkdir()
    vs.store_write(d, off, adj, "Person", "knows", directed=directed, graph_name="rt")
    sch = json.loads((d / "schema.json").read_text())
    assert sch["edge"][0]["directed"] is directed
    assert sch["vertex"][0]["statistics"]["count"] == V
    names = ost.store_names("Person", "knows", "Person", directed)
    for key in ("fwd_index", "fwd_data") + (("bwd_index", "bwd_data") if directed else ()):
        p = d / names[key]
        cnt, size, _ = ost.read_diskarray(p, np.int64 if "index" in key else np.uint32)
        assert os.path.getsize(p) == 64 + ost.capacity(cnt, size) * size
    V2, o, a, info = ost.read_store(d)
    assert V2 == V and np.array_equal(o, off) and np.array_equal(a, adj)
    assert info["duplicates"] == 0 and info["unsorted_lists"] == 0
    if directed:
        # reverse CSR = the importer's backward CSR, as sets
        _, _, rbi = ost.read_diskarray(d / names["bwd_index"], np.int64)
        _, _, rbd = ost.read_diskarray(d / names["bwd_data"], np.uint32)
        for u in range(V):
            assert np.array_equal(rbd[rbi[u]:rbi[u + 1]], np.unique(bd[bi[u]:bi[u + 1]]))
    o2, a2, _ = lib_open(d)
    assert np.array_equal(o2, off) and np.array_equal(a2, adj)


def test_importer_store_with_repeats_opens_as_sets(tmp_path):
    rng = np.random.default_rng(3)
    V = 500
    edges = rng.integers(0, V, size=(4000, 2))
    edges = np.concatenate([edges, edges[:500]])  # repeats far apart
    for directed in (False, True):
        fi, fd, bi, bd, n = ost.importer_csr(V, edges.tolist(), directed, allow_multiple=directed)
        d = tmp_path / f"imp{int(directed)}"
        d.mkdir()
        names = ost.store_names("Account", "transfer", "Account", directed)
        ost.write_diskarray(d / names["fwd_index"], fi)
        ost.write_diskarray(d / names["fwd_data"], fd)
        if directed:
            ost.write_diskarray(d / names["bwd_index"], bi)
            ost.write_diskarray(d / names["bwd_data"], bd)
        sch = {"edge": [{"allow_multiple_edge": directed, "directed": directed,
                         "dst": "Account", "properties": {}, "src": "Account",
                         "statistics": {"count": n},
                         "store_info": {"flats": [], "partition_counts": [],
                                        "partitioned_flats": []}, "type": "transfer"}],
               "graph_name": "", "graph_type": "", "vertex": [
                   {"properties": {}, "statistics": {"count": V}, "type": "Account"}]}
        (d / "schema.json").write_text(json.dumps(sch))
        o, a, info = lib_open(d)
        _, eo, ea, einfo = ost.read_store(d)
        assert np.array_equal(o, eo) and np.array_equal(a, ea)
        assert info["duplicates"] == einfo["duplicates"] > 0
        assert info["num_edges"] + info["duplicates"] == n


def test_empty_relation(tmp_path):
    d = tmp_path / "e"
    d.mkdir()
    off = np.zeros(11, dtype=np.uint64)
    vs.store_write(d, off, np.zeros(0, dtype=np.uint32), "V", "e")
    o, a, info = lib_open(d)
    assert o.tolist() == [0] * 11 and len(a) == 0 and info["num_edges"] == 0


def test_mutated_stores_fail_cleanly(tmp_path):
    """Seeded random damage to schema.json (byte flips, truncation, inserted
    characters) and to the DiskArray headers never crashes the reader: every
    open either succeeds or raises VSError with INVALID / IO / UNSUPPORTED."""
    base = STORE / "knows"
    text = (base / "schema.json").read_bytes()
    names = [n for n in os.listdir(base) if n != "schema.json"]
    rng = np.random.default_rng(11)
    ok = failed = 0
    for i in range(300):
        d = tmp_path / f"m{i}"
        _copy_store(base, d)
        kind = i % 3
        if kind == 0:  # flip bytes of the schema
            raw = bytearray(text)
            for p in rng.integers(0, len(raw), size=int(rng.integers(1, 4))):
                raw[p] = int(rng.integers(0, 256))
            (d / "schema.json").write_bytes(bytes(raw))
        elif kind == 1:  # truncate it, or insert JSON punctuation
            raw = bytearray(text)
            p = int(rng.integers(0, len(raw)))
            if rng.random() < 0.5:
                raw = raw[:p]
            else:
                raw[p:p] = bytes(rng.choice(list(b'{}[]",:\\-0e'), size=3))
            (d / "schema.json").write_bytes(bytes(raw))
        else:  # damage a DiskArray header (count or element size)
            name = names[int(rng.integers(0, len(names)))]
            raw = bytearray((d / name).read_bytes())
            field = int(rng.integers(0, 2)) * 8
            val = int(rng.choice([-1, 0, 1, 3, 7, 2**31, 2**62]))
            raw[field:field + 8] = np.array([val], dtype="<i8").tobytes()
            (d / name).write_bytes(bytes(raw))
        try:
            with vs.GraphStore(d) as st:
                o, a = st.csr()
                assert o[0] == 0 and o[-1] == len(a)
            ok += 1
        except vs.VSError as e:
            assert e.code in (vs._lib.VS_ERR_INVALID, vs._lib.VS_ERR_IO,
                              vs._lib.VS_ERR_UNSUPPORTED), (i, e)
            failed += 1
    assert failed > 100 and ok + failed == 300


# ------------------------------------------------------------ flat edge files
# The reference's curve, restated as its two tables (peano.hh:20-55): the
# order in which each traversal visits the quadrants q = 2 * xbit + ybit, and
# the traversal used inside the quadrant visited at each position.
_ORDER = {"RA": (0, 1, 3, 2), "RD": (3, 2, 0, 1), "CA": (0, 2, 3, 1), "CD": (3, 1, 0, 2)}
_SUB = {"RA": ("CA", "RA", "RA", "CD"), "RD": ("CD", "RD", "RD", "CA"),
        "CA": ("RA", "CA", "CA", "RD"), "CD": ("RD", "CD", "CD", "RA")}


def _curve_walk(n, x, y):
    key, d, h = 0, "RA", n // 2
    while h:
        q = (2 if x & h else 0) + (1 if y & h else 0)
        pos = _ORDER[d].index(q)
        key += pos * h * h
        d = _SUB[d][pos]
        h //= 2
    return key


def test_hilbert_key_hand_4x4():
    # quadrants in RA order A(0,0) B(0,1) D(1,1) C(1,0); A walked CA, B and D
    # walked RA, C walked CD
    hand = [[0, 3, 4, 5], [1, 2, 7, 6], [14, 13, 8, 9], [15, 12, 11, 10]]
    x, y = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    assert ost.hilbert_key(4, x.ravel(), y.ravel()).reshape(4, 4).tolist() == hand


def test_hilbert_key_equals_curve_tables():
    for n in (1, 2, 8, 64):
        x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        k = ost.hilbert_key(n, x.ravel(), y.ravel())
        assert k.tolist() == [_curve_walk(n, int(a), int(b))
                              for a, b in zip(x.ravel(), y.ravel())]
    rng = np.random.default_rng(5)
    n = 1 << 30
    x, y = rng.integers(0, n, 500), rng.integers(0, n, 500)
    assert ost.hilbert_key(n, x, y).tolist() == [_curve_walk(n, int(a), int(b))
                                                 for a, b in zip(x, y)]


def test_hilbert_full_grid_steps_are_unit():
    # the reference's Hilbert test property: in curve order, consecutive cells
    # of a full grid share a row or a column (and here: are neighbours)
    n = 32
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    o = np.argsort(ost.hilbert_key(n, x.ravel(), y.ravel()))
    px, py = x.ravel()[o], y.ravel()[o]
    assert np.all(np.abs(np.diff(px)) + np.abs(np.diff(py)) == 1)
    assert sorted(ost.hilbert_key(n, x.ravel(), y.ravel()).tolist()) == list(range(n * n))


def test_partition_by_dst_hand():
    dst = [0, 0, 0, 1, 2, 2, 2, 2, 3, 5]
    # 10 edges / 3: first cut at >= 3 and a new dst -> 3; then (7 / 2) + 3 = 6,
    # dst 2 runs to 8 -> 8; the rest -> 10
    assert ost.partition_by_dst(dst, 3) == [0, 3, 8, 10]
    assert ost.partition_by_dst(dst, 1) == [0, 10]
    # one destination cannot be split: later partitions are empty
    assert ost.partition_by_dst([4, 4, 4, 4], 3) == [0, 4, 4, 4]
    assert ost.partition_by_dst([], 2) == [0, 0, 0]


def _pairs(path):
    cnt, size, d = ost.read_diskarray(path, np.uint64)
    assert size == 8 and os.path.getsize(path) == 64 + ost.capacity(cnt, 8) * 8
    return np.stack([d & 0xFFFFFFFF, d >> 32], axis=1).astype(np.uint32)


@pytest.mark.parametrize("directed,parts", [(False, 4), (True, 3), (True, 1), (False, 64)])
def test_writer_flat_files_equal_oracle(tmp_path, directed, parts):
    g = vs.HostCSR.rmat(9, 8, seed=11)
    off, adj = np.asarray(g.offsets), np.asarray(g.adj)
    d = tmp_path / "flat"
    d.mkdir()
    vs.store_write(d, off, adj, "P", "knows", directed=directed, flat_partitions=parts)
    exp = ost.flat_files(g.V, off, adj, directed, parts, "P", "knows", "P")
    assert len(exp) == (2 if directed else 1) * (2 + 2 * parts)
    for name, pairs in exp.items():
        got = _pairs(d / name)
        assert np.array_equal(got, pairs), name
    si = json.loads((d / "schema.json").read_text())["edge"][0]["store_info"]
    assert si == {"flats": [1, 0], "partitioned_flats": [1, 0],
                  "partition_counts": [parts, parts]}
    # every partition family holds each edge once; no destination is split
    fwd = "P-knows->P" if directed else "P-knows-P"
    parts_h = [_pairs(d / f"{fwd}_hilbert_{i}") for i in range(parts)]
    allp = np.concatenate(parts_h)
    assert len(allp) == len(adj)
    assert np.array_equal(allp[np.lexsort((allp[:, 1], allp[:, 0]))], _pairs(d / f"{fwd}_src_dst"))
    owners = {}
    for i, p in enumerate(parts_h):
        for v in np.unique(p[:, 1]).tolist():
            assert owners.setdefault(v, i) == i
    # the store still opens as before
    o2, a2, _ = lib_open(d)
    assert np.array_equal(o2, off) and np.array_equal(a2, adj)


def test_writer_without_flats_lists_none(tmp_path):
    g = vs.HostCSR.rmat(6, 4, seed=2)
    vs.store_write(tmp_path, g.offsets, g.adj)
    si = json.loads((tmp_path / "schema.json").read_text())["edge"][0]["store_info"]
    assert si == {"flats": [], "partitioned_flats": [], "partition_counts": []}
    assert not [p for p in os.listdir(tmp_path) if "hilbert" in p or "src_dst" in p and "csr" not in p]
