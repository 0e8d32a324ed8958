"""Oracle (test infrastructure only): CPU restatement of the reference's
on-disk graph store, the data format in front of VExpand / Mintersect.

Only tests/ use this module; the product path never imports it.

* read_diskarray / write_diskarray -- DiskArray files (disk_array.hh:267-354:
  a 64-byte header whose first two int64 are the element count and size, then
  the elements; DiskArray(path, n, true) stretches the file to 64 + capacity
  * size with capacity = (4096 - 64) / size doubled until >= n,
  disk_array.hh:537-604).
* importer_csr -- the CSR an import writes for one relation
  (unify_edge_and_collect_statistics, graph_loader.hh:58-77: reverse edges
  appended for undirected relations, std::unique over the unsorted list
  unless multi-edges are allowed, then sort by (src, dst);
  sort_and_generate_index_from_src/_dst, graph_concepts.hh:114-160).
* store_names -- file names (schema.cpp:11-49).
* read_store -- what opening a relation yields: the source-side CSR with
  each list made an ascending set, plus the counts the library reports.
* hilbert_key -- position on the Peano-Hilbert curve (XYToIndex with
  PeanoDirection::RA, peano.hh:73-89), here as the textbook rotate-and-reflect
  walk over the coordinate bits.
* partition_by_dst -- destination partitions (graph_loader.hh:111-139).
* flat_files -- the flat edge files of one relation (gen_and_write_one_edge,
  graph_loader.hh:186-276, with flats / partitioned_flats {src_dst, hilbert}).

Pinned by: a hand-derived fixture (tests/golden/store/, made by
tests/golden/make_store_fixture.py from edge lists worked through the importer
steps by hand, including the adjacent-only duplicate removal), and by
round trips through the library's own writer; hilbert_key by the 4 x 4 curve
worked by hand from the reference's quadrant orders, by the reference's
quadrant-order / sub-curve tables walked cell by cell up to 64 x 64, and by
the reference's own Hilbert test property (consecutive cells of a full grid
share a row or a column, gtest.cpp:526-573); partition_by_dst by hand-worked
cuts.
"""
from __future__ import annotations

import json
import os

import numpy as np

HEADER = 64
PAGE = 4096


def capacity(n: int, size: int) -> int:
    cap = (PAGE - HEADER) // size
    while cap < n:
        cap *= 2
    return cap


def read_diskarray(path, dtype):
    """(count, element size, data) of a DiskArray file."""
    dt = np.dtype(dtype)
    raw = np.fromfile(path, dtype=np.uint8)
    count, size = np.frombuffer(raw[:16].tobytes(), dtype="<i8")
    assert size == dt.itemsize, (path, size, dt)
    data = np.frombuffer(raw[HEADER:HEADER + int(count) * dt.itemsize].tobytes(), dtype=dt)
    return int(count), int(size), data


def write_diskarray(path, arr) -> None:
    a = np.ascontiguousarray(arr)
    hdr = np.zeros(HEADER, dtype=np.uint8)
    hdr[:16] = np.frombuffer(np.array([len(a), a.dtype.itemsize], dtype="<i8").tobytes(),
                             dtype=np.uint8)
    total = HEADER + capacity(len(a), a.dtype.itemsize) * a.dtype.itemsize
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(a.tobytes())
        f.write(b"\0" * (total - HEADER - a.nbytes))


def importer_csr(V: int, edges, directed: bool, allow_multiple: bool):
    """Stored CSRs of one relation of a vertex type to itself.
    Returns (fwd_index int64[V+1], fwd_data uint32, bwd_index, bwd_data, count);
    the backward pair is None for undirected relations."""
    e = [tuple(map(int, x)) for x in edges]
    if not directed:  # add_reverse_edges
        e = e + [(d, s) for s, d in e]
    if not allow_multiple:  # std::unique: only adjacent repeats go
        u = []
        for x in e:
            if not u or u[-1] != x:
                u.append(x)
        e = u

    def csr(key_pairs):
        key_pairs = sorted(key_pairs)
        idx = np.zeros(V + 1, dtype=np.int64)
        for k, _ in key_pairs:
            idx[k + 1] += 1
        return np.cumsum(idx).astype(np.int64), np.array([v for _, v in key_pairs],
                                                         dtype=np.uint32)

    fi, fd = csr(e)
    if directed:
        bi, bd = csr([(d, s) for s, d in e])
    else:
        bi = bd = None
    return fi, fd, bi, bd, len(e)


def store_names(src: str, typ: str, dst: str, directed: bool):
    fwd = f"{src}-{typ}->{dst}" if directed else f"{src}-{typ}-{dst}"
    bwd = f"{dst}<-{typ}-{src}" if directed else f"{dst}-{typ}-{src}"
    return {"fwd_data": f"{fwd}_csr_src_dst", "fwd_index": f"{fwd}_csr_src_dst_index",
            "bwd_data": f"{bwd}_csr_dst_src", "bwd_index": f"{bwd}_csr_dst_src_index"}


def read_store(directory, edge: str | None = None):
    """(V, offsets uint64, adj uint32, info) of one relation, lists as sets."""
    with open(os.path.join(directory, "schema.json")) as f:
        sch = json.load(f)
    vcount = {v["type"]: int(v["statistics"]["count"]) for v in sch["vertex"]}
    rels = sch["edge"]
    if edge:
        hits = [r for r in rels if edge in (r["type"], f"{r['src']}_{r['type']}_{r['dst']}",
                                            f"{r['dst']}_{r['type']}_{r['src']}")]
    else:
        hits = [r for r in rels if r["src"] == r["dst"]]
    if len(hits) != 1:
        raise ValueError(f"{len(hits)} relations match {edge!r}")
    r = hits[0]
    if r["src"] != r["dst"]:
        raise NotImplementedError("relation between two vertex types")
    V = vcount[r["src"]]
    E = int(r["statistics"]["count"])
    names = store_names(r["src"], r["type"], r["dst"], bool(r["directed"]))
    ni, _, idx = read_diskarray(os.path.join(directory, names["fwd_index"]), np.int64)
    nd, _, data = read_diskarray(os.path.join(directory, names["fwd_data"]), np.uint32)
    assert ni == V + 1 and nd == E and idx[0] == 0 and idx[-1] == E
    offs = [0]
    out = []
    info = {"num_vertices": V, "stored_entries": E, "duplicates": 0, "unsorted_lists": 0,
            "self_loops": 0, "directed": int(bool(r["directed"])),
            "allow_multiple_edge": int(bool(r.get("allow_multiple_edge", False)))}
    for u in range(V):
        lst = data[idx[u]:idx[u + 1]].astype(np.int64)
        if len(lst) > 1 and np.any(np.diff(lst) < 0):
            info["unsorted_lists"] += 1
        info["self_loops"] += int(np.sum(lst == u))
        s = np.unique(lst)
        info["duplicates"] += len(lst) - len(s)
        out.append(s)
        offs.append(offs[-1] + len(s))
    adj = (np.concatenate(out) if out else np.zeros(0)).astype(np.uint32)
    info["num_edges"] = len(adj)
    return V, np.array(offs, dtype=np.uint64), adj, info


def hilbert_key(n: int, x, y) -> np.ndarray:
    """Curve position of (x, y) on an n x n grid, n a power of two."""
    x = np.asarray(x, dtype=np.uint64).copy()
    y = np.asarray(y, dtype=np.uint64).copy()
    d = np.zeros(x.shape, dtype=np.uint64)
    s = np.uint64(n // 2)
    while s > 0:
        rx = ((x & s) > 0).astype(np.uint64)
        ry = ((y & s) > 0).astype(np.uint64)
        d += s * s * ((np.uint64(3) * rx) ^ ry)
        low = s - np.uint64(1)
        flip = (ry == 0) & (rx == 1)
        x = np.where(flip, low - (x & low), x & low)
        y = np.where(flip, low - (y & low), y & low)
        swap = ry == 0
        x, y = np.where(swap, y, x), np.where(swap, x, y)
        s = s // np.uint64(2)
    return d


def partition_by_dst(dst, count: int) -> list:
    """Cut points (count + 1) of edges sorted by (dst, src): partition i grows
    until it holds (remaining / partitions left) edges and the destination
    changes.  Partitions after the edges run out are empty."""
    n = len(dst)
    cut, pos = [0], 0
    for i in range(count):
        if pos < n:
            expected = (n - pos) // (count - i) + pos
            while True:
                pos += 1
                if pos == n or (pos >= expected and dst[pos] != dst[pos - 1]):
                    break
        cut.append(pos)
    return cut


def flat_files(V: int, offsets, adj, directed: bool, parts: int, src: str = "Vertex",
               typ: str = "edge", dst: str = "Vertex") -> dict:
    """{file name: (n, 2) uint32 pairs} of the flat edge files."""
    off = np.asarray(offsets, dtype=np.int64)
    a = np.asarray(adj, dtype=np.uint32)
    s = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off))
    e = np.stack([s, a], axis=1)
    n = 1
    while n < V:
        n *= 2
    fwd = f"{src}-{typ}->{dst}" if directed else f"{src}-{typ}-{dst}"
    names = [(fwd, False)] + ([(f"{dst}<-{typ}-{src}", True)] if directed else [])

    def by_src_dst(x):
        return x[np.lexsort((x[:, 1], x[:, 0]))]

    def by_curve(x):
        k = hilbert_key(n, x[:, 0], x[:, 1])
        return x[np.argsort(k, kind="stable")]

    out = {}

    def put(name, bwd, x):
        out[name] = np.ascontiguousarray(x[:, ::-1] if bwd else x)

    for nm, bwd in names:
        put(nm + "_src_dst", bwd, by_src_dst(e))
        put(nm + "_hilbert", bwd, by_curve(e))
    ds = e[np.lexsort((e[:, 0], e[:, 1]))]
    cut = partition_by_dst(ds[:, 1].tolist(), parts)
    for i in range(parts):
        p = ds[cut[i]:cut[i + 1]]
        for nm, bwd in names:
            put(f"{nm}_src_dst_{i}", bwd, by_src_dst(p))
            put(f"{nm}_hilbert_{i}", bwd, by_curve(p))
    return out
