"""Oracle (test infrastructure only): CPU restatement of the bit-matrix table
operators between VExpand and Mintersect, on plain numpy boolean matrices.

Only tests/, the smoke entry and the bench's CPU leg use this module; the
product path never imports it.

A table here is (M, row_map, col_map): M a rows x cols bool array, row_map /
col_map None or uint32 arrays of vertex ids (MatrixTable, matrix_table.hh:4-37).

* colmain_words      -- DenseColMainMatrix word layout: word(r, c) =
                        c * align8(ceil(R/64)) + r / 64, bit r % 64
                        (bitmap.hh:262-326, bitmap.cpp:580-595).
* table_convert      -- TableConvert VarRow -> ColMain, with / without
                        transpose (big_operators2.cpp:16-86): the bits stay,
                        the maps travel (take_maps_from) and swap on transpose.
* select_cols        -- MatrixTableSelect (big_operators2.cpp:152-213).
* column_count       -- MatrixColumnCount (matrix_column_count.cpp:19-87),
                        optional (count, id) pair sort.
* row_or_sum         -- MatrixRowOrSum (matrix_row_or_sum.cpp:14-55).
* mintersect_colmain -- Mintersect, flat base and ColMain dependencies
                        (mintersect.cpp:48-243 with tackle_di_for_base_vid,
                        :139-156; find_col_idx, bit_matrix_table.cpp:26-57).
* mintersect_varrow  -- Mintersect, VarRow base (mintersect.cpp:244-333):
                        base pairs from the set bits (varrow_tuples), ColMain
                        or edge-relation dependencies.

Pinned by: the reference's own MintersectFlatMatMat fixture (3 tuples,
mintersect.cpp:411-512), a hand-computed word layout, the reference's
TableConvert properties (big_operators2.cpp:215-297: m(i,j) == c(i,j), and
== c(j,i) transposed), ColMain-of-adjacency intersections checked against the
sorted-merge C oracle, and (count, id) ordering with ties worked by hand;
mintersect_varrow by a hand-worked triangle case and a brute-force triple loop
(start, target, w) over every vertex w on small random graphs.
"""
from __future__ import annotations

import numpy as np


def colmain_words(M) -> np.ndarray:
    """rows x cols bool -> cols x pitch uint64 (bitmap.hh:277-281)."""
    M = np.asarray(M, dtype=bool)
    R, Cn = M.shape
    pitch = (((R + 63) // 64) + 7) // 8 * 8
    out = np.zeros((Cn, pitch), dtype=np.uint64)
    r, c = np.nonzero(M)
    np.bitwise_or.at(out, (c, r // 64), np.left_shift(np.uint64(1), (r % 64).astype(np.uint64)))
    return out


def table_convert(M, row_map, col_map, transpose: bool):
    """VarRow (rows x cols) -> ColMain; big_operators2.cpp:27-62."""
    M = np.asarray(M, dtype=bool)
    if transpose:
        return M.T.copy(), col_map, row_map
    return M.copy(), row_map, col_map


def select_cols(M, row_map, col_ids):
    """MatrixTableSelect: column k <- old column col_ids[k]; big_operators2.cpp:165-196."""
    ids = np.asarray(col_ids, dtype=np.int64)
    return np.asarray(M, dtype=bool)[:, ids], row_map, np.asarray(col_ids, dtype=np.uint32)


def find_col_idx(col_map, ncols: int, x: int):
    """bit_matrix_table.cpp:48-57 with set_col_hash_map's insert (first wins)."""
    if col_map is None:
        return x if x < ncols else None
    hits = np.nonzero(np.asarray(col_map) == x)[0]
    return int(hits[0]) if len(hits) else None


def column_count(M, col_map, order=None):
    """(ids, counts); order None / 'asc' / 'desc' sorts (count, id) pairs."""
    M = np.asarray(M, dtype=bool)
    counts = M.sum(axis=0).astype(np.int64)
    ids = np.arange(M.shape[1], dtype=np.int64)
    if col_map is not None:  # columns past a short map keep their index
        ids[:len(col_map)] = np.asarray(col_map, dtype=np.int64)
    pairs = list(zip(counts.tolist(), ids.tolist()))
    if order == "asc":
        pairs.sort()
    elif order == "desc":
        pairs.sort(reverse=True)
    return (np.array([p[1] for p in pairs], dtype=np.uint32),
            np.array([p[0] for p in pairs], dtype=np.int32))


def row_or_sum(M, col_map):
    M = np.asarray(M, dtype=bool)
    ids = np.arange(M.shape[1], dtype=np.int64)
    if col_map is not None:
        ids[:len(col_map)] = np.asarray(col_map, dtype=np.int64)
    return ids[np.nonzero(M.any(axis=0))[0]].astype(np.uint32)


def mintersect_colmain(deps, tuples):
    """deps: list of (M, row_map, col_map) with one row count; tuples: T x nd.
    Returns (counts, offsets, lists) in tuple order, rows ascending."""
    tup = np.asarray(tuples, dtype=np.int64).reshape(-1, len(deps))
    R = np.asarray(deps[0][0]).shape[0]
    target = deps[0][1]  # the first ColMain dependency's row map
    counts, lists = [], []
    for t in tup:
        ir = np.ones(R, dtype=bool)
        for d, (M, _, cmap) in enumerate(deps):
            M = np.asarray(M, dtype=bool)
            c = find_col_idx(cmap, M.shape[1], int(t[d]))
            if c is None:
                ir[:] = False
            else:
                ir &= M[:, c]
        rows = np.nonzero(ir)[0]
        if target is not None:
            rows = np.asarray(target)[rows]
        counts.append(len(rows))
        lists.append(rows.astype(np.uint32))
    counts = np.array(counts, dtype=np.uint32)
    offsets = np.zeros(len(counts) + 1, dtype=np.uint64)
    np.cumsum(counts, out=offsets[1:])
    lists = np.concatenate(lists) if lists else np.zeros(0, dtype=np.uint32)
    return counts, offsets, lists


def varrow_tuples(M, row_map):
    """The set bits of a VarRow matrix as (row vertex, column vertex) pairs in
    (row, column) order; the row vertex goes through row_map
    (VarRowMatrixTable::Iterator::next, bit_matrix_table.cpp:147-160)."""
    r, c = np.nonzero(np.asarray(M, dtype=bool))
    if row_map is not None:
        r = np.asarray(row_map, dtype=np.int64)[r]
    return np.stack([r, c], axis=1).astype(np.uint32)


def mintersect_varrow(M, row_map, deps, var_idx, adjacency=None):
    """Mintersect with a VarRow base (mintersect.cpp:244-333): for every base
    pair (start, target) dependency d intersects the column bound to
    base variable var_idx[d]; deps are ColMain triples (as mintersect_colmain)
    or, with adjacency = [(offsets, adj), ...], edge relations whose bound
    vertex's neighbours form the set.  Returns (start, target, new) arrays,
    ordered by base pair, then new vertex."""
    base = varrow_tuples(M, row_map)
    nd = len(var_idx)
    tup = base[:, np.asarray(var_idx, dtype=np.int64)].reshape(-1, nd)
    if adjacency is None:
        # mintersect_colmain over chunks of tuples at once: row t of `ir` is
        # the AND of the columns tuple t names (all False for a missing one)
        R = np.asarray(deps[0][0]).shape[0]
        cols = []
        for d, (Md, _, cmap) in enumerate(deps):
            Md = np.asarray(Md, dtype=bool)
            lut = {}
            ids = range(Md.shape[1]) if cmap is None else np.asarray(cmap).tolist()
            for c, x in enumerate(ids):
                lut.setdefault(int(x), c)
            cols.append((Md.T, np.array([lut.get(int(x), -1) for x in tup[:, d]],
                                        dtype=np.int64)))
        counts, lists = [], []
        for a in range(0, len(tup), 2048):
            b = min(len(tup), a + 2048)
            ir = np.ones((b - a, R), dtype=bool)
            for MT, ci in cols:
                ok = ci[a:b] >= 0
                ir[~ok] = False
                ir[ok] &= MT[ci[a:b][ok]]
            t_i, r_i = np.nonzero(ir)
            if deps[0][1] is not None:
                r_i = np.asarray(deps[0][1])[r_i]
            counts.append(np.bincount(t_i, minlength=b - a))
            lists.append(r_i.astype(np.uint32))
        counts = np.concatenate(counts) if counts else np.zeros(0, dtype=np.int64)
        lists = np.concatenate(lists) if lists else np.zeros(0, dtype=np.uint32)
    else:
        counts, lists = [], []
        for t in tup:
            s = None
            for d, (off, adj) in enumerate(adjacency):
                v = int(t[d])
                nb = np.asarray(adj[off[v]:off[v + 1]], dtype=np.uint32)
                s = nb if s is None else np.intersect1d(s, nb)
            counts.append(len(s))
            lists.append(np.sort(s).astype(np.uint32))
        counts = np.array(counts, dtype=np.int64)
        lists = np.concatenate(lists) if lists else np.zeros(0, dtype=np.uint32)
    rep = np.repeat(np.arange(len(base)), np.asarray(counts, dtype=np.int64))
    return base[rep, 0], base[rep, 1], np.asarray(lists, dtype=np.uint32)


def adjacency_colmain(offsets, adj, V: int):
    """ColMain of a CSR: rows = vertices, column u = N(u)."""
    M = np.zeros((V, V), dtype=bool)
    off = np.asarray(offsets, dtype=np.int64)
    for u in range(V):
        M[np.asarray(adj[off[u]:off[u + 1]], dtype=np.int64), u] = True
    return M
