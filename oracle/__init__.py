"""CPU oracle of the VExpand / Mintersect hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package, and only as the checker or as the timed CPU
baseline.  The product path (``vertexsurge_amd``) never imports it.

Contents:
  * ctypes wrappers of ``_build/liboracle.so`` (vs_oracle.cpp): oracle A
    (per-row semantic BFS), oracle B (the reference's stacked, dst-partitioned,
    Hilbert-blocked CPU algorithm -- the CPU baseline), the intersection
    oracles, a scalar BFS;
  * numpy restatements of the synthetic inputs (R-MAT, source sampling, path
    sampling) that pin the library's generators;
  * layout helpers (row-major <-> native tiled <-> reference DVSM).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
_LIB = None

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_vexpand_rows.argtypes = [u64p, u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_int, C.c_int, u64p, u32p, u64p, u64p]
        L.orc_bytes_alg.argtypes = [u64p, C.c_uint32, C.c_uint32, u64p, C.c_uint32,
                                    C.c_uint32, C.c_int, u64p, C.c_uint32]
        L.orc_bytes_alg.restype = C.c_uint64
        L.orc_blocks_build.argtypes = [u64p, u32p, C.c_uint32, C.c_uint32]
        L.orc_blocks_build.restype = C.c_void_p
        L.orc_blocks_free.argtypes = [C.c_void_p]
        L.orc_blocks_count.argtypes = [C.c_void_p]
        L.orc_blocks_count.restype = C.c_uint64
        L.orc_dvsm_words.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_dvsm_words.restype = C.c_uint64
        L.orc_vexpand_dvsm.argtypes = [C.c_void_p, u64p, C.c_uint32, u32p, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_int, u64p, u32p, u64p,
                                       C.POINTER(C.c_double)]
        L.orc_rowmajor_to_dvsm.argtypes = [u64p, C.c_uint32, C.c_uint32, u64p]
        L.orc_dvsm_row_counts.argtypes = [u64p, C.c_uint32, C.c_uint32, u64p]
        L.orc_mintersect_sorted.argtypes = [C.c_uint32, C.POINTER(u64p), C.POINTER(u32p),
                                            u32p, u32p, C.c_uint64, u32p, u64p, u32p]
        L.orc_mintersect_bitmap.argtypes = [C.c_uint32, C.POINTER(u64p), C.POINTER(u32p),
                                            u32p, C.c_uint32, u32p, C.c_uint64, u32p, u64p,
                                            u32p]
        L.orc_bfs_dist.argtypes = [u64p, u32p, C.c_uint32, C.c_uint32, i32p]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _csr(off, adj):
    return (np.ascontiguousarray(off, dtype=np.uint64),
            np.ascontiguousarray(adj if len(adj) else np.zeros(1, np.uint32), dtype=np.uint32))


# ------------------------------------------------------------------ oracle A
def vexpand_rows(off, adj, sources, k_min, k_max, path="shortest", with_distance=False):
    """Returns (matrices [n, S, ceil(V/64)] uint64 row-major, k_actual, te, layer_pop)."""
    off, adj = _csr(off, adj)
    V = len(off) - 1
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    S = len(src)
    nw = (V + 63) // 64
    nmat = (k_max - k_min + 1) if with_distance else 1
    out = np.zeros((max(nmat, 1), S, nw), dtype=np.uint64)
    ka = C.c_uint32()
    te = C.c_uint64()
    pop = np.zeros(k_max + 1, dtype=np.uint64)
    lib().orc_vexpand_rows(_p(off, C.c_uint64), _p(adj, C.c_uint32), V, _p(src, C.c_uint32), S,
                           k_min, k_max, 1 if path == "shortest" else 0, int(with_distance),
                           _p(out, C.c_uint64), C.byref(ka), C.byref(te), _p(pop, C.c_uint64))
    if with_distance:
        out = out[: max(0, ka.value - k_min + 1)]
    return out, ka.value, te.value, pop


def layers_rows(off, adj, sources, k_max, path="shortest"):
    """All layers L_0..L_k_max (row-major) plus k_actual, te."""
    mats, ka, te, pop = vexpand_rows(off, adj, sources, 0, k_max, path, with_distance=True)
    return mats, ka, te


def bytes_alg(off, sources_count, layers, k_actual, k_min, path, R, tile_rows):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    V = len(off) - 1
    lay = np.ascontiguousarray(layers, dtype=np.uint64)
    R = np.ascontiguousarray(R, dtype=np.uint64)
    return int(lib().orc_bytes_alg(_p(off, C.c_uint64), V, sources_count, _p(lay, C.c_uint64),
                                   k_actual, k_min, 1 if path == "shortest" else 0,
                                   _p(R, C.c_uint64), tile_rows))


def bfs_dist(off, adj, s):
    off, adj = _csr(off, adj)
    V = len(off) - 1
    d = np.zeros(V, dtype=np.int32)
    lib().orc_bfs_dist(_p(off, C.c_uint64), _p(adj, C.c_uint32), V, int(s), _p(d, C.c_int32))
    return d


# ------------------------------------------------------------------ oracle B
class EdgeBlocks:
    """Destination-partitioned, Hilbert-ordered edge blocks (offline import)."""

    def __init__(self, off, adj, nblocks=1000):
        self.off, self.adj = _csr(off, adj)
        self.V = len(self.off) - 1
        self._h = lib().orc_blocks_build(_p(self.off, C.c_uint64), _p(self.adj, C.c_uint32),
                                         self.V, nblocks)

    @property
    def count(self):
        return int(lib().orc_blocks_count(self._h))

    def vexpand(self, sources, k_min, k_max, path="shortest"):
        """Returns (dvsm words, k_actual, visit_count, ms)."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        words = int(lib().orc_dvsm_words(len(src), self.V))
        out = np.zeros(words, dtype=np.uint64)
        ka = C.c_uint32()
        vc = C.c_uint64()
        ms = C.c_double()
        lib().orc_vexpand_dvsm(self._h, _p(self.off, C.c_uint64), self.V, _p(src, C.c_uint32),
                               len(src), k_min, k_max, 1 if path == "shortest" else 0,
                               _p(out, C.c_uint64), C.byref(ka), C.byref(vc), C.byref(ms))
        return out, ka.value, vc.value, ms.value

    def close(self):
        if self._h:
            lib().orc_blocks_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rowmajor_to_dvsm(rm, V):
    rm = np.ascontiguousarray(rm, dtype=np.uint64)
    S = rm.shape[0]
    out = np.zeros(int(lib().orc_dvsm_words(S, V)), dtype=np.uint64)
    lib().orc_rowmajor_to_dvsm(_p(rm, C.c_uint64), S, V, _p(out, C.c_uint64))
    return out


def dvsm_row_counts(dv, S, V):
    """Per-row popcounts of a DVSM word array of S rows and V columns."""
    dv = np.ascontiguousarray(dv, dtype=np.uint64)
    if dv.size != int(lib().orc_dvsm_words(S, V)):
        raise ValueError("DVSM array size does not match S and V")
    out = np.zeros(S, dtype=np.uint64)
    lib().orc_dvsm_row_counts(_p(dv, C.c_uint64), S, V, _p(out, C.c_uint64))
    return out


def rowmajor_to_native(rm, V, tw):
    """Row-major bitsets (S x ceil(V/64)) -> native [tile][vertex][tw] words."""
    rm = np.asarray(rm, dtype=np.uint64)
    S = rm.shape[0]
    W = (S + 63) // 64
    T = (W + tw - 1) // tw
    bits = np.unpackbits(rm.view(np.uint8).reshape(S, -1), axis=1, bitorder="little")[:, :V]
    pad = np.zeros((T * tw * 64, V), dtype=np.uint8)
    pad[:S] = bits
    # word (w, v) = bits of rows 64w..64w+63 at column v, little-endian
    words = np.packbits(pad.reshape(T * tw, 64, V).transpose(0, 2, 1), axis=2,
                        bitorder="little").view(np.uint64)[..., 0]  # [W', V]
    return words.reshape(T, tw, V).transpose(0, 2, 1).reshape(-1).copy()


def row_counts(rm):
    rm = np.asarray(rm, dtype=np.uint64)
    return np.unpackbits(rm.view(np.uint8), axis=1).sum(axis=1).astype(np.uint64)


def digest_hash(c):
    c = np.asarray(c, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(0x5EEDC0DE) + (c + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def row_digests(rm, V):
    """digest[r] = sum of digest_hash(c) over set bits c of row r (mod 2^64)."""
    rm = np.asarray(rm, dtype=np.uint64)
    bits = np.unpackbits(rm.view(np.uint8).reshape(rm.shape[0], -1), axis=1,
                         bitorder="little")[:, :V]
    h = digest_hash(np.arange(V, dtype=np.uint64))
    out = np.zeros(rm.shape[0], dtype=np.uint64)
    with np.errstate(over="ignore"):
        for r in range(rm.shape[0]):
            out[r] = np.sum(h[bits[r].astype(bool)], dtype=np.uint64)
    return out


# ------------------------------------------------------------------ oracle C
def _dep_arrays(deps):
    offs = [np.ascontiguousarray(o, dtype=np.uint64) for o, _ in deps]
    adjs = [np.ascontiguousarray(a if len(a) else np.zeros(1, np.uint32), dtype=np.uint32)
            for _, a in deps]
    n = len(deps)
    po = (u64p * n)(*[_p(o, C.c_uint64) for o in offs])
    pa = (u32p * n)(*[_p(a, C.c_uint32) for a in adjs])
    Vs = np.array([len(o) - 1 for o in offs], dtype=np.uint32)
    return offs, adjs, po, pa, Vs


def mintersect_sorted(deps, tuples):
    """deps: list of (offsets, adj); returns (counts, offsets, lists)."""
    offs, adjs, po, pa, Vs = _dep_arrays(deps)
    tup = np.ascontiguousarray(tuples, dtype=np.uint32).reshape(-1, len(deps))
    T = tup.shape[0]
    counts = np.zeros(T, dtype=np.uint32)
    lib().orc_mintersect_sorted(len(deps), po, pa, _p(Vs, C.c_uint32), _p(tup, C.c_uint32), T,
                                _p(counts, C.c_uint32), None, None)
    offsets = np.zeros(T + 1, dtype=np.uint64)
    np.cumsum(counts, out=offsets[1:])
    lists = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
    lib().orc_mintersect_sorted(len(deps), po, pa, _p(Vs, C.c_uint32), _p(tup, C.c_uint32), T,
                                _p(counts, C.c_uint32), _p(offsets, C.c_uint64),
                                _p(lists, C.c_uint32))
    return counts, offsets, lists[: int(offsets[-1])]


def mintersect_bitmap(deps, tuples, V):
    offs, adjs, po, pa, Vs = _dep_arrays(deps)
    tup = np.ascontiguousarray(tuples, dtype=np.uint32).reshape(-1, len(deps))
    T = tup.shape[0]
    counts = np.zeros(T, dtype=np.uint32)
    lib().orc_mintersect_bitmap(len(deps), po, pa, _p(Vs, C.c_uint32), V, _p(tup, C.c_uint32),
                                T, _p(counts, C.c_uint32), None, None)
    offsets = np.zeros(T + 1, dtype=np.uint64)
    np.cumsum(counts, out=offsets[1:])
    lists = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint32)
    lib().orc_mintersect_bitmap(len(deps), po, pa, _p(Vs, C.c_uint32), V, _p(tup, C.c_uint32),
                                T, _p(counts, C.c_uint32), _p(offsets, C.c_uint64),
                                _p(lists, C.c_uint32))
    return counts, offsets, lists[: int(offsets[-1])]


# ------------------------------------------------------ synthetic-input pins
_G = np.uint64(0x9E3779B97F4A7C15)


def mix64(seed, ctr):
    """splitmix64 finaliser over seed + (ctr + 1) * golden, vectorised."""
    ctr = np.asarray(ctr, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (ctr + np.uint64(1)) * _G
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rmat_pairs(scale, edge_factor, a, b, c, seed):
    """The generated (u, v) pairs before cleanup (numpy restatement)."""
    npairs = edge_factor << scale
    nd = (scale + 1) // 2
    TA = np.uint64(round(a * 2**32))
    TAB = np.uint64(round((a + b) * 2**32))
    TABC = np.uint64(round((a + b + c) * 2**32))
    e = np.arange(npairs, dtype=np.uint64)
    u = np.zeros(npairs, dtype=np.uint64)
    v = np.zeros(npairs, dtype=np.uint64)
    d = None
    for lvl in range(scale):
        if lvl % 2 == 0:
            d = mix64(seed, e * np.uint64(nd) + np.uint64(lvl // 2))
        r = (d >> np.uint64(32)) if lvl % 2 else (d & np.uint64(0xFFFFFFFF))
        ub = (r >= TAB).astype(np.uint64)
        vb = (((r >= TA) & (r < TAB)) | (r >= TABC)).astype(np.uint64)
        u = (u << np.uint64(1)) | ub
        v = (v << np.uint64(1)) | vb
    return u.astype(np.uint32), v.astype(np.uint32)


def csr_from_pairs(V, u, v, symmetrize=True, drop_self=True):
    u = np.asarray(u, dtype=np.uint64)
    v = np.asarray(v, dtype=np.uint64)
    if drop_self:
        keep = u != v
        u, v = u[keep], v[keep]
    if symmetrize:
        u, v = np.concatenate([u, v]), np.concatenate([v, u])
    key = np.unique(u * np.uint64(V) + v)
    src = (key // np.uint64(V)).astype(np.int64)
    dst = (key % np.uint64(V)).astype(np.uint32)
    off = np.zeros(V + 1, dtype=np.uint64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off, dtype=np.uint64), dst


def sample_sources(V, S, seed):
    """Numpy restatement of the library's distinct uniform source sampler."""
    if 2 * S <= V:
        seen = set()
        out = []
        i = 0
        while len(out) < S:
            chunk = mix64(seed, np.arange(i, i + 4 * S + 64, dtype=np.uint64))
            xs = ((chunk >> np.uint64(32)) * np.uint64(V)) >> np.uint64(32)
            for x in xs.tolist():
                if x not in seen:
                    seen.add(x)
                    out.append(x)
                    if len(out) == S:
                        break
            i += 4 * S + 64
        return np.sort(np.array(out, dtype=np.uint32))
    perm = np.arange(V, dtype=np.uint32)
    for i in range(S):
        j = i + int(((int(mix64(seed, i)) >> 32) * (V - i)) >> 32)
        perm[i], perm[j] = perm[j], perm[i]
    return np.sort(perm[:S])
