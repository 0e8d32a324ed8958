"""vs_mintersect_device_noalloc: the Mintersect entry point that allocates
nothing (its working memory is a fixed device array) must give exactly what
vs_mintersect_device gives on the same device-resident inputs, on two calls in
a row -- the second smaller, so stale contents of the shared array from the
first call would show -- with and without lists, and reject what does not fit.
"""
import ctypes as C

import numpy as np
import pytest

import vertexsurge_amd as vs
from vertexsurge_amd import _lib

pytestmark = pytest.mark.gpu


class Dev:
    """Device buffers through the library's raw allocator (freed on close)."""

    def __init__(self):
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        _lib.call("vs_dev_alloc", max(8, nbytes), C.byref(p))
        self.ptrs.append(p)
        return p

    def put(self, a):
        p = self.alloc(a.nbytes)
        _lib.call("vs_memcpy_h2d", p, a.ctypes.data_as(C.c_void_p), a.nbytes)
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype=dtype)
        if n:
            _lib.call("vs_memcpy_d2h", out.ctypes.data_as(C.c_void_p), p, out.nbytes)
        return out

    def close(self):
        for p in self.ptrs:
            _lib.call("vs_dev_free", p)
        self.ptrs = []


def run(fn, deps, tup, dev, with_lists=True, cap=None):
    """Calls fn over device copies of tup; returns (counts, offsets, lists, total)."""
    T, arity = tup.shape
    d_t = dev.put(np.ascontiguousarray(tup))
    d_c = dev.alloc(T * 4)
    d_o = dev.alloc((T + 1) * 8)
    arr = (C.c_void_p * arity)(*[G.handle for G in deps])
    total = C.c_uint64()
    d_l = None
    if with_lists:
        # size query through the reference entry point, then the real call
        _lib.call("vs_mintersect_device", arr, arity, d_t, T, d_c, d_o, None, 0,
                  C.byref(total), None)
        cap = total.value if cap is None else cap
        d_l = dev.alloc(max(1, total.value) * 4)
    args = [arr, arity, d_t, T, d_c, d_o, d_l, cap or 0, C.byref(total)]
    if fn == "vs_mintersect_device":
        args.append(None)
    _lib.call(fn, *args)
    n = total.value
    return (dev.get(d_c, T, np.uint32), dev.get(d_o, T + 1, np.uint64),
            dev.get(d_l, n, np.uint32) if with_lists else None, n)


@pytest.fixture(scope="module")
def graphs():
    g = vs.HostCSR.rmat(16, 16, seed=21)
    h = vs.HostCSR.rmat(15, 8, seed=22)
    return g, vs.Graph.from_csr(g), h, vs.Graph.from_csr(h)


def test_noalloc_equals_device_entry_point(orc, graphs):
    g, G, h, H = graphs
    dev = Dev()
    try:
        # call 1: the full 4,194,304-triple batch of the intersection workload;
        # call 2: fewer triples, another arity and a second graph
        t1 = vs.sample_paths3(g.offsets, g.adj, 4 << 20, seed=31)
        t2 = vs.sample_paths3(g.offsets, g.adj, 70001, seed=32)[:, :2].copy()
        t2[:, 1] %= h.V
        t2[::1000, 0] = g.V + 5  # ids beyond a dependency: empty sets
        for deps, tup in (([G, G, G], t1), ([G, H], t2)):
            ref = run("vs_mintersect_device", deps, tup, dev)
            got = run("vs_mintersect_device_noalloc", deps, tup, dev)
            assert got[3] == ref[3]
            for a, b in zip(got[:3], ref[:3]):
                assert np.array_equal(a, b)
            cnt = run("vs_mintersect_device_noalloc", deps, tup, dev, with_lists=False)
            assert np.array_equal(cnt[0], ref[0]) and np.array_equal(cnt[1], ref[1])
            assert cnt[3] == ref[3]
        # the oracle on a slice of the second call
        c2, o2, l2 = orc.mintersect_sorted([(g.offsets, g.adj), (h.offsets, h.adj)], t2[:5000])
        assert np.array_equal(got[0][:5000], c2)
        assert np.array_equal(got[2][:int(o2[-1])], l2)
        assert got[3] > 0
    finally:
        dev.close()


def test_noalloc_giant_tuples(orc):
    """Tuples whose shortest list exceeds 8192 entries (counted apart in the
    shared array): a mixed call, then a smaller one of giants only."""
    from test_mintersect_gpu import giant_tuples, giants_graph
    off, adj = giants_graph()
    G = vs.Graph(off, adj, symmetric=False)
    dev = Dev()
    try:
        for T, seed in ((5000, 41), (30, 42)):
            tup = giant_tuples(T, len(off) - 1, seed, hi=5 if T > 30 else 3)
            ref = run("vs_mintersect_device", [G, G, G], tup, dev)
            got = run("vs_mintersect_device_noalloc", [G, G, G], tup, dev)
            assert got[3] == ref[3]
            for a, b in zip(got[:3], ref[:3]):
                assert np.array_equal(a, b)
            c2, o2, l2 = orc.mintersect_sorted([(off, adj)] * 3, tup)
            assert np.array_equal(got[0], c2) and np.array_equal(got[2], l2)
    finally:
        dev.close()


def test_noalloc_capacity_and_limits(graphs):
    g, G, _, _ = graphs
    dev = Dev()
    try:
        tup = vs.sample_paths3(g.offsets, g.adj, 20000, seed=33)
        ref = run("vs_mintersect_device", [G, G, G], tup, dev)
        assert ref[3] > 0
        with pytest.raises(vs.VSError) as e:
            run("vs_mintersect_device_noalloc", [G, G, G], tup, dev, cap=ref[3] - 1)
        assert e.value.code == _lib.VS_ERR_CAPACITY
        # host-side checks, before any device work
        arr = (C.c_void_p * 3)(G.handle, G.handle, G.handle)
        total = C.c_uint64()
        d = dev.alloc(64)
        for T, off in ((_lib.NOALLOC_MAX_TUPLES + 1, d), (10, None)):
            with pytest.raises(vs.VSError) as e:
                _lib.call("vs_mintersect_device_noalloc", arr, 3, d, T, d, off, None, 0,
                          C.byref(total))
            assert e.value.code == _lib.VS_ERR_INVALID
        # T = 0: offsets[0] = 0
        _lib.call("vs_mintersect_device_noalloc", arr, 3, None, 0, None, d, None, 0,
                  C.byref(total))
        assert total.value == 0 and dev.get(d, 1, np.uint64)[0] == 0
    finally:
        dev.close()
