"""apc_morton_sort_noalloc, the entry point that allocates nothing, against
apc_octree_build + apc_get_sorted (the context path) on the same particles:
keys and sorted order bit for bit, on calls in a row that share the fixed
workspace (the second with a smaller input), an odd number of radix passes,
and the size limit.  Device buffers come from the HIP runtime the library
itself links (by its soname), not from torch."""
import ctypes as C

import numpy as np
import pytest

import apcloud
from apcloud import ic

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    apcloud.lib()  # loads libamdhip64.so.7, the runtime the library uses
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


class Dev:
    """A device copy of a numpy array (or an empty one of its shape)."""

    def __init__(self, H, a, upload=True):
        self.H, self.a = H, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), max(1, self.a.nbytes)) == 0
        if upload and self.a.nbytes:
            assert H.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = C.c_void_p()


def context_sort(x, y, z, bbox, depth):
    g = apcloud.Context(apcloud.make_params(bbox=bbox, depth=depth, bc=1, c=0.02, lm=27))
    g.set_particles(x, y, z, np.zeros_like(x), np.zeros_like(x), np.zeros_like(x))
    g.octree_build()
    k, i = g.sorted()
    g.close()
    return k, i


def noalloc_sort(H, x, y, z, bbox, depth):
    n = len(x)
    bufs = [Dev(H, a) for a in (x, y, z)]
    kd = Dev(H, np.zeros(n, np.uint64), upload=False)
    idd = Dev(H, np.zeros(n, np.int32), upload=False)
    bb = np.ascontiguousarray(bbox, np.float64)
    L = apcloud.lib()
    print("noalloc: n=%d depth=%d" % (n, depth), flush=True)
    r = L.apc_morton_sort_noalloc(n, bufs[0].p, bufs[1].p, bufs[2].p,
                                  bb.ctypes.data_as(C.POINTER(C.c_double)), depth, kd.p, idd.p,
                                  None)
    assert r == 0, L.apc_last_error(None)
    assert H.hipDeviceSynchronize() == 0
    k, i = kd.get(), idd.get()
    for b in bufs + [kd, idd]:
        b.free()
    return k, i


def test_noalloc_sort_matches_the_context_path_twice():
    H = hip()
    # a clustered cloud (Zel'dovich lattice, RMS 0.8 spacings) at depth 20:
    # 60 key bits, 8 radix passes
    x, y, z, *_ = ic.zeldovich(64, seed=11, rms=0.05)
    bbox = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    k0, i0 = context_sort(x, y, z, bbox, 20)
    k1, i1 = noalloc_sort(H, x, y, z, bbox, 20)
    assert np.array_equal(k0, k1) and np.array_equal(i0, i1)
    assert np.all(np.diff(k1.astype(np.uint64)) >= 0)
    # second call, smaller input, another box and depth 7: 21 key bits, 3
    # passes (an odd count ends in the copy kernel); the workspace holds the
    # first call's contents
    rng = np.random.default_rng(12)
    n2 = 40_000
    x2 = rng.uniform(-0.5, 1.5, n2)
    y2 = rng.uniform(-0.5, 1.5, n2)
    z2 = rng.normal(0.5, 0.2, n2)
    bbox2 = (-0.5, 1.5, -0.5, 1.5, -0.5, 1.5)
    k0, i0 = context_sort(x2, y2, z2, bbox2, 7)
    k2, i2 = noalloc_sort(H, x2, y2, z2, bbox2, 7)
    assert np.array_equal(k0, k2) and np.array_equal(i0, i2)
    # ties are common at depth 7: the order within equal keys is the input
    # order (stable), as in the context path's first sort
    same = k2[1:] == k2[:-1]
    assert same.any() and np.all(i2[1:][same] > i2[:-1][same])


def test_noalloc_single_particle_and_size_limit():
    H = hip()
    k, i = noalloc_sort(H, np.array([0.3]), np.array([0.7]), np.array([0.1]),
                        (0, 1, 0, 1, 0, 1), 20)
    k0, i0 = context_sort(np.array([0.3]), np.array([0.7]), np.array([0.1]),
                          (0, 1, 0, 1, 0, 1), 20)
    assert np.array_equal(k, k0) and list(i) == [0]
    L = apcloud.lib()
    dummy = Dev(H, np.zeros(4))
    bb = np.array([0, 1, 0, 1, 0, 1], np.float64)
    r = L.apc_morton_sort_noalloc(L_MAX() + 1, dummy.p, dummy.p, dummy.p,
                                  bb.ctypes.data_as(C.POINTER(C.c_double)), 20, dummy.p, dummy.p,
                                  None)
    assert r == -1
    assert b"exceeds" in L.apc_last_error(None)
    r = L.apc_morton_sort_noalloc(4, dummy.p, dummy.p, dummy.p,
                                  bb.ctypes.data_as(C.POINTER(C.c_double)), 22, dummy.p, dummy.p,
                                  None)
    assert r == -1
    dummy.free()


def L_MAX():
    return 1 << 21  # APC_NOALLOC_MAX_N in include/apcloud.h
