"""cd_tritri_batch_noalloc (the allocation-free triangle-pair stream) against
cd_tritri_batch, through the C ABI.

The allocation-free call takes device pointers the caller owns and keeps its
scratch (counters, the exact mode's rerun lists) in a fixed device array of
the library.  Its flags and segments must equal the allocating entry point's
bit for bit on the same pairs, in both modes, on two calls in a row where the
second is smaller (the array keeps the first call's contents), and the call
must leave the device's free memory unchanged.  Device buffers come from the
HIP runtime the library itself is linked against.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import case
from inmyroom_vulkan_amd import MODE_BATCHED, MODE_PLAIN, scenes
from inmyroom_vulkan_amd.collide import CD_E_CAPACITY, CD_E_INVALID, CD_NOALLOC_MAX_PAIRS, CD_OK, _ptr

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")  # the runtime the collision library already loaded
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(int(nbytes), 4)) == 0
        return p.value

    def free(self, p):
        assert self.lib.hipFree(C.c_void_p(p)) == 0

    def put(self, dst, arr):
        assert self.lib.hipMemcpy(C.c_void_p(dst), C.c_void_p(arr.ctypes.data), arr.nbytes, H2D) == 0

    def get(self, arr, src):
        assert self.lib.hipMemcpy(C.c_void_p(arr.ctypes.data), C.c_void_p(src), arr.nbytes, D2H) == 0

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        assert self.lib.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value


def _noalloc(gpu, hip, bufs, n, mode, with_seg=True):
    a, b, fl, sg = bufs
    rc = gpu.lib.cd_tritri_batch_noalloc(gpu.h, C.c_void_p(a), C.c_void_p(b), n, mode, C.c_void_p(fl),
                                         C.c_void_p(sg) if with_seg else None)
    gpu._check(rc, "cd_tritri_batch_noalloc")
    flags = np.zeros(n, np.int32)
    seg = np.zeros((n, 6), np.float32)
    hip.get(flags, fl)
    if with_seg:
        hip.get(seg, sg)
    return flags, seg


def _same(ref, got, mode):
    fr, sr = ref
    fg, sg = got
    assert np.array_equal(fr, fg)
    hit = (fr & 1) != 0 if mode == MODE_BATCHED else (fr & 3) == 1
    assert hit.any()
    assert sr[hit].tobytes() == sg[hit].tobytes()  # segments bit-identical where defined


def test_noalloc_matches_tritri_batch_two_calls():
    from inmyroom_vulkan_amd import Collider

    hip = _Hip()
    n1, n2 = 300_000, 70_000
    a1, b1, _ = scenes.config4(n1, seed=41)
    a2, b2, _ = scenes.config4(n2, seed=42)
    with Collider(leaf_max=4) as gpu:
        bufs = (hip.alloc(n1 * 36), hip.alloc(n1 * 36), hip.alloc(n1 * 4), hip.alloc(n1 * 24))
        try:
            for mode in (MODE_BATCHED, MODE_PLAIN):
                for call, (a, b) in enumerate(((a1, b1), (a2, b2))):
                    n = len(a)
                    case(f"noalloc tritri mode={mode} call={call} n={n}")
                    ref = gpu.tritri(a, b, mode)
                    ref_ctr = gpu.counters()
                    hip.put(bufs[0], a)
                    hip.put(bufs[1], b)
                    got = _noalloc(gpu, hip, bufs, n, mode)
                    _same(ref, got, mode)
                    ctr = gpu.counters()
                    assert ctr["tri_tests"] == n
                    assert ctr["fp64_fallbacks"] == ref_ctr["fp64_fallbacks"]
                    assert ctr["exact_fallbacks"] == ref_ctr["exact_fallbacks"]
                    if mode == MODE_BATCHED:
                        assert ctr["fp64_fallbacks"] > 0  # the rerun lists in the scratch array were used
            # flags only (no segment buffer), after the larger calls
            case("noalloc tritri without segments")
            fl, _ = _noalloc(gpu, hip, bufs, n2, MODE_BATCHED, with_seg=False)
            assert np.array_equal(fl, gpu.tritri(a2, b2, MODE_BATCHED)[0])
            # a warm call allocates no device memory
            hip.put(bufs[0], a1)
            hip.put(bufs[1], b1)
            before = hip.free_bytes()
            _noalloc(gpu, hip, bufs, n1, MODE_BATCHED)
            assert hip.free_bytes() == before
        finally:
            for p in bufs:
                hip.free(p)


def test_noalloc_rejects():
    from inmyroom_vulkan_amd import Collider

    hip = _Hip()
    case("noalloc tritri argument checks")
    with Collider(leaf_max=4) as gpu:
        f = gpu.lib.cd_tritri_batch_noalloc
        d = hip.alloc(64 * 36)
        try:
            # more pairs than the scratch array holds: refused before any device work
            assert f(gpu.h, C.c_void_p(d), C.c_void_p(d), CD_NOALLOC_MAX_PAIRS + 1, MODE_BATCHED, C.c_void_p(d),
                     None) == CD_E_CAPACITY
            # buffers shorter than n pairs, and host memory, are refused
            assert f(gpu.h, C.c_void_p(d), C.c_void_p(d), 65, MODE_BATCHED, C.c_void_p(d), None) == CD_E_INVALID
            host = np.zeros(64 * 9, np.float32)
            assert f(gpu.h, C.c_void_p(_ptr(host)), C.c_void_p(d), 64, MODE_PLAIN, C.c_void_p(d),
                     None) == CD_E_INVALID
            # a refused pointer leaves no error behind for the next call
            rc = f(gpu.h, None, None, 0, MODE_PLAIN, None, None)  # nothing to do
            assert rc == CD_OK, gpu.lib.cd_last_error(gpu.h)
            assert f(gpu.h, C.c_void_p(d), C.c_void_p(d), 1, 7, C.c_void_p(d), None) == CD_E_INVALID
            # and a real call after the refusals works
            a, b, _ = scenes.config4(64, seed=43)
            hip.put(d, a)
            dd = hip.alloc(64 * 36 + 64 * 4)
            try:
                hip.put(dd, b)
                rc = f(gpu.h, C.c_void_p(d), C.c_void_p(dd), 64, MODE_PLAIN, C.c_void_p(dd + 64 * 36), None)
                assert rc == CD_OK, gpu.lib.cd_last_error(gpu.h)
                fl = np.zeros(64, np.int32)
                hip.get(fl, dd + 64 * 36)
                assert np.array_equal(fl, gpu.tritri(a, b, MODE_PLAIN)[0])
            finally:
                hip.free(dd)
        finally:
            hip.free(d)
