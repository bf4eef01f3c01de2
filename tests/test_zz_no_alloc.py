"""psba_generate_idxs_device_noalloc (no allocation: device pointers in and
out, a fixed device scratch array) against psba_generate_idxs_device (the
allocating entry point) on the same visibility masks: iidx, jidx, blk_idx and
the count bit for bit, on two calls in a row, the second with a smaller
input; sizes past the scratch limit are rejected before any device work."""
import ctypes as ct
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip():
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ct.CDLL(name)
        except OSError:
            continue
    pytest.skip("HIP runtime library not found")


class DevBuf:
    def __init__(self, H, nbytes):
        self.H = H
        self.p = ct.c_void_p()
        assert H.hipMalloc(ct.byref(self.p), ct.c_size_t(max(nbytes, 1))) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.H.hipMemcpy(self.p, a.ctypes.data_as(ct.c_void_p), ct.c_size_t(a.nbytes), 1) == 0  # H2D

    def get(self, dtype, n):
        out = np.empty(n, dtype=dtype)
        assert self.H.hipMemcpy(out.ctypes.data_as(ct.c_void_p), self.p, ct.c_size_t(out.nbytes), 2) == 0  # D2H
        return out

    def free(self):
        self.H.hipFree(self.p)


def noalloc_fn(P):
    L = P.lib()
    f = L.psba_generate_idxs_device_noalloc
    vp, ip = ct.c_void_p, ct.POINTER(ct.c_int)
    f.argtypes = [ct.c_int, vp, ct.c_int, ct.c_int, ct.c_int, vp, vp, vp, vp, vp]
    f.restype = ct.c_int
    return f


def reference(P, vm, nc, npt):
    L = P.lib()
    no = int(np.count_nonzero(vm))
    ii = np.zeros(no, np.int32)
    jj = np.zeros(no, np.int32)
    bk = np.zeros(nc * npt, np.int32)
    ip = ct.POINTER(ct.c_int)
    rc = L.psba_generate_idxs_device(0, nc, npt, no, vm.tobytes(), None, ii.ctypes.data_as(ip),
                                     jj.ctypes.data_as(ip), bk.ctypes.data_as(ip))
    assert rc == 0
    return no, ii, jj, bk


def test_noalloc_matches_allocating_entry_point(P, gpu):
    H = hip()
    f = noalloc_fn(P)
    rng = np.random.default_rng(9)
    # the second call is smaller: stale scratch from the first must not leak
    for nc, npt, dens in [(88, 64298, 0.07), (7, 1000, 0.5)]:
        vm = (rng.random((npt, nc)) < dens).astype(np.int8)
        vm[np.arange(npt), rng.integers(0, nc, npt)] = 1
        no, ii, jj, bk = reference(P, vm, nc, npt)
        dvm = DevBuf(H, vm.nbytes)
        dii, djj = DevBuf(H, 4 * no), DevBuf(H, 4 * no)
        dbk, dcnt = DevBuf(H, 4 * nc * npt), DevBuf(H, 4)
        try:
            dvm.put(vm.reshape(-1))
            rc = f(0, None, nc, npt, no, dvm.p, dii.p, djj.p, dbk.p, dcnt.p)
            assert rc == 0, P.lib().psba_last_error()
            assert H.hipDeviceSynchronize() == 0
            assert int(dcnt.get(np.int32, 1)[0]) == no
            assert np.array_equal(dii.get(np.int32, no), ii)
            assert np.array_equal(djj.get(np.int32, no), jj)
            assert np.array_equal(dbk.get(np.int32, nc * npt), bk)
        finally:
            for b in (dvm, dii, djj, dbk, dcnt):
                b.free()


def test_noalloc_rejects_oversized_inputs(P):
    f = noalloc_fn(P)
    # 1024 * 65537 mask entries: one workgroup count more than the scratch holds
    dummy = ct.c_void_p(1)
    rc = f(0, None, 1024, 65537, 10, dummy, dummy, dummy, dummy, None)
    assert rc == -1
    assert b"exceeds the limit" in P.lib().psba_last_error()
