"""The allocation-free batched entry point (cdft_converge_batch_device_noalloc)
against cdft_converge_batch_device, through the C ABI: two calls in a row, the second
with fewer systems, must give the same results and profiles bit for bit (same
kernel, same inputs; only where the descriptors and results live differs).  Needs a
GPU."""
import ctypes as C
import sys

import numpy as np
import pytest

from cdft_amd import DFTOptions, _lib, configs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceBuffer:
    """Device memory through the HIP runtime the library itself uses (the test
    allocates; the entry point under test does not)."""

    def __init__(self, host):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = host.size
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * self.n)) == 0
        self.put(host)

    def put(self, host):
        h = np.ascontiguousarray(host, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0   # H2D

    def get(self):
        h = np.empty(self.n)
        assert self.hip.hipMemcpy(h.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0   # D2H
        return h

    def free(self):
        self.hip.hipFree(self.ptr)


def solve(ctx, pores, noalloc):
    systems = [s.device_system for s, _ in pores]
    sizes = [s.N * s.nc for s in systems]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    buf = DeviceBuffer(np.concatenate([r.reshape(-1) for _, r in pores]))
    try:
        ptrs = [buf.ptr.value + 8 * int(o) for o in offs[:-1]]
        res = _lib.converge_batch(ctx, systems, ptrs, tol=1e-8, device_ptrs=True, noalloc=noalloc)
        return res, buf.get()
    finally:
        buf.free()


def test_noalloc_matches_allocating_entry_point():
    idx_a = [3, 9, 40, 77, 114, 200, 252, 269, 300, 343, 400, 480]      # 12 pores, a blow-up and a cap among them
    idx_b = [11, 13, 252, 426]                                           # the smaller second call
    mixed = DFTOptions(precision="mixed")
    pa = configs.slit_pores(n=512, seed=0, indices=idx_a, options=mixed)
    pb = configs.slit_pores(n=512, seed=0, indices=idx_b, options=mixed)
    ctx = pa[0][0].device_system.ctx
    for name, pores in (("first call, 12 pores", pa), ("second call, 4 pores", pb)):
        say(f"noalloc vs device entry point: {name}")
        r0, x0 = solve(ctx, pores, noalloc=True)
        r1, x1 = solve(ctx, pores, noalloc=False)
        for a, b in zip(r0, r1):
            assert a["errcode"] == b["errcode"] and a["nevals"] == b["nevals"] and a["iters"] == b["iters"]
            assert a["nevals_fp32"] == b["nevals_fp32"]
            assert (a["resnorm"] == b["resnorm"]) or (np.isnan(a["resnorm"]) and np.isnan(b["resnorm"]))
        assert np.array_equal(x0, x1, equal_nan=True)
        say(f"  errcodes {[r['errcode'] for r in r0]}, profiles identical")


def test_noalloc_rejects_oversized_and_unprepared_batches():
    """More than CDFT_NOALLOC_MAX_SYS systems, or a system the batched path cannot
    take, raise before any launch."""
    s, rho = configs.slit_pores(n=512, seed=0, indices=[5], options=DFTOptions(precision="mixed"))[0]
    dev = s.device_system
    n = _lib.NOALLOC_MAX_SYS + 1
    L = _lib.load()
    hs = (C.c_void_p * n)(*([dev.handle.value] * n))
    ps = (_lib._dp * n)(*([C.cast(C.c_void_p(1), _lib._dp)] * n))
    o = _lib.SolverOpts(10000, 1e-2, 1e-8, 1e-1, 5, 0, None)
    res = (_lib.Result * n)()
    rc = L.cdft_converge_batch_device_noalloc(dev.ctx.handle, n, hs, ps, C.byref(o), res)
    assert rc == -1 and b"at most" in L.cdft_last_error(dev.ctx.handle)
    wall, _ = configs.c1c10_interface(ngrid=128, box=6.0, options=DFTOptions(precision="mixed"))
    hs1 = (C.c_void_p * 1)(wall.device_system.handle.value)
    rc = L.cdft_converge_batch_device_noalloc(dev.ctx.handle, 1, hs1, ps, C.byref(o), res)
    assert rc < 0
