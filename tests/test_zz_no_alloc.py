"""The allocation-free H*v entry point (nss_sb_hv_device_noalloc) through the
C ABI: on two superblocks in a row, the second smaller, its output equals
nss_sb_hv_device's bit for bit (same kernels, same accumulation order) and
the oracle's direct product to 1e-12 relative.  Device vectors come from the
HIP runtime the library itself links (not torch, which ships a runtime of its
own); the library's stream is drained by a device-wide synchronisation."""
import ctypes as C
import os

import numpy as np
import pytest

import nssdmrg_amd as N
from nssdmrg_amd import _ffi as F
from oracle import dmrg_oracle as O

pytestmark = pytest.mark.gpu

SPIN_HIJ = [[1.0, 0.0], [0.0, 0.5]]
HUB_HIJ = [[-1.0, 0.0], [0.0, -1.0]]


class _Hip:
    """hipMalloc / hipMemcpy / hipDeviceSynchronize of the runtime that
    libnssdmrg.so is linked against (/opt/rocm/lib)."""

    def __init__(self):
        path = "/opt/rocm/lib/libamdhip64.so.7"
        if not os.path.exists(path):
            path = "/opt/rocm/lib/libamdhip64.so"
        self.h = C.CDLL(path)
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipSetDevice.argtypes = [C.c_int]

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), nbytes))
        return p

    def put(self, dst, a):  # hipMemcpyHostToDevice = 1
        self.ok(self.h.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1))

    def get(self, src, n):  # hipMemcpyDeviceToHost = 2
        a = np.empty(n, dtype=np.float64)
        self.ok(self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2))
        return a

    def sync(self):
        self.ok(self.h.hipDeviceSynchronize())


def _superblock(model, steps, M):
    if model == "spin1":
        dot, hij, tq = O.spin_site(3), SPIN_HIJ, (0.0,)
    else:
        dot, hij, tq = O.electron_site(4.0, True), HUB_HIJ, (0.5, 0.5)
    p = O.Params(ldmrg=steps + 2, mdmrg=M, target_qn=tq, hij=hij, neigen=1, lanc_dim_threshold=64)
    d = O.DMRG(dot, p)
    for _ in range(steps):
        d.step("i")
    left = O.enlarge_block(d.left, dot, "left", hij)
    right = O.enlarge_block(d.right, dot, "right", hij)
    sb = O.sb_get_states(left, right, O.current_target(left.length + right.length, tq))
    O.setup_direct(sb, left, right, hij)
    return sb


def test_noalloc_hv_equals_device_hv_on_two_calls():
    lib = F.lib()
    with N.Context(0) as ctx:
        hip = _Hip()
        hip.ok(hip.h.hipSetDevice(0))
        bufs = []
        # the second superblock is the smaller one
        for model, steps, M in (("hubbard", 4, 48), ("spin1", 3, 24)):
            sb = _superblock(model, steps, M)
            print(f"case noalloc {model} n={sb.n} nq={len(sb.qns)} terms={len(sb.terms)}", flush=True)
            ctx.set_layout(sb.Dl, sb.Dr, sb.offset)
            ctx.upload_blocks(sb.HL, sb.HR, [t[0] for t in sb.terms], [t[1] for t in sb.terms],
                              [t[2] for t in sb.terms], [t[3] for t in sb.terms])
            v = O.splitmix64_uniform_fast(sb.n, 5)
            x, y_ref, y_new = (hip.alloc(8 * sb.n) for _ in range(3))
            bufs += [x, y_ref, y_new]
            hip.put(x, v)
            nan = np.full(sb.n, np.nan)
            hip.put(y_ref, nan)
            hip.put(y_new, nan)
            F.check(lib.nss_sb_hv_device(ctx.handle, x, y_ref, sb.n))
            hip.sync()
            F.check(lib.nss_sb_hv_device_noalloc(ctx.handle, x, y_new, sb.n))
            hip.sync()
            a, b = hip.get(y_ref, sb.n), hip.get(y_new, sb.n)
            assert np.all(np.isfinite(b))
            assert np.array_equal(a, b)
            ref = O.hv_direct(sb, v)
            assert np.linalg.norm(b - ref) < 1e-12 * np.linalg.norm(ref)
        # a length that does not match the layout is refused
        rc = lib.nss_sb_hv_device_noalloc(ctx.handle, x, y_new, sb.n + 1)
        assert rc != 0 and "does not match" in F.lib().nss_last_error().decode()
        hip.sync()
        for p in bufs:
            hip.ok(hip.h.hipFree(p))
