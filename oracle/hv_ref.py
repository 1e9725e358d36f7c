"""ctypes wrapper of oracle/hv_direct.c -- TEST INFRASTRUCTURE / CPU BASELINE ONLY.

Converts dense sector blocks to the reference's row-list sparse form (exact
zeros dropped, as sp_filter stores only nonzeros) and runs the scalar C
restatement of spMatVec_direct_main (src/DMRG_SUPERBLOCK_SETUP.f90:740-865).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hv_direct.c")
LIB = os.path.join(HERE, "_build", "libhvref.so")


def build() -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O3", "-fPIC", "-shared", "-o", tmp, SRC], check=True)
    os.replace(tmp, LIB)
    return LIB


class _Csr(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("cols", C.POINTER(C.c_int)), ("vals", C.POINTER(C.c_double))]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.hv_direct_csr.restype = None
        _lib = L
    return _lib


def _to_csr(m: np.ndarray, keep: list):
    m = np.asarray(m, dtype=np.float64)
    nz = m != 0.0
    rowptr = np.concatenate([[0], np.cumsum(nz.sum(axis=1))]).astype(np.int32)
    rr, cc = np.nonzero(nz)
    cols = cc.astype(np.int32)
    vals = m[rr, cc].astype(np.float64)
    keep += [rowptr, cols, vals]
    s = _Csr()
    s.nrow, s.ncol = m.shape
    s.rowptr = rowptr.ctypes.data_as(C.POINTER(C.c_int))
    s.cols = cols.ctypes.data_as(C.POINTER(C.c_int))
    s.vals = vals.ctypes.data_as(C.POINTER(C.c_double))
    return s


class CsrSuperblock:
    """Superblock in the reference's sparse-row representation."""

    def __init__(self, Dl, Dr, HL, HR, dst, src, A, B):
        self.keep = []
        self.nq = len(Dl)
        self.Dl = np.ascontiguousarray(Dl, dtype=np.int32)
        self.Dr = np.ascontiguousarray(Dr, dtype=np.int32)
        self.off = np.concatenate([[0], np.cumsum(self.Dl.astype(np.int64) * self.Dr)[:-1]]).astype(np.int64)
        self.n = int((self.Dl.astype(np.int64) * self.Dr).sum())
        self.HL = (_Csr * self.nq)(*[_to_csr(h, self.keep) for h in HL])
        self.HR = (_Csr * self.nq)(*[_to_csr(h, self.keep) for h in HR])
        nt = len(dst)
        self.nt = nt
        self.dst = np.ascontiguousarray(dst, dtype=np.int32)
        self.src = np.ascontiguousarray(src, dtype=np.int32)
        self.A = (_Csr * max(1, nt))(*[_to_csr(a, self.keep) for a in A])
        self.B = (_Csr * max(1, nt))(*[_to_csr(b, self.keep) for b in B])

    def hv(self, v: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        if out is None:
            out = np.empty_like(v)
        L = _load()
        L.hv_direct_csr(C.c_int(self.nq), self.Dl.ctypes.data_as(C.POINTER(C.c_int)),
                        self.Dr.ctypes.data_as(C.POINTER(C.c_int)), self.off.ctypes.data_as(C.POINTER(C.c_long)),
                        self.HL, self.HR, C.c_int(self.nt), self.dst.ctypes.data_as(C.POINTER(C.c_int)),
                        self.src.ctypes.data_as(C.POINTER(C.c_int)), self.A, self.B,
                        v.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)),
                        C.c_long(self.n))
        return out


def from_oracle(sb) -> CsrSuperblock:
    return CsrSuperblock(sb.Dl, sb.Dr, sb.HL, sb.HR, [t[0] for t in sb.terms], [t[1] for t in sb.terms],
                         [t[2] for t in sb.terms], [t[3] for t in sb.terms])
