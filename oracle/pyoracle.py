"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

COO, CSR, BITMAP, DENSE = 0, 1, 2, 3
PREPROC, MSMF = 0, 1


class OrcAbhsf(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64),
        ("s", C.c_int32), ("k", C.c_int32),
        ("M", C.c_int64), ("N", C.c_int64),
        ("policy", C.c_int32), ("is_binary", C.c_int32), ("val_bytes", C.c_int32),
        ("cost_b", C.c_uint64),
        ("bcol_bytes", C.c_int32), ("coo_idx_bytes", C.c_int32),
        ("csr_rp_bytes", C.c_int32), ("bm_nw", C.c_int32),
        ("nnz", C.c_uint64),
        ("nb", C.c_uint64 * 4), ("ne", C.c_uint64 * 4),
        ("sptr", C.c_void_p * 4), ("eptr", C.c_void_p * 4),
        ("bcol", C.c_void_p * 4), ("z1", C.c_void_p * 4),
        ("coo_idx", C.c_void_p), ("csr_rowptr", C.c_void_p),
        ("csr_col", C.c_void_p), ("bm_words", C.c_void_p),
        ("val", C.c_void_p * 4),
        ("smf_bits", C.c_uint64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE])
        L = C.CDLL(LIB_PATH)
        u64, i64, i32, vp, d = C.c_uint64, C.c_int64, C.c_int32, C.c_void_p, C.c_double
        L.orc_ceil_log2.restype = u64
        L.orc_ceil_log2.argtypes = [u64]
        L.orc_choose.restype = C.c_int
        L.orc_choose.argtypes = [C.c_int, u64, u64, u64, C.c_int, C.POINTER(u64)]
        L.orc_smf_preproc.restype = u64
        L.orc_smf_preproc.argtypes = [u64, u64, u64, vp, vp, u64, u64, C.c_int, vp]
        L.orc_block_hist.restype = u64
        L.orc_block_hist.argtypes = [u64, vp, vp, C.c_int, C.c_int, vp, vp, u64]
        L.orc_mmf_row.restype = None
        L.orc_mmf_row.argtypes = [u64, u64, u64, C.c_int, u64, C.c_int, u64, u64, u64, vp, vp, vp]
        L.orc_build.restype = C.c_int
        L.orc_build.argtypes = [C.POINTER(OrcAbhsf), i64, i64, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.POINTER(OrcAbhsf)]
        L.orc_spmm.restype = None
        L.orc_spmm.argtypes = [C.POINTER(OrcAbhsf), C.c_int, d, vp, i64, d, vp, i64]
        L.orc_csr_spmv.restype = None
        L.orc_csr_spmv.argtypes = [i64, vp, vp, vp, d, vp, d, vp]
        L.orc_decode.restype = u64
        L.orc_decode.argtypes = [C.POINTER(OrcAbhsf), vp, vp, vp, u64]
        L.orc_varint_encode.restype = u64
        L.orc_varint_encode.argtypes = [u64, vp, vp]
        L.orc_varint_decode.restype = u64
        L.orc_varint_decode.argtypes = [u64, u64, vp, vp]
        L.orc_delta32_encode.restype = u64
        L.orc_delta32_encode.argtypes = [C.c_int, C.c_int64, vp, vp, u64, vp, vp, vp, vp, vp]
        L.orc_delta32_decode.restype = None
        L.orc_delta32_decode.argtypes = [C.c_int, u64, u64, vp, vp, vp, vp, vp, vp]
        L.orc_rows_stdev.restype = d
        L.orc_rows_stdev.argtypes = [u64, vp]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def ceil_log2(x: int) -> int:
    return int(lib().orc_ceil_log2(x))


def choose(policy: int, s: int, z: int, b: int = 64, is_binary: bool = False):
    bits = C.c_uint64(0)
    sc = lib().orc_choose(policy, s, z, b, int(is_binary), C.byref(bits))
    return int(sc), int(bits.value)


def smf_preproc(m, n, rows, cols, s, b=64, is_binary=False):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    nb = np.zeros(4, dtype=np.uint64)
    bits = lib().orc_smf_preproc(m, n, rows.size, _p(rows), _p(cols), s, b, int(is_binary), _p(nb))
    return int(bits), [int(v) for v in nb]


def block_hist(rows, cols, kr, ks):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    cap = 1 << 17
    z = np.zeros(cap, dtype=np.uint64)
    c = np.zeros(cap, dtype=np.uint64)
    nd = lib().orc_block_hist(rows.size, _p(rows), _p(cols), kr, ks, _p(z), _p(c), cap)
    assert nd <= cap
    return {int(a): int(b) for a, b in zip(z[:nd], c[:nd])}


def mmf_row(m, n, nnz, r, s, hist, is_binary=False, bits_single=32, mult=2):
    hz = np.array(sorted(hist), dtype=np.uint64)
    hc = np.array([hist[z] for z in sorted(hist)], dtype=np.uint64)
    out = np.zeros(8, dtype=np.uint64)
    lib().orc_mmf_row(m, n, nnz, int(is_binary), bits_single, mult, r, s, hz.size, _p(hz), _p(hc), _p(out))
    return [int(v) for v in out]


def rows_stdev(rows_nnz):
    a = np.ascontiguousarray(rows_nnz, dtype=np.uint64)
    return float(lib().orc_rows_stdev(a.size, _p(a)))


def varint_encode(keys):
    """Reference tagged varint delta stream of ascending keys (bytes after
    the first key): spmv/unified/spmv.cpp:859-896."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    n = lib().orc_varint_encode(keys.size, _p(keys), None)
    out = np.zeros(n, dtype=np.uint8)
    lib().orc_varint_encode(keys.size, _p(keys), _p(out))
    return out


def varint_decode(stream, first, nnz):
    stream = np.ascontiguousarray(stream, dtype=np.uint8)
    keys = np.zeros(nnz, dtype=np.uint64)
    used = lib().orc_varint_decode(nnz, int(first), _p(stream), _p(keys))
    return keys, int(used)


DELTA32_ROWMAJOR, DELTA32_CACHEBLOCK = 1, 2
DELTA_CHUNK = 512


def delta32_encode(layout, m, rowptr, col, chunk=DELTA_CHUNK):
    """u32 delta stream of the keys of layout 1 (row << 31 | col,
    spmv/orig/delta.cpp:124-151) or 2 (8x8 cache blocks,
    spmv/unified/spmv.cpp:636-689); returns dict da, perm, base, xstart, exc."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nnz = int(rowptr[m])
    nch = (nnz + chunk - 1) // chunk
    da = np.zeros(max(nnz, 1), dtype=np.uint32)
    perm = np.zeros(max(nnz, 1), dtype=np.uint64)
    base = np.zeros(max(nch, 1), dtype=np.uint64)
    xstart = np.zeros(nch + 1, dtype=np.uint32)
    exc = np.zeros(max(nnz, 1), dtype=np.uint64)
    nx = lib().orc_delta32_encode(layout, m, _p(rowptr), _p(col), chunk, _p(da), _p(perm), _p(base), _p(xstart),
                                  _p(exc))
    return {"da": da[:nnz], "perm": perm[:nnz], "base": base[:nch], "xstart": xstart, "exc": exc[:nx]}


def delta32_decode(layout, enc, chunk=DELTA_CHUNK):
    nnz = enc["da"].size
    rows = np.zeros(max(nnz, 1), dtype=np.int64)
    cols = np.zeros(max(nnz, 1), dtype=np.int64)
    arrs = [np.ascontiguousarray(enc[k], dtype=t) for k, t in
            (("da", np.uint32), ("base", np.uint64), ("xstart", np.uint32), ("exc", np.uint64))]
    arrs = [a if a.size else np.zeros(1, dtype=a.dtype) for a in arrs]
    lib().orc_delta32_decode(layout, nnz, chunk, *(_p(a) for a in arrs), _p(rows), _p(cols))
    return rows[:nnz], cols[:nnz]


def csr_spmv(rowptr, col, val, x, alpha=1.0, beta=0.0, y=None):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    m = rowptr.size - 1
    y = np.zeros(m) if y is None else np.array(y, dtype=np.float64, copy=True)
    lib().orc_csr_spmv(m, _p(rowptr), _p(col), _p(val), alpha, _p(x), beta, _p(y))
    return y


_U = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


class Abhsf:
    """CPU-built ABHSF storage; arrays are copied out as numpy arrays."""

    def __init__(self, m, n, rowptr, col, val, k, policy=PREPROC, val_bytes=8, is_binary=False):
        self._keep = (np.ascontiguousarray(rowptr, dtype=np.int64),
                      np.ascontiguousarray(col, dtype=np.int32),
                      np.ascontiguousarray(val, dtype=np.float64))
        self.st = OrcAbhsf()
        rc = lib().orc_build(C.byref(self.st), m, n, _p(self._keep[0]), _p(self._keep[1]),
                             _p(self._keep[2]), k, policy, val_bytes, int(is_binary))
        if rc != 0:
            raise ValueError("orc_build rejected the input")

    def __del__(self):
        try:
            lib().orc_free(C.byref(self.st))
        except Exception:
            pass

    def _arr(self, ptr, n, dt):
        if n == 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).copy()

    def arrays(self):
        """All storage arrays by name (names match abhsf_export in include/abhsf_c.h)."""
        st = self.st
        M = st.M
        vdt = np.float64 if st.val_bytes == 8 else np.float32
        out = {}
        names = ["coo", "csr", "bm", "dn"]
        for x in range(4):
            nb, ne = st.nb[x], st.ne[x]
            out[f"{names[x]}_sptr"] = self._arr(st.sptr[x], M + 1, np.uint32)
            out[f"{names[x]}_eptr"] = self._arr(st.eptr[x], M + 1, np.uint32)
            out[f"{names[x]}_bcol"] = self._arr(st.bcol[x], nb, _U[st.bcol_bytes])
            out[f"{names[x]}_z1"] = self._arr(st.z1[x], nb, np.uint16)
            out[f"{names[x]}_val"] = self._arr(st.val[x], ne, vdt)
        out["coo_idx"] = self._arr(st.coo_idx, st.ne[0], _U[st.coo_idx_bytes])
        out["csr_rowptr"] = self._arr(st.csr_rowptr, st.nb[1] * (st.s + 1), _U[st.csr_rp_bytes])
        out["csr_col"] = self._arr(st.csr_col, st.ne[1], np.uint8)
        out["bm_words"] = self._arr(st.bm_words, st.nb[2] * st.bm_nw, np.uint64)
        return out

    def spmm(self, X, alpha=1.0, beta=0.0, Y=None):
        X = np.ascontiguousarray(X, dtype=np.float64)
        nrhs = 1 if X.ndim == 1 else X.shape[1]
        shape = (self.st.m,) if X.ndim == 1 else (self.st.m, nrhs)
        Y = np.zeros(shape) if Y is None else np.array(Y, dtype=np.float64, copy=True)
        lib().orc_spmm(C.byref(self.st), nrhs, alpha, _p(X), nrhs, beta, _p(Y), nrhs)
        return Y

    def decode(self):
        cap = int(self.st.nnz + sum(self.st.ne))
        r = np.zeros(cap, dtype=np.int64)
        c = np.zeros(cap, dtype=np.int64)
        v = np.zeros(cap, dtype=np.float64)
        n = lib().orc_decode(C.byref(self.st), _p(r), _p(c), _p(v), cap)
        return r[:n], c[:n], v[:n]

    @property
    def smf_bits(self):
        return int(self.st.smf_bits)

    @property
    def nb(self):
        return [int(v) for v in self.st.nb]
