"""Host-side mirror of the reference's matrix-class interface for this path.

The reference's kernel interface (spmv/unified/spmv.cpp:71-242) is a class with
  from_elements(elements, props)  -- build from (row, col, val) triplets
  spmv(x, y)                      -- y += A*x (accumulating), synchronous
  release()                       -- free storage
AbhsfMatrix keeps those names and meanings on top of the C ABI, plus the
extensions of this library (alpha/beta SpMV, multi-RHS SpMM, statistics and
array export).  x and y are device arrays (hip.DeviceArray or raw addresses),
as for the reference's GPU variant spmv_cusparse_hyb (spmv/unified/spmv.cpp:199-207).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import capi, hip


class matrix_type_t(enum.IntEnum):  # include/abhsf/utils/matrix_properties.h:8-14
    BINARY = 0
    INTEGER = 1
    REAL = 2
    COMPLEX = 3


class matrix_symmetry_t(enum.IntEnum):  # include/abhsf/utils/matrix_properties.h:16-22
    SYMMETRIC = 0
    HERMITIAN = 1
    SKEW_SYMMETRIC = 2
    UNSYMMETRIC = 3


@dataclass
class matrix_properties:  # include/abhsf/utils/matrix_properties.h:24-71
    m: int = 0
    n: int = 0
    nnz: int = 0
    type: matrix_type_t = matrix_type_t.REAL
    symmetry: matrix_symmetry_t = matrix_symmetry_t.UNSYMMETRIC


_SYM = {matrix_symmetry_t.UNSYMMETRIC: capi.GENERAL, matrix_symmetry_t.SYMMETRIC: capi.SYMMETRIC,
        matrix_symmetry_t.HERMITIAN: capi.SYMMETRIC, matrix_symmetry_t.SKEW_SYMMETRIC: capi.SKEW_SYMMETRIC,
        "general": capi.GENERAL, "symmetric": capi.SYMMETRIC, "skew": capi.SKEW_SYMMETRIC}


def _addr(a) -> int:
    if isinstance(a, hip.DeviceArray):
        return a.address
    return int(a)


_POLICY = {"preproc": capi.POLICY_PREPROC, "msmf": capi.POLICY_MSMF}
_PREC = {"f64": capi.F64, "f32": capi.F32}
_VAR = {"auto": capi.VARIANT_AUTO, "strip": capi.VARIANT_STRIP, "balanced": capi.VARIANT_BALANCED,
        "panel": capi.VARIANT_PANEL, "tiled": capi.VARIANT_TILED}


class AbhsfMatrix:
    """One ABHSF matrix resident in HBM."""

    def __init__(self, block_size: int = 64, policy: str = "preproc", precision: str = "f64",
                 variant: str = "auto", stream=None, is_binary: bool = False):
        self.opts = capi.Options(block_size, _POLICY[policy], _PREC[precision], _VAR[variant],
                                 int(is_binary), 0,
                                 stream.handle if isinstance(stream, hip.Stream) else stream)
        self._h = C.c_void_p(0)
        capi.call("abhsf_create", C.byref(self.opts), C.byref(self._h))
        self.props = None

    # -- build ---------------------------------------------------------------
    def from_elements(self, elements, props: matrix_properties) -> None:
        """Build from triplets (rows, cols, vals) arrays or a sequence of
        (row, col, val) tuples.  Unlike spmv/unified/spmv.cpp:77 the caller's
        data is not reordered (a private copy is sorted)."""
        if isinstance(elements, tuple) and len(elements) == 3 and hasattr(elements[0], "__len__"):
            rows, cols, vals = elements
        else:
            arr = np.asarray(elements, dtype=np.float64).reshape(-1, 3)
            rows, cols, vals = arr[:, 0], arr[:, 1], arr[:, 2]
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = None if vals is None else np.ascontiguousarray(vals, dtype=np.float64)
        # symmetric / skew-symmetric: one stored triangle, SpMV applies both
        capi.call("abhsf_set_symmetry", self._h, _SYM[props.symmetry])
        capi.call("abhsf_build_coo", self._h, props.m, props.n, rows.size,
                  rows.ctypes.data_as(C.c_void_p), cols.ctypes.data_as(C.c_void_p),
                  None if vals is None else vals.ctypes.data_as(C.c_void_p))
        self.props = props

    def from_csr(self, m: int, n: int, rowptr, col, val, on_device: bool = False, nnz: int = None,
                 symmetry="general") -> None:
        """Build from CSR (host numpy arrays, or device arrays with on_device).
        symmetry "symmetric" / "skew" (or a matrix_symmetry_t): the CSR holds
        one triangle and SpMV applies the whole matrix."""
        capi.call("abhsf_set_symmetry", self._h, _SYM[symmetry])
        if on_device:
            if nnz is None:
                nnz = col.shape[0]
            capi.call("abhsf_build_csr", self._h, m, n, nnz, _addr(rowptr), _addr(col),
                      None if val is None else _addr(val), 1)
        else:
            rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
            col = np.ascontiguousarray(col, dtype=np.int32)
            val = None if val is None else np.ascontiguousarray(val, dtype=np.float64)
            capi.call("abhsf_build_csr", self._h, m, n, int(rowptr[-1]),
                      rowptr.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                      None if val is None else val.ctypes.data_as(C.c_void_p), 0)
        sym = {capi.GENERAL: matrix_symmetry_t.UNSYMMETRIC, capi.SYMMETRIC: matrix_symmetry_t.SYMMETRIC,
               capi.SKEW_SYMMETRIC: matrix_symmetry_t.SKEW_SYMMETRIC}[_SYM[symmetry]]
        self.props = matrix_properties(m, n, self.stats()["nnz"], symmetry=sym)

    # -- apply ---------------------------------------------------------------
    def spmv(self, x, y) -> None:
        """y += A*x and wait (spmv/unified/spmv.cpp:140-159, 199-207)."""
        capi.call("abhsf_spmv", self._h, 1.0, _addr(x), 1.0, _addr(y))
        capi.call("abhsf_synchronize", self._h)

    def spmv_async(self, alpha: float, x, beta: float, y) -> None:
        """y = alpha*A*x + beta*y, enqueued on the handle's stream."""
        capi.call("abhsf_spmv", self._h, float(alpha), _addr(x), float(beta), _addr(y))

    def spmm_async(self, nrhs: int, alpha: float, X, ldx: int, beta: float, Y, ldy: int) -> None:
        capi.call("abhsf_spmm", self._h, int(nrhs), float(alpha), _addr(X), int(ldx), float(beta),
                  _addr(Y), int(ldy))

    def synchronize(self) -> None:
        capi.call("abhsf_synchronize", self._h)

    # -- inspection ------------------------------------------------------------
    def stats(self) -> dict:
        st = capi.Stats()
        capi.call("abhsf_get_stats", self._h, C.byref(st))
        return st.as_dict()

    def export(self, name: str, dtype) -> np.ndarray:
        nbytes = C.c_size_t(0)
        capi.call("abhsf_export", self._h, name.encode(), None, 0, C.byref(nbytes))
        out = np.empty(nbytes.value // np.dtype(dtype).itemsize, dtype=dtype)
        if nbytes.value:
            capi.call("abhsf_export", self._h, name.encode(), out.ctypes.data_as(C.c_void_p),
                      nbytes.value, C.byref(nbytes))
        return out

    def export_all(self) -> dict:
        """All storage arrays with their stored element types."""
        st = self.stats()
        s, N = st["s"], st["N"]
        bdt = np.uint16 if N <= 65536 else np.uint32
        vdt = np.float64 if self.opts.precision == capi.F64 else np.float32
        out = {}
        for p in ("coo", "csr", "bm", "dn"):
            out[f"{p}_sptr"] = self.export(f"{p}_sptr", np.uint32)
            out[f"{p}_eptr"] = self.export(f"{p}_eptr", np.uint32)
            out[f"{p}_bcol"] = self.export(f"{p}_bcol", bdt)
            out[f"{p}_z1"] = self.export(f"{p}_z1", np.uint16)
            out[f"{p}_val"] = self.export(f"{p}_val", vdt)
        out["coo_idx"] = self.export("coo_idx", np.uint8 if s <= 16 else np.uint16)
        out["csr_rowptr"] = self.export("csr_rowptr", np.uint8 if s <= 8 else np.uint16)
        out["csr_col"] = self.export("csr_col", np.uint8)
        out["bm_words"] = self.export("bm_words", np.uint64)
        return out

    def save(self, path: str) -> None:
        """Serialised ABHSF blob: every storage array plus shape and options (npz)."""
        st = self.stats()
        arrays = self.export_all()
        meta = np.array([st["m"], st["n"], st["nnz"], self.opts.block_size, self.opts.policy,
                         self.opts.precision, self.opts.is_binary], dtype=np.int64)
        np.savez(path, __meta__=meta, **arrays)

    @classmethod
    def load(cls, path: str, variant: str = "auto", stream=None) -> "AbhsfMatrix":
        """Recreate a matrix from save() output without re-running the conversion."""
        with np.load(path, allow_pickle=False) as f:
            m, n, nnz, s, pol, prec, binary = (int(v) for v in f["__meta__"])
            names = [k for k in f.files if k != "__meta__"]
            arrays = [np.ascontiguousarray(f[k]) for k in names]
        inv_pol = {v: k for k, v in _POLICY.items()}
        inv_prec = {v: k for k, v in _PREC.items()}
        M = cls(block_size=s, policy=inv_pol[pol], precision=inv_prec[prec], variant=variant, stream=stream,
                is_binary=bool(binary))
        cn = (C.c_char_p * len(names))(*[nm.encode() for nm in names])
        cd = (C.c_void_p * len(names))(*[a.ctypes.data for a in arrays])
        cb = (C.c_size_t * len(names))(*[a.nbytes for a in arrays])
        capi.call("abhsf_import", M._h, m, n, nnz, len(names), cn, cd, cb)
        M.props = matrix_properties(m, n, nnz)
        return M

    # -- release ---------------------------------------------------------------
    def release(self) -> None:
        if self._h and self._h.value:
            capi.call("abhsf_destroy", self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class AbhsfBatch:
    """Many built matrices driven by one launch (optionally a replayed hipGraph).

    The members must stay alive and unchanged while the batch exists.
    """

    def __init__(self, mats, stream=None):
        self.mats = list(mats)
        arr = (C.c_void_p * len(self.mats))(*[m._h.value for m in self.mats])
        self._h = C.c_void_p(0)
        capi.call("abhsf_batch_create", arr, len(self.mats),
                  stream.handle if isinstance(stream, hip.Stream) else stream, C.byref(self._h))
        self._vec = None

    def set_vectors(self, xs, ys) -> None:
        xa = (C.c_void_p * len(xs))(*[_addr(v) for v in xs])
        ya = (C.c_void_p * len(ys))(*[_addr(v) for v in ys])
        capi.call("abhsf_batch_set_vectors", self._h, xa, ya)
        self._vec = (xs, ys)  # keep the device vectors alive

    def spmv_async(self, alpha: float = 1.0, beta: float = 0.0, graph: bool = False) -> None:
        capi.call("abhsf_batch_spmv", self._h, float(alpha), float(beta), int(graph))

    def synchronize(self) -> None:
        capi.call("abhsf_batch_synchronize", self._h)

    def profile_cycles(self, alpha: float = 1.0, beta: float = 0.0) -> np.ndarray:
        """One eager batched SpMV returning, per member, the shader-clock
        cycles its work items occupied (its share of the launch)."""
        out = np.zeros(len(self.mats), dtype=np.uint64)
        capi.call("abhsf_batch_profile", self._h, float(alpha), float(beta), out.ctypes.data_as(C.c_void_p))
        return out

    def release(self) -> None:
        if self._h and self._h.value:
            capi.call("abhsf_batch_destroy", self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


_CODEC = {"varint": capi.DELTA_VARINT, "u32_row": capi.DELTA_U32_ROW, "u32_block": capi.DELTA_U32_BLOCK}


class DeltaMatrix:
    """Delta-coded coordinate storage resident in HBM: the reference's
    delta_matrix family (spmv/unified/spmv.cpp:628-1033, spmv/orig/delta.cpp:
    118-160) laid out for wavefront decode (see include/abhsf_c.h).  codec:
    "varint" (tagged varint words of row-major keys), "u32_row" (u32 deltas of
    row << 31 | col) or "u32_block" (u32 deltas of 8 x 8 cache-block keys)."""

    def __init__(self, stream=None, codec: str = "varint"):
        self._stream = stream
        self.codec = codec
        self._h = C.c_void_p(0)
        capi.call("abhsf_delta_create_codec", stream.handle if isinstance(stream, hip.Stream) else stream,
                  _CODEC[codec], C.byref(self._h))

    def from_elements(self, elements, props: matrix_properties) -> None:
        """Build from (rows, cols, vals) arrays or (row, col, val) tuples; a
        private copy is sorted row-major (the reference sorts in place)."""
        if isinstance(elements, tuple) and len(elements) == 3 and hasattr(elements[0], "__len__"):
            rows, cols, vals = (np.asarray(a) for a in elements)
        else:
            arr = np.asarray(elements, dtype=np.float64).reshape(-1, 3)
            rows, cols, vals = arr[:, 0], arr[:, 1], arr[:, 2]
        order = np.lexsort((cols, rows))
        rows = rows[order].astype(np.int64)
        cols = cols[order].astype(np.int32)
        vals = np.ascontiguousarray(vals[order], dtype=np.float64)
        rowptr = np.zeros(props.m + 1, dtype=np.int64)
        np.add.at(rowptr, rows + 1, 1)
        self.from_csr(props.m, props.n, np.cumsum(rowptr), cols, vals)

    def from_csr(self, m: int, n: int, rowptr, col, val, on_device: bool = False, nnz: int = None) -> None:
        if on_device:
            capi.call("abhsf_delta_build_csr", self._h, m, n, col.shape[0] if nnz is None else nnz,
                      _addr(rowptr), _addr(col), None if val is None else _addr(val), 1)
            return
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = None if val is None else np.ascontiguousarray(val, dtype=np.float64)
        capi.call("abhsf_delta_build_csr", self._h, m, n, int(rowptr[-1]), rowptr.ctypes.data_as(C.c_void_p),
                  col.ctypes.data_as(C.c_void_p), None if val is None else val.ctypes.data_as(C.c_void_p), 0)

    def spmv(self, x, y) -> None:
        """y += A*x and wait (delta_matrix::spmv, spmv/unified/spmv.cpp:712-789)."""
        capi.call("abhsf_delta_spmv", self._h, 1.0, _addr(x), 1.0, _addr(y))
        capi.call("abhsf_delta_synchronize", self._h)

    def spmv_async(self, alpha: float, x, beta: float, y) -> None:
        capi.call("abhsf_delta_spmv", self._h, float(alpha), _addr(x), float(beta), _addr(y))

    def synchronize(self) -> None:
        capi.call("abhsf_delta_synchronize", self._h)

    def stats(self) -> dict:
        st = capi.DeltaStats()
        capi.call("abhsf_delta_get_stats", self._h, C.byref(st))
        return st.as_dict()

    def export(self, name: str) -> np.ndarray:
        dt = {"tags": np.uint32, "w8": np.uint8, "w16": np.uint16, "w32": np.uint32, "w64": np.uint64,
              "base": np.uint64, "start": np.uint32, "val": np.float64, "da": np.uint32, "xstart": np.uint32,
              "exc": np.uint64}[name]
        nbytes = C.c_size_t(0)
        capi.call("abhsf_delta_export", self._h, name.encode(), None, 0, C.byref(nbytes))
        out = np.empty(nbytes.value // np.dtype(dt).itemsize, dtype=dt)
        capi.call("abhsf_delta_export", self._h, name.encode(), out.ctypes.data_as(C.c_void_p), nbytes.value,
                  C.byref(nbytes))
        return out

    def release(self) -> None:
        if self._h and self._h.value:
            capi.call("abhsf_delta_destroy", self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
