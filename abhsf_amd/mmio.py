"""Matrix Market coordinate input with the reference's semantics.

include/abhsf/utils/matrix_market_reader.h:41-200: banner and size line of
the NIST format; 1-based indices become 0-based; entries stored with value
0.0 are dropped (read_mtx_real_or_binary, :179-187); pattern matrices get the
value 1.0; for symmetric / skew-symmetric / Hermitian matrices only the stored
triangle is kept and mixing both triangles is an error (:191-197).
`expand_symmetric` mirrors the stored triangle into a general matrix for
y = A*x over all entries (the reference counts those as nnz_all,
spmv/unified/spmv.cpp:1160-1166); without it, build symmetric storage
(AbhsfMatrix.from_elements with the returned props).  read_mm / write_mm: the
binary layout of utils/mtx2mm/mtx2mm.cpp:101-136 (cached conversions).
"""
from __future__ import annotations

import os

import numpy as np

from .matrix import matrix_properties, matrix_symmetry_t, matrix_type_t
from .synth import Csr, from_triplets


class MatrixMarketError(RuntimeError):
    pass


def read(path: str, expand_symmetric: bool = False):
    """Returns (Csr, matrix_properties, nnz_all)."""
    with open(path) as f:
        banner = f.readline().strip().split()
        if len(banner) < 5 or banner[0] != "%%MatrixMarket" or banner[1].lower() != "matrix":
            raise MatrixMarketError("Could not process Matrix Market banner.")
        fmt, field, sym = banner[2].lower(), banner[3].lower(), banner[4].lower()
        if fmt != "coordinate":
            raise MatrixMarketError("Only sparse matrices in coordinate format are supported.")
        types = {"pattern": matrix_type_t.BINARY, "integer": matrix_type_t.INTEGER, "real": matrix_type_t.REAL,
                 "complex": matrix_type_t.COMPLEX}
        syms = {"general": matrix_symmetry_t.UNSYMMETRIC, "symmetric": matrix_symmetry_t.SYMMETRIC,
                "skew-symmetric": matrix_symmetry_t.SKEW_SYMMETRIC, "hermitian": matrix_symmetry_t.HERMITIAN}
        if field not in types:
            raise MatrixMarketError("Unsupported matrix type.")
        if sym not in syms:
            raise MatrixMarketError("Unsupported type of matrix symmetry.")
        if field not in ("real", "pattern"):  # matrix_market_reader.h:155-156
            raise MatrixMarketError("read_matrix_real_or_binary() invoked for non-real-non-binary matrix.")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        try:
            m, n, nnz = (int(t) for t in line.split())
        except ValueError as e:
            raise MatrixMarketError("Could not retrieve matrix sizes.") from e
        ncol = 2 if field == "pattern" else 3
        data = np.loadtxt(f, ndmin=2, max_rows=nnz) if nnz else np.zeros((0, ncol))
    if data.shape[0] != nnz or (nnz and data.shape[1] < ncol):
        raise MatrixMarketError("Could not read matrix element.")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else data[:, 2].astype(np.float64)
    if nnz and (rows.min() < 0 or cols.min() < 0 or rows.max() >= m or cols.max() >= n):
        raise MatrixMarketError("Matrix element index out of range.")
    return _finish(m, n, rows, cols, vals, types[field], syms[sym], expand_symmetric, path)


def _finish(m, n, rows, cols, vals, mtype, msym, expand_symmetric, name):
    """Zeros dropped, one stored triangle for symmetric kinds (the triangle
    test sees every stored element, zeros included:
    matrix_market_reader.h:163-197)."""
    if msym != matrix_symmetry_t.UNSYMMETRIC:
        if np.any(rows > cols) and np.any(cols > rows):
            raise MatrixMarketError("Elements from both L and U parts stored for not-unsymmetric matrix.")
    if mtype == matrix_type_t.REAL:
        keep = vals != 0.0
        rows, cols, vals = rows[keep], cols[keep], vals[keep]
    props = matrix_properties(m, n, int(rows.size), mtype, msym)
    off = rows != cols
    nnz_all = int(rows.size + (off.sum() if props.symmetry != matrix_symmetry_t.UNSYMMETRIC else 0))
    if expand_symmetric and props.symmetry != matrix_symmetry_t.UNSYMMETRIC:
        sign = -1.0 if props.symmetry == matrix_symmetry_t.SKEW_SYMMETRIC else 1.0
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, sign * vals[off]]))
    return from_triplets(m, n, rows, cols, vals, name), props, nnz_all


def write(path: str, A: Csr, field: str = "real") -> None:
    r, c, v = A.triplets()
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} general\n{A.m} {A.n} {A.nnz}\n")
        if field == "pattern":
            for a, b in zip(r, c):
                f.write(f"{a + 1} {b + 1}\n")
        else:
            for a, b, x in zip(r, c, v):
                f.write(f"{a + 1} {b + 1} {float(x)!r}\n")


# ---- binary .mm layout of utils/mtx2mm/mtx2mm.cpp:101-136 --------------------
# u64 m, n, nnz, type (matrix_type_t: BINARY 0, INTEGER 1, REAL 2, COMPLEX 3),
# symmetry (matrix_symmetry_t: SYMMETRIC 0, HERMITIAN 1, SKEW_SYMMETRIC 2,
# UNSYMMETRIC 3); u32 rows[nnz], u32 cols[nnz] (0-based, as read); values:
# i64[nnz] (INTEGER), f64[nnz] (REAL), f64 re[nnz] + f64 im[nnz] (COMPLEX),
# none (BINARY).  Little-endian, no padding.

def write_mm(path: str, A: Csr, props: matrix_properties = None) -> None:
    """Write a real (or, with props.type BINARY, pattern) matrix in the
    binary layout; A holds the elements as stored (one triangle for
    symmetric props)."""
    props = props or matrix_properties(A.m, A.n, A.nnz)
    r, c, v = A.triplets()
    with open(path, "wb") as f:
        np.array([A.m, A.n, A.nnz, int(props.type), int(props.symmetry)], dtype="<u8").tofile(f)
        r.astype("<u4").tofile(f)
        c.astype("<u4").tofile(f)
        if props.type == matrix_type_t.REAL:
            np.asarray(v, dtype="<f8").tofile(f)
        elif props.type != matrix_type_t.BINARY:
            raise MatrixMarketError("write_mm writes real or pattern matrices")


def read_mm(path: str, expand_symmetric: bool = False):
    """Read the binary layout with read()'s semantics (real or pattern only,
    zeros dropped, one triangle for symmetric kinds); returns (Csr,
    matrix_properties, nnz_all)."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype="<u8", count=5)
        if hdr.size != 5:
            raise MatrixMarketError("binary matrix header truncated")
        m, n, nnz, mtype, msym = (int(x) for x in hdr)
        if mtype > 3 or msym > 3:
            raise MatrixMarketError("binary matrix header holds an unknown type or symmetry")
        mtype, msym = matrix_type_t(mtype), matrix_symmetry_t(msym)
        if mtype not in (matrix_type_t.REAL, matrix_type_t.BINARY):
            raise MatrixMarketError("read_matrix_real_or_binary() invoked for non-real-non-binary matrix.")
        want = 40 + 8 * nnz + (8 * nnz if mtype == matrix_type_t.REAL else 0)
        if size != want:
            raise MatrixMarketError(f"binary matrix size {size} differs from the header's {want}")
        rows = np.fromfile(f, dtype="<u4", count=nnz).astype(np.int64)
        cols = np.fromfile(f, dtype="<u4", count=nnz).astype(np.int64)
        vals = np.fromfile(f, dtype="<f8", count=nnz) if mtype == matrix_type_t.REAL else np.ones(nnz)
    if nnz and (rows.max() >= m or cols.max() >= n):
        raise MatrixMarketError("Matrix element index out of range.")
    return _finish(m, n, rows, cols, vals, mtype, msym, expand_symmetric, path)
