"""Matrix Market coordinate input with the reference's semantics.

include/abhsf/utils/matrix_market_reader.h:41-200: banner and size line of
the NIST format; 1-based indices become 0-based; entries stored with value
0.0 are dropped (read_mtx_real_or_binary, :179-187); pattern matrices get the
value 1.0; for symmetric / skew-symmetric / Hermitian matrices only the stored
triangle is kept and mixing both triangles is an error (:191-197).
`expand_symmetric` mirrors the stored triangle into a general matrix for
y = A*x over all entries (the reference counts those as nnz_all,
spmv/unified/spmv.cpp:1160-1166).
"""
from __future__ import annotations

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
        if field == "complex":
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
    keep = vals != 0.0
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    props = matrix_properties(m, n, int(rows.size), types[field], syms[sym])
    off = rows != cols
    if props.symmetry != matrix_symmetry_t.UNSYMMETRIC:
        if np.any(rows > cols) and np.any(cols > rows):
            raise MatrixMarketError("Elements from both L and U parts stored for not-unsymmetric matrix.")
    nnz_all = int(rows.size + (off.sum() if props.symmetry != matrix_symmetry_t.UNSYMMETRIC else 0))
    if expand_symmetric and props.symmetry != matrix_symmetry_t.UNSYMMETRIC:
        sign = -1.0 if props.symmetry == matrix_symmetry_t.SKEW_SYMMETRIC else 1.0
        rows, cols, vals = (np.concatenate([rows, cols[off]]), np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, sign * vals[off]]))
    return from_triplets(m, n, rows, cols, vals, path), props, nnz_all


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
