"""Seeded synthetic matrices (CSR, host numpy) for tests and benchmarks.

Backed by libabhsf_synth.so (abhsf_amd/csrc/synth.c, OpenMP); every family is
a deterministic function of its parameters and seed.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _L():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libabhsf_synth.so")
        if not os.path.exists(path):
            raise OSError(f"{path} is missing: run __graft_entry__.build()")
        L = C.CDLL(path)
        i64, vp, d, u64 = C.c_int64, C.c_void_p, C.c_double, C.c_uint64
        L.synth_laplacian2d.argtypes = [i64, vp, vp, vp]
        L.synth_stencil27.argtypes = [i64, vp, vp, vp]
        L.synth_block_mix.argtypes = [i64, d, d, u64, vp, vp, vp]
        L.synth_banded.argtypes = [i64, i64, u64, vp, vp, vp]
        L.synth_random_blocks.argtypes = [i64, i64, i64, d, u64, vp, vp, vp]
        L.synth_rmat.argtypes = [C.c_int, i64, d, d, d, u64, vp, vp, vp]
        L.synth_uniform_blocks.argtypes = [i64, i64, i64, u64, vp, vp, vp]
        for f in ("synth_laplacian2d", "synth_stencil27", "synth_block_mix", "synth_banded",
                  "synth_random_blocks", "synth_rmat", "synth_uniform_blocks"):
            getattr(L, f).restype = i64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Csr:
    m: int
    n: int
    rowptr: np.ndarray  # int64[m+1]
    col: np.ndarray     # int32[nnz]
    val: np.ndarray     # float64[nnz]
    name: str = ""

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def checksum(self) -> int:
        """Order-sensitive 64-bit checksum of the structure (rows, cols)."""
        h = np.uint64(1469598103934665603)
        c = self.col.astype(np.uint64)
        r = np.repeat(np.arange(self.m, dtype=np.uint64), np.diff(self.rowptr).astype(np.int64))
        with np.errstate(over="ignore"):
            mix = (r * np.uint64(0x9E3779B97F4A7C15)) ^ (c * np.uint64(0xC2B2AE3D27D4EB4F))
            w = np.arange(1, c.size + 1, dtype=np.uint64)
            return int((h + np.sum(mix * w, dtype=np.uint64)) & np.uint64(0xFFFFFFFFFFFFFFFF))

    def triplets(self):
        r = np.repeat(np.arange(self.m, dtype=np.int64), np.diff(self.rowptr).astype(np.int64))
        return r, self.col.astype(np.int64), self.val


def _two_call(fn, m, *args):
    L = _L()
    rowptr = np.zeros(m + 1, dtype=np.int64)
    nnz = getattr(L, fn)(*args, _p(rowptr), None, None)
    if nnz < 0:
        raise MemoryError(f"{fn} failed")
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    got = getattr(L, fn)(*args, _p(rowptr), _p(col), _p(val))
    assert got == nnz
    return rowptr, col, val


def laplacian2d(g: int) -> Csr:
    """5-point Laplacian on a g x g grid: diagonal 4, off-diagonals -1."""
    n = g * g
    L = _L()
    nnz = L.synth_laplacian2d(g, None, None, None)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    L.synth_laplacian2d(g, _p(rowptr), _p(col), _p(val))
    return Csr(n, n, rowptr, col, val, f"laplacian2d_{g}")


def stencil27(g: int) -> Csr:
    """27-point stencil on a g^3 grid: diagonal 26, off-diagonals -1."""
    n = g * g * g
    L = _L()
    nnz = L.synth_stencil27(g, None, None, None)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    L.synth_stencil27(g, _p(rowptr), _p(col), _p(val))
    return Csr(n, n, rowptr, col, val, f"stencil27_{g}")


def block_mix(n: int, seed: int = 0, p_dense: float = 0.2, p_band: float = 0.3) -> Csr:
    """32-row strips: 20% three dense 32x32 blocks, 30% block-tridiagonal of
    8x8 sub-blocks, 50% 12 uniform columns per row; values N(0,1)."""
    rowptr, col, val = _two_call("synth_block_mix", n, n, p_dense, p_band, seed)
    return Csr(n, n, rowptr, col, val, f"block_mix_{n}_s{seed}")


def banded(n: int, bw: int, seed: int = 0) -> Csr:
    rowptr, col, val = _two_call("synth_banded", n, n, bw, seed)
    return Csr(n, n, rowptr, col, val, f"banded_{n}_{bw}")


def random_blocks(n: int, bs: int, nbr: int, density: float, seed: int = 0) -> Csr:
    rowptr, col, val = _two_call("synth_random_blocks", n, n, bs, nbr, density, seed)
    return Csr(n, n, rowptr, col, val, f"random_blocks_{n}_{bs}_{density:.2f}")


def rmat(scale: int, ef: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 0) -> Csr:
    """R-MAT graph adjacency (duplicates merged), values U[-1,1]."""
    n = 1 << scale
    rowptr, col, val = _two_call("synth_rmat", n, scale, ef, a, b, c, seed)
    return Csr(n, n, rowptr, col, val, f"rmat_{scale}_{ef}")


def uniform_blocks(nb: int, bs: int, bnnz: int, seed: int = 0x0498BC04) -> Csr:
    """nb x nb copies of one random bs x bs pattern with bnnz entries, values 1
    (the micro-benchmark matrices of spmv/orig/88.cpp:43-202)."""
    n = nb * bs
    rowptr, col, val = _two_call("synth_uniform_blocks", n, nb, bs, bnnz, seed)
    return Csr(n, n, rowptr, col, val, f"uniform_blocks_{nb}x{bs}_{bnnz}")


def batch_member(rng, i: int, seed: int) -> Csr:
    """One matrix of the multi-matrix batch workload: n log-uniform in
    [1e3, 2e5], family drawn uniformly from 5-point Laplacian / banded
    (bandwidth 4-64) / R-MAT edge factor 8 / random blocks (4-32, 2-8 blocks
    per block row, density 1-30%); rng draws the parameters, seed + i seeds
    the generator."""
    import math
    n = int(math.exp(rng.uniform(math.log(1e3), math.log(2e5))))
    fam = int(rng.integers(0, 4))
    if fam == 0:
        return laplacian2d(max(2, int(round(math.sqrt(n)))))
    if fam == 1:
        return banded(n, int(rng.integers(4, 65)), seed=seed + i)
    if fam == 2:
        return rmat(max(4, int(round(math.log2(n)))), 8, seed=seed + i)
    return random_blocks(n, int(rng.choice([4, 8, 16, 32])), int(rng.integers(2, 9)),
                         float(rng.uniform(0.01, 0.30)), seed=seed + i)


def from_triplets(m: int, n: int, rows, cols, vals, name: str = "") -> Csr:
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(m + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    return Csr(m, n, rowptr, cols.astype(np.int32), vals, name)


def random_sparse(m: int, n: int, density: float, seed: int = 0) -> Csr:
    """Uniform random pattern (numpy; for small test matrices)."""
    rng = np.random.default_rng(seed)
    k = max(1, int(m * n * density))
    flat = np.unique(rng.integers(0, m * n, size=k, dtype=np.int64))
    rows, cols = flat // n, flat % n
    vals = rng.standard_normal(flat.size)
    return from_triplets(m, n, rows, cols, vals, f"random_{m}x{n}_{density}")
