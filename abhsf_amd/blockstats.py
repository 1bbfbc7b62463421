"""Block statistics and storage-footprint tables on the GPU.

The reference's analysis pipeline (block_statistics/) reads a matrix, sorts it
64 times to build the histograms "entries per nonzero r x s block -> number of
blocks" for (r, s) in {2..256}^2 (block_stats.cpp:69-170), counts per-row and
per-column entries (block_stats.cpp:174-221), and derives from the histograms
the memory footprint of 8 storage variants per block size (mmf.cpp:53-138),
the best block size (mmf.h:77-102, min-bss.cpp, min-schemes.cpp) and the
regret of restricting the block-size set (consistency.cpp:108-190).

Here the histograms come from abhsf_block_hist (GPU sort + run-length count);
the tables are the reference's formulas evaluated on them.
"""
from __future__ import annotations

import numpy as np

from . import tuner

POWERS = tuple(range(1, 9))  # r, s = 2 .. 256


def ceil_log2(x: int) -> int:
    """ceil(log2 x), log2(1) = 0 (include/abhsf/utils/math.h:17-31)."""
    return 0 if x <= 1 else (x - 1).bit_length()


def histograms(A, pairs=None, stream=None):
    """{(r, s): {z: count}} for every (r, s) = (2^kr, 2^ks) pair (default all 64)."""
    import ctypes as C
    from . import capi
    out = {}
    pairs = pairs or [(1 << a, 1 << b) for a in POWERS for b in POWERS]
    for r, s in pairs:
        cap = r * s + 2
        h = np.zeros(cap, dtype=np.uint64)
        nb, ns = C.c_uint64(0), C.c_uint64(0)
        capi.call("abhsf_block_hist", A.rowptr.ctypes.data_as(C.c_void_p), A.col.ctypes.data_as(C.c_void_p),
                  A.m, A.n, A.nnz, 0, r.bit_length() - 1, s.bit_length() - 1, 1, h.ctypes.data_as(C.c_void_p),
                  cap, C.byref(nb), C.byref(ns), None if stream is None else stream.handle)
        if sum(int(z) * int(h[z]) for z in np.nonzero(h)[0]) != A.nnz:
            raise RuntimeError("Nonzero elements counts do not match!")  # block_stats.cpp:167-168
        out[(r, s)] = {int(z): int(h[z]) for z in np.nonzero(h)[0]}
    return out


def rows_cols_nnz(A, symmetric=False):
    """rows/cols entry counts, stored and all (block_stats.cpp:174-199)."""
    r, c, _ = A.triplets()
    rs = np.bincount(r, minlength=A.m)
    cs = np.bincount(c, minlength=A.n)
    if not symmetric:
        return rs, cs, rs.copy(), cs.copy()
    off = r != c
    ra = rs + np.bincount(c[off], minlength=A.m)
    ca = cs + np.bincount(r[off], minlength=A.n)
    return rs, cs, ra, ca


def rows_nnz_stdev(rows_nnz) -> float:
    """Population standard deviation (rows-nnz-stdev.cpp:11-30)."""
    a = np.asarray(rows_nnz, dtype=np.float64)
    return float(np.sqrt(np.mean((a - a.mean()) ** 2)))


def _msmf_coo(r, s, z):
    return ceil_log2(r * s) + z * (ceil_log2(r) + ceil_log2(s))


def _msmf_csr(r, s, z):
    return r * ceil_log2(r * s) + z * ceil_log2(s)


def _msmf_min(r, s, z, b, binary, with_csr):
    m = _msmf_coo(r, s, z)
    if with_csr:
        m = min(m, _msmf_csr(r, s, z))
    m = min(m, r * s)
    if not binary:
        m = min(m, (r * s - z) * b)
    return m


def mmf_row(m, n, nnz, r, s, hist, bits_single=32, mult=2, binary=False):
    """[coo32, csr32, coo, csr, bitmap, dense, adaptive, adaptive134] in bits,
    values included (block_statistics/mmf.cpp:53-138)."""
    M, N = -(-m // r), -(-n // s)
    NNZ = sum(hist.values())
    csr0 = M * ceil_log2(N + 1) + NNZ * ceil_log2(N)
    b = bits_single * mult
    coo = csr = bm = dense = ab = ab134 = csr0
    if binary:
        dense = 2 ** 64 - 1
    for z, cnt in hist.items():
        coo += _msmf_coo(r, s, z) * cnt
        csr += _msmf_csr(r, s, z) * cnt
        bm += r * s * cnt
        if not binary:
            dense += (r * s - z) * b * cnt
        ab += (_msmf_min(r, s, z, b, binary, True) + 2) * cnt
        ab134 += (_msmf_min(r, s, z, b, binary, False) + 2) * cnt
    valf = nnz * b
    return [(v + valf) % (2 ** 64) for v in (nnz * 64, (m + 1 + nnz) * 32, coo, csr, bm, dense, ab, ab134)]


def mmf_table(A, hists, mult=2, binary=False):
    return {k: mmf_row(A.m, A.n, A.nnz, k[0], k[1], h, 32, mult, binary) for k, h in hists.items()}


def block_opt(table):
    """Smallest footprint over the variants 2..6 (coo, csr, bitmap, dense,
    adaptive) and all block sizes (block_statistics/mmf.h:77-102)."""
    best, arg = None, None
    for bs in sorted(table):
        for k in range(2, 7):
            if best is None or table[bs][k] < best:
                best, arg = table[bs][k], (bs, k)
    return best, arg


def min_per_block_size(table):
    """min over variants 2..6 per block size (min-bss.cpp:34-58)."""
    return {bs: min(row[2:7]) for bs, row in table.items()}


def subset_regret(tables, scheme_adaptive=True):
    """Regret (t_min - min) / min * 100 of restricting the block sizes to the
    sets B_64, B_20, B_14, B_8 (block_statistics/consistency.cpp:108-190),
    averaged and maximised over matrices (each table an mmf table)."""
    B8 = {(h, h) for h in (2, 4, 8, 16, 32, 64, 128, 256)}
    B14 = B8 | {(h, w) for h in (4, 8, 16) for w in (4, 8, 16)}
    B20 = B8 | {(h, w) for h in (4, 8, 16, 32) for w in (4, 8, 16, 32)}
    B64 = {(h, w) for h in (2, 4, 8, 16, 32, 64, 128, 256) for w in (2, 4, 8, 16, 32, 64, 128, 256)}
    sets = {"B_64": B64, "B_20": B20, "B_14": B14, "B_8": B8}
    out = {k: [] for k in sets}
    for t in tables:
        mn = min(min(min(r[2], r[3], r[4], r[5]), r[6]) for r in t.values())  # consistency.cpp:77-82
        key = {bs: (r[7] if scheme_adaptive else min(r[2], r[4], r[5]) + 2) for bs, r in t.items()}
        for name, S in sets.items():
            tmin = min(key[bs] for bs in S)
            out[name].append((tmin - mn) / mn * 100.0)
    return {k: {"average": float(np.mean(v)), "maximum": float(np.max(v))} for k, v in out.items()}


def storage_optimal_block_size(A, candidates=tuner.S_CANDIDATES):
    """Square block size with the smallest stored bytes of this library's layout."""
    best = None
    for s in candidates:
        h = histograms(A, [(s, s)])[(s, s)]
        hist = np.zeros(s * s + 2, dtype=np.uint64)
        for z, c in h.items():
            hist[z] = c
        b = tuner.size_features(A.rowptr, A.m, A.n, s, hist, sum(h.values()), A.nnz)["stored_bytes"]
        if best is None or b < best[1]:
            best = (s, b)
    return best[0]


__all__ = ["histograms", "rows_cols_nnz", "rows_nnz_stdev", "mmf_row", "mmf_table", "block_opt",
           "min_per_block_size", "subset_regret", "storage_optimal_block_size", "ceil_log2"]


# ---- gaps between consecutive linearised keys (the delta codecs' input) -------

def morton_keys(rows, cols):
    """Bit-interleaved keys: bit i of the row goes to bit 2i, bit i of the
    column to bit 2i + 1 (spmv/orig/morton.cpp:22-30)."""
    r = np.asarray(rows, dtype=np.uint64)
    c = np.asarray(cols, dtype=np.uint64)
    key = np.zeros(r.shape, dtype=np.uint64)
    for i in range(32):
        bit = np.uint64(1) << np.uint64(i)
        key |= ((r & bit) << np.uint64(i)) | ((c & bit) << np.uint64(i + 1))
    return key


def delta_widths(A, order="morton", top=50):
    """Gaps between consecutive keys in the given order ("morton", or "row":
    key = (row << K) + col with K = ceil_log2(n), the delta format's key) and
    their histogram by required bits, ceil_log2(gap) with -1 for a zero gap;
    plus the `top` largest gaps (spmv/orig/morton.cpp:119-147, whose
    distances array is one element short; here it has nnz - 1 entries)."""
    r, c, _ = A.triplets()
    if order == "morton":
        keys = np.sort(morton_keys(r, c))
    elif order == "row":
        K = max(1, ceil_log2(A.n))
        keys = (r.astype(np.uint64) << np.uint64(K)) + c.astype(np.uint64)
    else:
        raise ValueError(order)
    gaps = np.diff(keys).astype(np.uint64)
    hist = {}
    if gaps.size:
        zero = int(np.count_nonzero(gaps == 0))
        if zero:
            hist[-1] = zero
        nz = gaps[gaps > 0]
        # ceil_log2 of a positive gap: bit length of gap - 1
        bits = np.zeros(nz.shape, dtype=np.int64)
        v = nz - np.uint64(1)
        while np.any(v):
            bits += (v > 0)
            v >>= np.uint64(1)
        for b, cnt in zip(*np.unique(bits, return_counts=True)):
            hist[int(b)] = int(cnt)
    largest = sorted((int(g) for g in gaps), reverse=True)[:top] if gaps.size <= 1_000_000 else \
        [int(g) for g in np.sort(gaps)[::-1][:top]]
    return {"hist": dict(sorted(hist.items())), "largest": largest}
