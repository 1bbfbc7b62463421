"""Learned block-size / kernel-variant tuner.

A 2 x 64 hidden-layer FP32 MLP (abhsf_amd/csrc/mlp.hip, trained and run on
the GPU through the C ABI) predicts log(SpMV time per entry) for every
candidate (s, variant) from cheap matrix statistics; the tuner picks the
candidate with the smallest prediction.  Two models share the features: one
trained on single-launch latency (MODEL_PATH), one on per-matrix throughput
inside a batched launch (BATCH_MODEL_PATH), where launch-level parallelism
comes from the other matrices and per-item overheads dominate.  Labels come from exhaustive timing
sweeps of this library (`sweep`), which also measure the tuner's regret
(block_statistics/consistency.cpp:147-149: (t_pick - t_min) / t_min * 100).

Features (the statistics block_statistics/ computes, block_stats.cpp:78-100,
174-221 and rows-nnz-stdev.cpp:11-30, taken on the GPU):
  global      log2 n, log2 nnz, log2 mean row nnz, row-nnz stdev / mean,
              log2 (max row nnz / mean), fraction of empty rows
  per (s, v)  log2 s, log2 (blocks / entries), fraction of blocks and of
              entries in each scheme, log2 predicted stored bytes per entry,
              strip-work stdev / mean, log2 (max strip work / mean),
              fraction of entries in strips above the split threshold,
              log2 strips, fraction of entries in light strips whose blocks
              are all COO (what the tiled variant gathers into tiles),
              variant flags (balanced, panel, delta, tiled)
Block-size histograms come from abhsf_block_hist on a sample of strips.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np

from . import capi, hip

NAME = "mlp-2x64"
S_CANDIDATES = (8, 16, 32, 64, 128, 256)
VARIANTS = ("strip", "balanced", "panel", "tiled")
DELTA = (0, "delta")  # the delta-coded coordinate format (no block size), one more candidate
# the delta formats: candidate variant -> DeltaMatrix codec (spmv/unified/spmv.cpp:628-1033,
# spmv/orig/delta.cpp:118-160)
DELTA_CODECS = {"delta": "varint", "delta_u32_row": "u32_row", "delta_u32_block": "u32_block"}
HIDDEN = 64
_HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(_HERE, "tuner_model.json")          # single-matrix SpMV latency
BATCH_MODEL_PATH = os.path.join(_HERE, "tuner_model_batch.json")  # throughput inside a batched launch
BATCH_TARGET_NNZ = 30_000_000  # entries per timed batch of replicas
SAMPLE_NNZ = 2_000_000  # entries counted per block-size histogram


# ---- cost model of the stored layout (mirrors Storage::stored_bytes) --------

def _ceil_log2(x: int) -> int:
    return 0 if x <= 1 else (x - 1).bit_length()


def choose(s: int, z: int, b: int = 64) -> int:
    """Scheme index (0 COO, 1 CSR, 2 bitmap, 3 dense) under the
    preprocessing/test.cpp:55-77 model."""
    k, lss = _ceil_log2(s), _ceil_log2(s * s)
    coo = 2 * k * z + lss
    csr = k * z + (s + 1) * lss
    best, sc = coo, 0
    if csr < best:
        best, sc = csr, 1
    if s * s < best:
        best, sc = s * s, 2
    if (s * s - z) * b < best:
        sc = 3
    return sc


def layout_bytes(s, m, n, nb, ne, val_bytes=8):
    """Stored bytes of the HBM layout for per-scheme block/value counts."""
    M, N = -(-m // s), -(-n // s)
    bcol = 2 if N <= 65536 else 4
    idx = 1 if s <= 16 else 2
    rp = 1 if s <= 8 else 2
    W = max(1, s * s // 64)
    tot = 4 * 2 * 4 * (M + 1)
    for x in range(4):
        tot += nb[x] * (bcol + 2) + ne[x] * val_bytes
    tot += ne[0] * idx + nb[1] * (s + 1) * rp + ne[1] + nb[2] * W * 8
    return tot


# ---- features -----------------------------------------------------------------

def block_hist(rowptr, col, m, n, s, stride, on_device=True, nnz=None, stream=None):
    """GPU block histogram of s x s blocks over every stride-th strip."""
    cap = s * s + 2
    hist = np.zeros(cap, dtype=np.uint64)
    nbk, ns = C.c_uint64(0), C.c_uint64(0)
    k = s.bit_length() - 1
    if on_device:
        rp_p, c_p = rowptr.ptr, col.ptr
    else:
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        rp_p, c_p = rowptr.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p)
        nnz = int(rowptr[-1])
    capi.call("abhsf_block_hist", rp_p, c_p, m, n, nnz, int(on_device), k, k, stride,
              hist.ctypes.data_as(C.c_void_p), cap, C.byref(nbk), C.byref(ns),
              None if stream is None else stream.handle)
    return hist, int(nbk.value), int(ns.value)


def global_features(rowptr, m, n):
    deg = np.diff(rowptr).astype(np.float64)
    nnz = float(rowptr[-1])
    mean = nnz / m
    sd = float(np.std(deg))
    return {
        "log2_n": math.log2(max(n, 1)), "log2_nnz": math.log2(max(nnz, 1)),
        "log2_mean_row": math.log2(max(mean, 1e-9)), "row_cv": sd / max(mean, 1e-9),
        "log2_max_row": math.log2(max(deg.max(initial=0.0), 1.0) / max(mean, 1e-9)),
        "empty_rows": float(np.mean(deg == 0)),
    }


def unit_entries(s):
    return max(2048, 8 * s)


def coo_max_z(s: int) -> int:
    """Largest z whose block is stored as COO (the scheme is monotone in z)."""
    z = 1
    while z < s * s and choose(s, z + 1) == 0:
        z += 1
    return z


def tile_fraction(rowptr, col, m, s, stride):
    """Fraction of the sampled strips' entries (every stride-th strip) that
    lie in light strips whose blocks are all COO."""
    k = s.bit_length() - 1
    M = -(-m // s)
    strips = np.arange(0, M, stride, dtype=np.int64)
    e0 = rowptr[strips * s]
    e1 = rowptr[np.minimum(strips * s + s, m)]
    cnt = (e1 - e0).astype(np.int64)
    tot = int(cnt.sum())
    if tot == 0:
        return 0.0
    idx = np.repeat(e0 - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(tot, dtype=np.int64)
    sid = np.repeat(np.arange(strips.size, dtype=np.int64), cnt)
    key = (sid << np.int64(32)) | (col[idx].astype(np.int64) >> np.int64(k))
    u, z = np.unique(key, return_counts=True)
    mixed = np.zeros(strips.size, dtype=bool)
    mixed[(u >> np.int64(32))[z > coo_max_z(s)]] = True
    pure = ~mixed & (cnt <= 2 * unit_entries(s))
    return float(cnt[pure].sum()) / tot


def size_features(rowptr, m, n, s, hist, nblk, sampled, tile_frac=0.0):
    nnz = int(rowptr[-1])
    scale = nnz / max(sampled, 1)
    nb = [0.0] * 4
    ne = [0.0] * 4
    z = np.nonzero(hist)[0]
    for zz in z:
        c = float(hist[zz]) * scale
        sc = choose(s, int(zz))
        nb[sc] += c
        ne[sc] += c * (s * s if sc == 3 else int(zz))
    totb = max(sum(nb), 1.0)
    tote = max(sum(ne), 1.0)
    stored = layout_bytes(s, m, n, nb, ne)
    M = -(-m // s)
    rp = rowptr[np.minimum(np.arange(0, M + 1) * s, m)]
    work = np.diff(rp).astype(np.float64)
    mean = max(float(work.mean()), 1e-9)
    H = 2 * unit_entries(s)
    return {
        "log2_s": math.log2(s), "log2_blk_per_entry": math.log2(max(nblk, 1) / max(sampled, 1)),
        "blk_frac": [v / totb for v in nb], "ent_frac": [v / tote for v in ne],
        "log2_bytes_per_entry": math.log2(max(stored, 1) / max(nnz, 1)),
        "strip_cv": float(work.std()) / mean, "log2_max_strip": math.log2(max(float(work.max()), 1.0) / mean),
        "heavy_frac": float(work[work > H].sum()) / max(float(work.sum()), 1.0), "log2_M": math.log2(M),
        "tile_frac": tile_frac, "stored_bytes": stored,
    }


def feature_vector(g, sf, variant):
    return np.array([
        g["log2_n"], g["log2_nnz"], g["log2_mean_row"], g["row_cv"], g["log2_max_row"], g["empty_rows"],
        sf["log2_s"], sf["log2_blk_per_entry"], *sf["blk_frac"], *sf["ent_frac"],
        sf["log2_bytes_per_entry"], sf["strip_cv"], sf["log2_max_strip"], sf["heavy_frac"], sf["log2_M"],
        sf["tile_frac"], 1.0 if variant == "balanced" else 0.0, 1.0 if variant == "panel" else 0.0,
        1.0 if variant in DELTA_CODECS else 0.0, 1.0 if variant == "tiled" else 0.0,
    ], dtype=np.float64)


NFEAT = 6 + 20  # 6 global + 20 per candidate (fractions are 4 values each)


def _band_sample(rowptr, col, limit=SAMPLE_NNZ, band=64):
    """Elements of a sample of bands of `band` rows (band-aligned), drawn by
    systematic sampling over element positions (a band is taken when one of
    the evenly spaced sample points falls into it, so with probability
    pi = min(1, entries / step)), about `limit` elements in all.  Returns
    rows, cols (uint64), the element weights 1 / pi (Horvitz-Thompson: sums
    of weights estimate totals without bias), each element's predecessor in
    the whole matrix (row and column; -1 before the first) and the band of
    every element.  Small matrices: every element, weight 1."""
    m = rowptr.size - 1
    nnz = int(rowptr[-1])
    nb = -(-m // band)
    bptr = rowptr[np.minimum(np.arange(nb + 1, dtype=np.int64) * band, m)]
    size = np.diff(bptr).astype(np.float64)
    if nnz <= limit:
        take = np.arange(nb, dtype=np.int64)
        pi = np.ones(nb)
    else:
        # sampling measure: entries plus rows (at the mean row length), so
        # that per-row events (row gaps, escapes) in sparse bands are sampled
        # as well as per-entry ones
        rowsz = np.diff(np.minimum(np.arange(nb + 1, dtype=np.int64) * band, m)).astype(np.float64)
        meas = size + rowsz * (nnz / max(m, 1))
        mptr = np.concatenate([[0.0], np.cumsum(meas)])
        lo, hi = 1.0, float(mptr[-1])  # step with about `limit` expected elements
        for _ in range(40):
            step = 0.5 * (lo + hi)
            if float(np.sum(size * np.minimum(1.0, meas / step))) > limit:
                lo = step
            else:
                hi = step
        step = hi
        pts = (np.arange(int(mptr[-1] // step) + 1, dtype=np.float64) + 0.5) * step
        take = np.unique(np.searchsorted(mptr, pts[pts < mptr[-1]], side="right") - 1)
        pi = np.minimum(1.0, meas[take] / step)
    e0 = bptr[take]
    cnt = (bptr[take + 1] - e0).astype(np.int64)
    tot = int(cnt.sum())
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]]) if take.size else np.zeros(0, dtype=np.int64)
    idx = np.repeat(e0 - first, cnt) + np.arange(tot, dtype=np.int64)
    rows = (np.searchsorted(rowptr, idx, side="right") - 1).astype(np.int64)
    prev = idx - 1
    prow = np.where(prev >= 0, np.searchsorted(rowptr, np.maximum(prev, 0), side="right") - 1, -1)
    pcol = np.where(prev >= 0, col[np.maximum(prev, 0)], -1).astype(np.int64)
    w = np.repeat(1.0 / np.maximum(pi, 1e-12), cnt)
    return (rows.astype(np.uint64), col[idx].astype(np.uint64), w, prow, pcol, np.repeat(take, cnt), idx)


def delta_features(rowptr, col, n, codec="varint", sample=None):
    """Per-candidate features of a delta format from the CSR, the element
    statistics estimated on a sample of row bands (_band_sample, unbiased):
    the fractions of elements stored in 1/2/4/8-byte words (in the
    scheme-fraction slots; the u32 codecs use the 4-byte slot for their words
    and the 8-byte slot for exceptions), stored bytes per entry, log2 of the
    runs of equal rows per entry (the SpMV makes one atomic update per run;
    in the blocks-per-entry slot), log2 of the chunk count."""
    nnz = int(rowptr[-1])
    tot = max(nnz, 1)
    chunks = -(-nnz // 512)
    rows, c, w, prow, pcol, _, idx = sample if sample is not None else _band_sample(rowptr, col)
    nonempty = int(np.count_nonzero(np.diff(rowptr)))
    has_prev = prow >= 0
    pr = np.maximum(prow, 0).astype(np.uint64)
    pc = np.maximum(pcol, 0).astype(np.uint64)
    if codec == "varint":
        K = np.uint64(max(1, _ceil_log2(n)))
        d = ((rows << K) + c) - np.where(has_prev, (pr << K) + pc, np.uint64(0))
        cls = [float(np.sum(w[d < (1 << 6)])), float(np.sum(w[(d >= (1 << 6)) & (d < (1 << 14))])),
               float(np.sum(w[(d >= (1 << 14)) & (d < (1 << 30))])), float(np.sum(w[d >= (1 << 30)]))]
        stored = 8 * nnz + cls[0] + 2 * cls[1] + 4 * cls[2] + 8 * cls[3] + 4 * -(-nnz // 16) + 24 * chunks
        runs = nonempty
    else:
        chunk_first = idx % 512 == 0  # a chunk's first element never escapes (its position in key order:
        # the u32_row order is the CSR order; the cache-block order keeps each 8-row band's positions)
        if codec == "u32_row":
            d = ((rows << np.uint64(31)) + c) - np.where(has_prev, (pr << np.uint64(31)) + pc, np.uint64(0))
            nexc = float(np.sum(w[(d >= np.uint64(1 << 32)) & ~chunk_first]))
            runs = nonempty
        else:
            # cache-block order: an element escapes only as the first (smallest
            # rest key) of an 8-row band following the previous nonempty band's
            # last (largest rest key) by 2^32 or more; the band's first key in
            # key order is taken at the band's first CSR position
            rest = lambda r, cc: ((cc >> np.uint64(3)) << np.uint64(6)) + ((r & np.uint64(7)) << np.uint64(3)) + \
                (cc & np.uint64(7))
            band8 = rows >> np.uint64(3)
            nexc = 0.0
            if rows.size:
                bstart = np.flatnonzero(np.diff(band8, prepend=~np.uint64(0)) != 0)
                rs = rest(rows, c)
                kfirst = (band8[bstart] << np.uint64(31)) + np.minimum.reduceat(rs, bstart)
                # the previous nonempty 8-row band in the whole matrix: its last key
                r0 = (band8[bstart] << np.uint64(3)).astype(np.int64)
                e0 = rowptr[r0]
                has = e0 > 0
                lr = np.searchsorted(rowptr, np.maximum(e0 - 1, 0), side="right") - 1
                pb = (np.maximum(lr, 0) >> 3).astype(np.int64)
                q0, q1 = rowptr[pb * 8], rowptr[np.minimum(pb * 8 + 8, rowptr.size - 1)]
                klast = np.zeros(bstart.size, dtype=np.uint64)
                for i in np.flatnonzero(has):  # one previous band per sampled 8-row band
                    rr = np.repeat(np.arange(pb[i] * 8, min(pb[i] * 8 + 8, rowptr.size - 1)),
                                   np.diff(rowptr[pb[i] * 8:min(pb[i] * 8 + 8, rowptr.size - 1) + 1]))
                    cc = col[q0[i]:q1[i]].astype(np.uint64)
                    klast[i] = (np.uint64(pb[i]) << np.uint64(31)) + np.max(rest(rr.astype(np.uint64), cc))
                esc = has & ((kfirst - klast) >= np.uint64(1 << 32)) & ~chunk_first[bstart]
                nexc = float(np.sum(w[bstart][esc]))
            # runs: distinct (row, col / 8) pairs
            same = np.zeros(rows.size, dtype=bool)
            if rows.size > 1:
                same[1:] = (rows[1:] == rows[:-1]) & ((c[1:] >> np.uint64(3)) == (c[:-1] >> np.uint64(3)))
            runs = float(np.sum(w[~same]))
        cls = [0.0, 0.0, nnz - nexc, nexc]
        stored = 12 * nnz + 8 * nexc + 8 * chunks + 4 * (chunks + 1)
    return {
        "log2_s": 0.0, "log2_blk_per_entry": math.log2(max(runs, 1) / tot), "blk_frac": [0.0] * 4,
        "ent_frac": [v / tot for v in cls], "log2_bytes_per_entry": math.log2(max(stored, 1) / tot),
        "strip_cv": 0.0, "log2_max_strip": 0.0, "heavy_frac": 0.0, "log2_M": math.log2(max(chunks, 1)),
        "tile_frac": 0.0, "stored_bytes": stored,
    }


# Batched launches use the balanced and tiled variants: inside one launch a
# whole heavy strip on one wave (strip variant) sets the launch's tail, which
# the per-member cycle label cannot see, and column panels only add partial
# rows.
BATCH_VARIANTS = ("balanced", "tiled")


def candidates(variants=VARIANTS + tuple(DELTA_CODECS)):
    """(s, variant) of the ABHSF configurations, then the delta formats
    (block size 0) that are allowed."""
    out = [(s, v) for s in S_CANDIDATES for v in VARIANTS if v in variants]
    return out + [(0, v) for v in DELTA_CODECS if v in variants]


def matrix_features(A, d_rowptr=None, d_col=None, stream=None):
    """Feature rows for every candidate of matrix A (synth.Csr)."""
    g = global_features(A.rowptr, A.m, A.n)
    stride = max(1, A.nnz // SAMPLE_NNZ)
    rows = []
    per_s = {}
    for s in S_CANDIDATES:
        if d_rowptr is not None:
            hist, nblk, ns = block_hist(d_rowptr, d_col, A.m, A.n, s, stride, True, A.nnz, stream)
        else:
            hist, nblk, ns = block_hist(A.rowptr, A.col, A.m, A.n, s, stride, False)
        per_s[s] = size_features(A.rowptr, A.m, A.n, s, hist, nblk, ns,
                                 tile_fraction(A.rowptr, A.col, A.m, s, stride))
    sample = _band_sample(A.rowptr, A.col)  # shared by the delta formats' features
    for s, v in candidates():
        sf = per_s[s] if v not in DELTA_CODECS else delta_features(A.rowptr, A.col, A.n, DELTA_CODECS[v], sample)
        rows.append(feature_vector(g, sf, v))
    return np.array(rows), per_s


# ---- exhaustive timing sweep (labels and regret) -------------------------------

def build_candidate(A, s, variant, d_rp, d_c, d_v, stream):
    """The matrix object of one candidate (ABHSF at (s, variant), or the
    delta format), built from a device CSR."""
    from .matrix import AbhsfMatrix, DeltaMatrix
    if variant in DELTA_CODECS:
        M = DeltaMatrix(stream=stream, codec=DELTA_CODECS[variant])
    else:
        M = AbhsfMatrix(block_size=s, variant=variant, stream=stream)
    M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True, nnz=A.nnz)
    return M


def time_candidate(A, s, variant, d_rp, d_c, d_v, x, y, stream, reps=10):
    M = build_candidate(A, s, variant, d_rp, d_c, d_v, stream)
    for _ in range(2):
        M.spmv_async(1.0, x, 1.0, y)
    stream.synchronize()
    times = []
    for _ in range(3):
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(reps):
            M.spmv_async(1.0, x, 1.0, y)
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_ms(e1) / reps)
    M.release()
    return float(np.median(times))


def time_candidate_batched(A, s, variant, d_rp, d_c, d_v, x, stream, reps=5):
    """Throughput label: ms per copy of A inside one batched launch (hipGraph
    replay) of R references to the same built matrix, R chosen so that the
    launch holds about BATCH_TARGET_NNZ entries.  The copies share x and the
    heavy-strip partial buffer (so their y values are not meaningful: this
    is a timing run); each has its own y."""
    from .matrix import AbhsfBatch, AbhsfMatrix
    M = AbhsfMatrix(block_size=s, variant=variant, stream=stream)
    M.from_csr(A.m, A.n, d_rp, d_c, d_v, on_device=True, nnz=A.nnz)
    R = int(min(256, max(1, -(-BATCH_TARGET_NNZ // max(A.nnz, 1)))))
    ys = [hip.DeviceArray.zeros(A.m) for _ in range(R)]
    B = AbhsfBatch([M] * R, stream)
    B.set_vectors([x] * R, ys)
    for _ in range(2):
        B.spmv_async(1.0, 1.0, graph=True)
    B.synchronize()
    times = []
    for _ in range(3):
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(reps):
            B.spmv_async(1.0, 1.0, graph=True)
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_ms(e1) / reps / R)
    B.release()
    M.release()
    return float(np.median(times))


def sweep(A, stream=None, batched=False):
    """Time every candidate on A; returns {(s, variant): ms} (single launch
    latency, or per-matrix throughput inside a batch when batched)."""
    stream = stream or hip.Stream()
    d_rp, d_c, d_v = (hip.DeviceArray.from_numpy(a) for a in (A.rowptr, A.col, A.val))
    x = hip.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, A.n))
    if batched:
        return {(s, v): time_candidate_batched(A, s, v, d_rp, d_c, d_v, x, stream) for s, v in candidates()}
    y = hip.DeviceArray.zeros(A.m)
    return {(s, v): time_candidate(A, s, v, d_rp, d_c, d_v, x, y, stream) for s, v in candidates()}


def regret(times, pick):
    """(t_pick - t_min) / t_min * 100 (block_statistics/consistency.cpp:147-149)."""
    tmin = min(times.values())
    return (times[pick] - tmin) / tmin * 100.0


# ---- the model ---------------------------------------------------------------------

def init_theta(F, seed=0):
    rng = np.random.default_rng(seed)
    P = HIDDEN * F + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
    t = np.zeros(P, dtype=np.float32)
    o = 0
    t[o:o + HIDDEN * F] = rng.uniform(-1, 1, HIDDEN * F) * math.sqrt(6.0 / F)
    o += HIDDEN * F + HIDDEN
    t[o:o + HIDDEN * HIDDEN] = rng.uniform(-1, 1, HIDDEN * HIDDEN) * math.sqrt(6.0 / HIDDEN)
    o += HIDDEN * HIDDEN + HIDDEN
    t[o:o + HIDDEN] = rng.uniform(-1, 1, HIDDEN) * math.sqrt(6.0 / HIDDEN)
    return t


class Model:
    """An ensemble of MLPs sharing the input normalisation: theta is one
    parameter vector or a [members, P] array; predictions are the members'
    mean."""

    def __init__(self, theta, mu, sd, ymu=0.0, meta=None):
        self.theta = np.asarray(theta, dtype=np.float32)
        self.mu = np.asarray(mu, dtype=np.float64)
        self.sd = np.asarray(sd, dtype=np.float64)
        self.ymu = float(ymu)
        self.meta = meta or {}

    @classmethod
    def load(cls, path=MODEL_PATH):
        d = json.load(open(path))
        return cls(d["theta"], d["mu"], d["sd"], d.get("ymu", 0.0), d.get("meta"))

    def save(self, path=MODEL_PATH):
        json.dump({"theta": self.theta.astype(np.float64).tolist(), "mu": list(map(float, self.mu)),
                   "sd": list(map(float, self.sd)), "ymu": self.ymu, "meta": self.meta},
                  open(path, "w"))

    def compatible(self) -> bool:
        """True when the model was trained on this feature layout (NFEAT inputs)."""
        return self.mu.shape == (NFEAT,)

    def normalize(self, X):
        X = np.asarray(X, dtype=np.float64)
        if X.shape[-1] != self.mu.shape[0]:
            raise ValueError(f"model expects {self.mu.shape[0]} features, got {X.shape[-1]}: retrain it "
                             "(python -m abhsf_amd.train_tuner)")
        return ((X - self.mu) / self.sd).astype(np.float32)

    def predict(self, X):
        """log(ms per entry) predictions: the mean over the ensemble members,
        all members and candidates in one launch of the GPU inference kernel
        (abhsf_mlp_predict), one copy back."""
        Xn = self.normalize(X)
        B, F = Xn.shape
        members = self.theta.reshape(-1, self.theta.shape[-1])
        if getattr(self, "_dtheta", None) is None:  # parameters stay on the device
            self._dtheta = hip.DeviceArray.from_numpy(np.ascontiguousarray(members))
        dx = hip.DeviceArray.from_numpy(Xn)
        out = hip.DeviceArray((B,), np.float64)
        capi.call("abhsf_mlp_predict", self._dtheta.ptr, members.shape[0], F, dx.ptr, B, out.ptr, None)
        return out.numpy() + self.ymu


def train(X, y, w=None, steps=700, lr=1e-2, seed=0, mu=None, sd=None, members=1):
    """Fit the MLP on the GPU (full-batch Adam by abhsf_mlp_train); with
    members > 1 an ensemble of that many MLPs (seeds seed, seed + 1, ...)."""
    if members > 1:
        fits = [train(X, y, w, steps, lr, seed + i, mu, sd) for i in range(members)]
        model = fits[0][0]
        model.theta = np.stack([m.theta for m, _ in fits])
        model.meta["members"] = members
        model.meta["train_s"] = sum(m.meta["train_s"] for m, _ in fits)
        return model, np.mean([l for _, l in fits], axis=0)
    X = np.asarray(X, dtype=np.float64)
    mu = X.mean(0) if mu is None else mu
    sd = X.std(0) if sd is None else sd
    sd = np.where(sd > 1e-9, sd, 1.0)
    ymu = float(np.mean(y))
    model = Model(init_theta(X.shape[1], seed), mu, sd, ymu)
    Xn = model.normalize(X)
    dt = hip.DeviceArray.from_numpy(model.theta)
    dx = hip.DeviceArray.from_numpy(Xn)
    dy = hip.DeviceArray.from_numpy(np.asarray(y, dtype=np.float32) - np.float32(ymu))
    dw = None if w is None else hip.DeviceArray.from_numpy(np.asarray(w, dtype=np.float32))
    losses = np.zeros(steps, dtype=np.float32)
    t0 = time.perf_counter()
    capi.call("abhsf_mlp_train", dt.ptr, X.shape[1], dx.ptr, dy.ptr, None if dw is None else dw.ptr,
              X.shape[0], steps, float(lr), losses.ctypes.data_as(C.c_void_p), None)
    model.theta = dt.numpy()
    model.meta = {"steps": steps, "lr": lr, "samples": int(X.shape[0]), "final_loss": float(losses[-1]),
                  "first_loss": float(losses[0]), "train_s": time.perf_counter() - t0}
    return model, losses


def pick(model, Xc):
    """Candidate with the smallest predicted time among the variants the model
    was trained on (Xc: feature rows of every candidate), and the
    predictions."""
    allowed = model.meta.get("variants", VARIANTS + tuple(DELTA_CODECS))
    pred = model.predict(Xc)
    cands = candidates()
    ok = [i for i, (_, v) in enumerate(cands) if v in allowed]
    return cands[min(ok, key=lambda i: pred[i])], pred


def pick_storage_optimal(per_s):
    """The reference's exhaustive search: the block size with the fewest
    stored bits (block_statistics/mmf.h:77-102), default variant."""
    s = min(per_s, key=lambda k: per_s[k]["stored_bytes"])
    return (s, "balanced")


def pick_block_size(A, candidates=S_CANDIDATES, model_path=MODEL_PATH):
    """Block size chosen by the trained model (None when no model is present)."""
    if not os.path.exists(model_path):
        return None
    model = Model.load(model_path)
    Xc, _ = matrix_features(A)
    cand = globals()["candidates"]()
    pred = model.predict(Xc)
    allowed = [i for i, (s, _) in enumerate(cand) if s in candidates]
    best = min(allowed, key=lambda i: pred[i])
    return cand[best][0]
