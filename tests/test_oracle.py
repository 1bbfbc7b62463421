"""CPU tests pinning the oracle (oracle/abhsf_oracle.c) before it is trusted.

Pins, per function (each chosen so that a dropped term, wrong sign/index or a
transposed operand fails at least one):
  ceil_log2     brute force against both reference loops (math.h:17-31,
                test.cpp:88-98), including powers of two and their neighbours.
  choose        closed-form crossovers solved by hand for s = 8, 64, 256.
  smf_preproc   golden known answers of preprocessing/test.cpp (test.mtx at
                s = 2..1024; 256^2 Laplacian at s = 64), the Laplacian at the
                other s from an independent restatement of test.cpp:55-77, and
                a count by hand at s = 1024.
  block_hist    oracle/_ref/block_stats (the reference's binary) on test.mtx and
                seeded random matrices; golden bstats; sum z*count = nnz.
  mmf_row       oracle/_ref/mmf (the reference's binary); golden rows.
  rows_stdev    golden value; numpy population stdev.
  build/decode  round trip; popcount(bitmap word) = z; hand-built 8x8 layouts.
  spmm          textbook CSR SpMV; brute-force dense product; closed forms.
  varint        hand-encoded words at every width boundary of
                spmv/unified/spmv.cpp:859-896 (2^6, 2^14, 2^30), the byte
                order, decode(encode(k)) = k on random and 62-bit keys.
  delta32       keys and deltas of hand-placed elements worked out by hand for
                spmv/orig/delta.cpp:124-151 (row << 31 | col) and the 8 x 8
                cache-block key of spmv/unified/spmv.cpp:636-689 (including
                an element the block order moves), escapes at known row gaps
                and at a chunk's first element, the element order checked
                against a lexicographic sort, round trip on seeded matrices.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from abhsf_amd import synth
from oracle import pyoracle as O

from conftest import ROOT, read_mtx

KA = json.load(open(os.path.join(ROOT, "tests", "golden", "known_answers.json")))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def load_test_mtx():
    m, n, r, c, v = read_mtx(os.path.join(ROOT, "tests", "golden", "test.mtx"))
    return m, n, r, c, v


# ---- ceil_log2 --------------------------------------------------------------
def ref_log2_math(n):  # include/abhsf/utils/math.h:17-31
    k = 0
    while (1 << k) < n:
        k += 1
    return k


def ref_log2_preproc(arg):  # preprocessing/test.cpp:88-98
    e = 0
    while (arg >> e) != 0:
        e += 1
    if (arg & (arg - 1)) == 0:
        e -= 1
    return e


def test_ceil_log2_brute_force():
    xs = list(range(1, 5000)) + [(1 << p) + d for p in range(12, 40) for d in (-1, 0, 1)]
    for x in xs:
        assert O.ceil_log2(x) == ref_log2_math(x) == ref_log2_preproc(x) == (x - 1).bit_length(), x


# ---- scheme choice ------------------------------------------------------------
def test_choose_crossovers_by_hand():
    # PREPROC, s = 64, b = 64: COO 12z+12, CSR 6z+780, bitmap 4096, dense (4096-z)*64.
    # COO < CSR  <=>  z < 128; CSR <= COO at z = 128 is a tie -> COO (strict <).
    assert O.choose(O.PREPROC, 64, 127)[0] == O.COO
    assert O.choose(O.PREPROC, 64, 128) == (O.COO, 12 * 128 + 12)
    assert O.choose(O.PREPROC, 64, 129) == (O.CSR, 6 * 129 + 780)
    # CSR < bitmap <=> 6z + 780 < 4096 <=> z <= 552
    assert O.choose(O.PREPROC, 64, 552) == (O.CSR, 6 * 552 + 780)
    assert O.choose(O.PREPROC, 64, 553) == (O.BITMAP, 4096)
    # dense < bitmap <=> (4096 - z)*64 < 4096 <=> z > 4032; z = 4032 ties -> bitmap
    assert O.choose(O.PREPROC, 64, 4032) == (O.BITMAP, 4096)
    assert O.choose(O.PREPROC, 64, 4033) == (O.DENSE, 63 * 64)
    assert O.choose(O.PREPROC, 64, 4096) == (O.DENSE, 0)
    # binary matrices never choose dense (test.cpp:70)
    assert O.choose(O.PREPROC, 64, 4096, b=0, is_binary=True)[0] == O.BITMAP
    # s = 8: COO 6z+6, CSR 3z+54, bitmap 64: COO<CSR <=> z<16; COO<64 <=> z<=9
    assert O.choose(O.PREPROC, 8, 9) == (O.COO, 60)
    assert O.choose(O.PREPROC, 8, 10) == (O.BITMAP, 64)
    assert O.choose(O.PREPROC, 8, 64) == (O.DENSE, 0)
    assert O.choose(O.PREPROC, 8, 63) == (O.BITMAP, 64)  # dense 64 ties bitmap
    # s = 16 never picks CSR under PREPROC (4z+136 < min(8z+8, 256) has no solution)
    assert all(O.choose(O.PREPROC, 16, z)[0] != O.CSR for z in range(1, 257))
    # MSMF, s = 64: COO 12+12z, CSR 64*12+6z = 768+6z, bitmap 4096
    assert O.choose(O.MSMF, 64, 126) == (O.COO, 12 + 12 * 126)
    assert O.choose(O.MSMF, 64, 127) == (O.CSR, 768 + 6 * 127)
    assert O.choose(O.MSMF, 64, 554) == (O.CSR, 768 + 6 * 554)
    assert O.choose(O.MSMF, 64, 555) == (O.BITMAP, 4096)
    # f32 values (b = 32): dense < bitmap <=> (4096-z)*32 < 4096 <=> z > 3968
    assert O.choose(O.PREPROC, 64, 3969, b=32) == (O.DENSE, 127 * 32)


def test_scheme_monotone_in_z():
    for pol in (O.PREPROC, O.MSMF):
        for s in (2, 4, 8, 16, 32, 64, 128, 256):
            prev = -1
            for z in range(1, s * s + 1):
                sc = O.choose(pol, s, z)[0]
                assert sc >= prev, (pol, s, z)
                prev = sc


# ---- footprints against the reference's known answers ---------------------------
def test_smf_preproc_test_mtx_golden():
    m, n, r, c, _ = load_test_mtx()
    for s, bits in KA["test_mtx"]["preproc_smf_bits"].items():
        assert O.smf_preproc(m, n, r, c, int(s))[0] == bits, s


def test_smf_preproc_test_mtx_by_hand():
    # s = 1024: one block (M = N = 1), z = 15: header 0+0+2, COO 2*10*15 + 20 = 320
    m, n, r, c, _ = load_test_mtx()
    assert O.smf_preproc(m, n, r, c, 1024)[0] == 2 + 320


def test_smf_preproc_laplacian_golden():
    A = synth.laplacian2d(256)
    assert A.nnz == KA["laplacian_256"]["nnz"]
    r, c, _ = A.triplets()
    for s, bits in KA["laplacian_256"]["preproc_smf_bits"].items():
        got, nb = O.smf_preproc(A.m, A.n, r, c, int(s))
        assert got == bits, s
        want = KA["laplacian_256"]["preproc_blocks"][s]
        assert nb[O.COO] == want.get("coo", 0) and nb[O.CSR] == want.get("csr", 0)
        assert nb[O.BITMAP] == want.get("bitmap", 0) and nb[O.DENSE] == want.get("dense", 0)


# ---- block statistics against the reference's own binaries -----------------------
def run_ref(tmp_path, rows, cols, m, n, vals=None):
    """Run oracle/_ref/block_stats + mmf on a matrix written as .mtx."""
    if not (os.path.exists(os.path.join(REF_BIN, "block_stats")) and os.path.exists(os.path.join(REF_BIN, "mmf"))):
        pytest.skip("reference binaries not built (needs /root/reference; make -C oracle ref)")
    p = tmp_path / "a.mtx"
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{m} {n} {len(rows)}\n")
        for i, (a, b) in enumerate(zip(rows, cols)):
            v = 1.0 if vals is None else vals[i]
            f.write(f"{a + 1} {b + 1} {v!r}\n")
    subprocess.run([os.path.join(REF_BIN, "block_stats"), str(p)], cwd=tmp_path, check=True, capture_output=True)
    subprocess.run([os.path.join(REF_BIN, "mmf")], cwd=tmp_path, check=True, capture_output=True)
    return tmp_path


def read_bstats(d, r, s):
    out = {}
    for line in open(d / f"{r}x{s}.bstats"):
        z, c = line.split()
        out[int(z)] = int(c)
    return out


def read_mmf(d, name):
    rows = {}
    for line in open(d / name):
        t = [int(v) for v in line.split()]
        rows[(t[0], t[1])] = t[2:]
    return rows


@pytest.mark.parametrize("case", ["test_mtx", "random_a", "random_b"])
def test_block_hist_and_mmf_vs_reference_binary(tmp_path, case):
    if case == "test_mtx":
        m, n, r, c, _ = load_test_mtx()
    else:
        seed = {"random_a": 1, "random_b": 2}[case]
        A = synth.random_sparse(300 + seed * 37, 280 - seed * 11, 0.03 * seed, seed=seed)
        m, n = A.m, A.n
        r, c, _ = A.triplets()
    d = run_ref(tmp_path, r, c, m, n)
    mmf_s, mmf_d = read_mmf(d, "mmf-single"), read_mmf(d, "mmf-double")
    for kr in range(1, 9):
        for ks in range(1, 9):
            h = O.block_hist(r, c, kr, ks)
            assert h == read_bstats(d, 1 << kr, 1 << ks), (kr, ks)
            assert sum(z * cnt for z, cnt in h.items()) == len(r)
            key = (1 << kr, 1 << ks)
            assert O.mmf_row(m, n, len(r), key[0], key[1], h, mult=1) == mmf_s[key], key
            assert O.mmf_row(m, n, len(r), key[0], key[1], h, mult=2) == mmf_d[key], key


def test_block_hist_and_mmf_golden():
    m, n, r, c, _ = load_test_mtx()
    for k, want in KA["test_mtx"]["bstats"].items():
        a, b = (int(v) for v in k.split("x"))
        got = O.block_hist(r, c, a.bit_length() - 1, b.bit_length() - 1)
        assert got == {int(z): cnt for z, cnt in want.items()}
    for k, want in KA["test_mtx"]["mmf_double"].items():
        a, b = (int(v) for v in k.split())
        h = O.block_hist(r, c, a.bit_length() - 1, b.bit_length() - 1)
        assert O.mmf_row(m, n, 15, a, b, h, mult=2) == want
    # 256x256 totals printed by mmf: CSR-ABHSF(64) 259 = structure only (values excluded)
    h = O.block_hist(r, c, 8, 8)
    row = O.mmf_row(m, n, 15, 256, 256, h, mult=2)
    assert row[6] - 15 * 64 == KA["test_mtx"]["mmf_256"]["CSR-ABHSF(64)"]
    assert row[2] - 15 * 64 == KA["test_mtx"]["mmf_256"]["CSR-COO"]
    assert row[3] - 15 * 64 == KA["test_mtx"]["mmf_256"]["CSR-CSR"]
    assert row[4] - 15 * 64 == KA["test_mtx"]["mmf_256"]["CSR-bitmap"]
    assert row[5] - 15 * 64 == KA["test_mtx"]["mmf_256"]["CSR-dense(64)"]


def test_rows_stdev():
    want = KA["test_mtx"]["rows_nnz_stdev"]
    rn = KA["test_mtx"]["rows_nnz_stored"]
    sd = O.rows_stdev(rn)
    assert abs(sd - want[0]) < 1e-6
    assert abs(sd / 7 * 100 - want[2]) < 1e-4  # stdev / n * 100 (rows-nnz-stdev.cpp:41-44)
    x = np.random.default_rng(0).integers(0, 50, 1000)
    assert abs(O.rows_stdev(x) - float(np.std(x))) < 1e-9


# ---- builder ---------------------------------------------------------------------
def roundtrip(A, k, policy=O.PREPROC):
    B = O.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, k, policy)
    r, c, v = B.decode()
    order = np.lexsort((c, r))
    r0, c0, v0 = A.triplets()
    keep = v0 != 0  # dense blocks cannot tell explicit zeros from padding
    assert np.array_equal(r[order], r0[keep]) and np.array_equal(c[order], c0[keep])
    assert np.array_equal(v[order], v0[keep])
    return B


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 8])
def test_build_roundtrip(k):
    for A in (synth.laplacian2d(24), synth.random_sparse(301, 190, 0.08, seed=k),
              synth.block_mix(1024, seed=k), synth.rmat(10, 8, seed=k)):
        for pol in (O.PREPROC, O.MSMF):
            B = roundtrip(A, k, pol)
            st = B.st
            assert sum(B.nb) > 0 and st.nnz == A.nnz
            a = B.arrays()
            s = 1 << k
            # popcount(bitmap) = z for every bitmap block
            if st.nb[O.BITMAP]:
                w = a["bm_words"].reshape(st.nb[O.BITMAP], -1)
                pc = np.array([sum(bin(int(x)).count("1") for x in row) for row in w])
                assert np.array_equal(pc, a["bm_z1"].astype(int) + 1)
            # CSR row pointers end at z
            if st.nb[O.CSR]:
                rp = a["csr_rowptr"].reshape(st.nb[O.CSR], s + 1).astype(int)
                assert np.array_equal(rp[:, -1], a["csr_z1"].astype(int) + 1)
                assert np.all(np.diff(rp, axis=1) >= 0) and np.all(rp[:, 0] == 0)
            # footprint equals the independent smf computation
            if pol == O.PREPROC:
                r, c, _ = A.triplets()
                assert B.smf_bits == O.smf_preproc(A.m, A.n, r, c, s)[0]


def test_build_layout_by_hand():
    """An 8x8 block with the 88.cpp generator layout: bit 8*row+col, COO byte
    (row<<4)|col for s <= 16 becomes (row<<3)|col here (k = 3)."""
    rows = [0, 0, 3, 7]
    cols = [1, 6, 3, 7]
    A = synth.from_triplets(8, 8, rows, cols, [1.0, 2.0, 3.0, 4.0])
    B = O.Abhsf(8, 8, A.rowptr, A.col, A.val, 3)
    a = B.arrays()
    assert B.nb == [1, 0, 0, 0]  # z = 4: COO 6*4+6 = 30 < 64
    assert list(a["coo_idx"]) == [(0 << 3) | 1, (0 << 3) | 6, (3 << 3) | 3, (7 << 3) | 7]
    assert list(a["coo_val"]) == [1.0, 2.0, 3.0, 4.0]
    # 12 entries -> bitmap (COO 78 > 64)
    rr = [0] * 8 + [5, 5, 6, 7]
    cc = list(range(8)) + [0, 7, 2, 5]
    A = synth.from_triplets(8, 8, rr, cc, np.arange(1, 13, dtype=float))
    B = O.Abhsf(8, 8, A.rowptr, A.col, A.val, 3)
    a = B.arrays()
    assert B.nb == [0, 0, 1, 0]
    want = sum(1 << (8 * r + c) for r, c in zip(rr, cc))
    assert int(a["bm_words"][0]) == want
    assert list(a["bm_val"]) == list(np.arange(1, 13, dtype=float))


# ---- SpMV/SpMM ----------------------------------------------------------------------
def test_spmm_matches_dense_brute_force():
    rng = np.random.default_rng(0)
    for k in (1, 2, 3, 5):
        A = synth.random_sparse(37, 45, 0.3, seed=k)
        D = np.zeros((A.m, A.n))
        r, c, v = A.triplets()
        D[r, c] = v
        B = O.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, k)
        X = rng.standard_normal((A.n, 4))
        Y0 = rng.standard_normal((A.m, 4))
        got = B.spmm(X, alpha=0.5, beta=2.0, Y=Y0)
        assert np.allclose(got, 0.5 * D @ X + 2.0 * Y0, rtol=1e-13, atol=1e-13)
        x = X[:, 0].copy()
        assert np.allclose(O.csr_spmv(A.rowptr, A.col, A.val, x), D @ x, rtol=1e-13, atol=1e-13)


def test_spmm_laplacian_closed_form():
    g = 50
    A = synth.laplacian2d(g)
    B = O.Abhsf(A.m, A.n, A.rowptr, A.col, A.val, 4)
    y = B.spmm(np.ones(A.n))
    i, j = np.divmod(np.arange(A.n), g)
    deg = (i > 0).astype(int) + (i < g - 1) + (j > 0) + (j < g - 1)
    assert np.array_equal(y, 4.0 - deg)


def test_planted_mistake_is_caught():
    """A plausible mistake planted in a copy of the cost formula (CSR with s
    instead of s+1 row pointers, or COO without its z-count field) no longer
    reproduces the reference's 3,695,344 bits for the Laplacian at s = 64."""
    A = synth.laplacian2d(256)
    r, c, _ = A.triplets()
    s, lg = 64, O.ceil_log2
    keys, z = np.unique((r // s) * 4096 + c // s, return_counts=True)
    hdr = lg(4096 // s * 0 + 1024) + lg(1024) + 2  # M = N = 1024
    k = lg(s)

    def total(coo_fixed, csr_ptrs):
        coo = 2 * k * z + coo_fixed
        csr = k * z + csr_ptrs * lg(s * s)
        dense = (s * s - z) * 64
        return int(np.sum(hdr + np.minimum(np.minimum(coo, csr), np.minimum(s * s, dense))))

    assert total(lg(s * s), s + 1) == KA["laplacian_256"]["preproc_smf_bits"]["64"]
    assert total(lg(s * s), s) != KA["laplacian_256"]["preproc_smf_bits"]["64"]
    assert total(0, s + 1) != KA["laplacian_256"]["preproc_smf_bits"]["64"]


def test_varint_codec_by_hand():
    """Widths and words of spmv/unified/spmv.cpp:859-896: (delta << 2) | tag,
    little-endian, 1/2/4/8 bytes for delta < 2^6 / 2^14 / 2^30 / more."""
    deltas = [0, 63, 64, 16383, 16384, (1 << 30) - 1, 1 << 30, (1 << 62) - 1]
    want = [[0x00], [0xFC], [0x01, 0x01], [0xFD, 0xFF], [0x02, 0x00, 0x01, 0x00],
            [0xFE, 0xFF, 0xFF, 0xFF], [0x03, 0, 0, 0, 1, 0, 0, 0], [0xFF] * 8]
    for d, w in zip(deltas, want):
        keys = np.array([7, 7 + d], dtype=np.uint64)
        assert list(O.varint_encode(keys)) == w, d
    # a 1-byte word is the delta shifted left by two with tag 0
    keys = np.array([0, 5, 6, 70], dtype=np.uint64)
    assert list(O.varint_encode(keys)) == [5 << 2, 1 << 2, ((64 << 2) | 1) & 0xFF, (64 << 2) >> 8]


def test_varint_codec_roundtrip():
    rng = np.random.default_rng(3)
    for hi in (1 << 5, 1 << 13, 1 << 29, 1 << 40, 1 << 61):
        n = int(min(2000, ((1 << 62) - 1) // hi))  # keys stay below 2^62
        keys = np.cumsum(rng.integers(0, hi, n, dtype=np.int64)).astype(np.uint64)
        s = O.varint_encode(keys)
        got, used = O.varint_decode(s, keys[0], keys.size)
        assert used == s.size and np.array_equal(got, keys)


def _csr(m, entries):
    """CSR of (row, col) entries given in row-major order."""
    rows = [r for r, _ in entries]
    rowptr = np.zeros(m + 1, dtype=np.int64)
    for r in rows:
        rowptr[r + 1] += 1
    return np.cumsum(rowptr), np.array([c for _, c in entries], dtype=np.int32)


def test_delta32_row_major_by_hand():
    """spmv/orig/delta.cpp:124-151: key = (row << 31) + col, deltas from 0.
    Keys by hand: (0,1) 1, (0,5) 5, (1,0) 2^31 = 2147483648, (3,2)
    3*2^31 + 2 = 6442450946, (9,17) 19327352849, (19,19) 40802189331; the
    last three deltas are 2^32 or more (empty rows skipped), so they escape."""
    rp, col = _csr(20, [(0, 1), (0, 5), (1, 0), (3, 2), (9, 17), (19, 19)])
    e = O.delta32_encode(O.DELTA32_ROWMAJOR, 20, rp, col)
    assert e["da"].tolist() == [1, 4, 2147483643, 0, 0, 0]
    assert e["exc"].tolist() == [6442450946, 19327352849, 40802189331]
    assert e["base"].tolist() == [0] and e["xstart"].tolist() == [0, 3]
    assert e["perm"].tolist() == [0, 1, 2, 3, 4, 5]
    r, c = O.delta32_decode(O.DELTA32_ROWMAJOR, e)
    assert r.tolist() == [0, 0, 1, 3, 9, 19] and c.tolist() == [1, 5, 0, 2, 17, 19]


def test_delta32_cache_block_by_hand():
    """spmv/unified/spmv.cpp:636-689: elements ordered by (row/8, col/8,
    row%8, col%8), key ((row/8) << 31) + ((col/8) << 6) + ((row%8) << 3) +
    col%8.  By hand: (0,1) 1, (1,0) 8, (7,7) 63, (0,9) 64 + 1 = 65 (block
    column 1), (8,0) 2^31; so (0,9) moves behind (1,0) and (7,7)."""
    rp, col = _csr(9, [(0, 1), (0, 9), (1, 0), (7, 7), (8, 0)])
    e = O.delta32_encode(O.DELTA32_CACHEBLOCK, 9, rp, col)
    assert e["perm"].tolist() == [0, 2, 3, 1, 4]
    assert e["da"].tolist() == [1, 7, 55, 2, 2147483583]
    assert e["exc"].size == 0
    r, c = O.delta32_decode(O.DELTA32_CACHEBLOCK, e)
    assert r.tolist() == [0, 1, 7, 0, 8] and c.tolist() == [1, 0, 7, 9, 0]
    # (9,17) -> block row 1, block column 2, local (1,1): 2^31 + 128 + 8 + 1
    rp, col = _csr(10, [(9, 17)])
    assert O.delta32_encode(O.DELTA32_CACHEBLOCK, 10, rp, col)["da"].tolist() == [2147483785]


def test_delta32_chunk_first_escape_by_hand():
    """Chunks of 2: a chunk's first element whose delta needs 33+ bits puts
    its key into the chunk base (delta 0); any other element escapes."""
    rp, col = _csr(10, [(0, 0), (0, 1), (5, 0), (9, 3)])
    e = O.delta32_encode(O.DELTA32_ROWMAJOR, 10, rp, col, chunk=2)
    assert e["da"].tolist() == [0, 1, 0, 0]
    assert e["base"].tolist() == [0, 5 * 2 ** 31]
    assert e["xstart"].tolist() == [0, 0, 1]
    assert e["exc"].tolist() == [9 * 2 ** 31 + 3]
    r, c = O.delta32_decode(O.DELTA32_ROWMAJOR, e, chunk=2)
    assert r.tolist() == [0, 0, 5, 9] and c.tolist() == [0, 1, 0, 3]


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("chunk", [512, 7])
def test_delta32_roundtrip_and_order(layout, chunk):
    from abhsf_amd import synth
    for A in (synth.rmat(12, 8, seed=1), synth.laplacian2d(40), synth.random_sparse(300, 5000, 0.01, seed=2)):
        e = O.delta32_encode(layout, A.m, A.rowptr, A.col, chunk)
        r, c = O.delta32_decode(layout, e, chunk)
        rr = np.repeat(np.arange(A.m), np.diff(A.rowptr))
        p = e["perm"].astype(np.int64)
        assert np.array_equal(r, rr[p]) and np.array_equal(c, A.col[p].astype(np.int64))
        assert np.array_equal(np.sort(p), np.arange(A.nnz))  # a permutation
        # order: row-major, or (row/8, col/8, row%8, col%8) lexicographic
        key = (r, c) if layout == 1 else (r // 8, c // 8, r % 8, c % 8)
        order = np.lexsort(key[::-1])
        assert np.array_equal(order, np.arange(A.nnz))
