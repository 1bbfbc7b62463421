/*
 * abhsf_oracle.c -- CPU restatement of the reference's ABHSF algorithms.
 *
 * TEST INFRASTRUCTURE ONLY (see abhsf_oracle.h).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Pinning (what each function is checked against, in tests/):
 *   orc_ceil_log2      brute force against the definition for 1..2^20 and both
 *                      reference loops (math.h:17-31, test.cpp:88-98).
 *   orc_choose         closed-form crossover points solved by hand; App. B
 *                      footprints of preprocessing/test.cpp (test.mtx s=2..1024,
 *                      256^2 Laplacian s=8..256).
 *   orc_smf_preproc    the same known answers (output of preprocessing/test.cpp
 *                      recorded in tests/golden/known_answers.json; its AQsort
 *                      dependency is absent here, so test.cpp itself cannot be
 *                      rebuilt).
 *   orc_block_hist     oracle/_ref/block_stats (the reference's own binary,
 *                      compiled from block_statistics/block_stats.cpp) on
 *                      test.mtx and seeded random matrices; sum z*count = nnz.
 *   orc_mmf_row        oracle/_ref/mmf (block_statistics/mmf.cpp) on the same
 *                      inputs; App. B rows "2 2" and "4 4".
 *   orc_build/decode   round trip to the input triplets; popcount(bitmap) = z;
 *                      hand-built 4x4/8x8 layouts (spmv/orig/88.cpp:103-130).
 *   orc_spmm           orc_csr_spmv (textbook) and closed forms (Laplacian row
 *                      sums, 88.cpp y-sum); brute-force dense product on tiny
 *                      matrices.
 *   orc_rows_stdev     App. B "0.763763" for test.mtx; two-pass formula.
 *   orc_varint_*       hand-encoded words at every width boundary; round trip.
 *   orc_delta32_*      keys of hand-placed elements worked out by hand for both
 *                      layouts (tests/test_oracle.py), the element order of a
 *                      hand-listed 16 x 16 case, escapes at a known row gap,
 *                      round trip on seeded matrices.
 */
#include "abhsf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* preprocessing/test.cpp:88-98: count significant bits, minus one for powers of 2. */
uint64_t orc_ceil_log2(uint64_t x)
{
    if (x == 0) return 0; /* the reference asserts x > 0 */
    uint64_t e = 0;
    while (e < 64 && (x >> e) != 0) e++;
    if ((x & (x - 1)) == 0) e--;
    return e;
}

int orc_choose(int policy, uint64_t s, uint64_t z, uint64_t b, int is_binary, uint64_t *bits)
{
    uint64_t ss = s * s;
    uint64_t lg_s = orc_ceil_log2(s), lg_ss = orc_ceil_log2(ss);
    uint64_t coo, csr;
    if (policy == ORC_POLICY_PREPROC) {
        /* preprocessing/test.cpp:48-51,59-62 */
        coo = 2 * lg_s * z + lg_ss;
        csr = lg_s * z + (s + 1) * lg_ss;
    } else {
        /* include/abhsf/msmf.h:9-21 with m = n = s */
        coo = lg_ss + z * (lg_s + lg_s);
        csr = s * lg_ss + z * lg_s;
    }
    uint64_t bitmap = ss; /* test.cpp:66, msmf.h:23-26 */
    int scheme = ORC_COO;
    uint64_t min = coo;
    if (csr < min) { min = csr; scheme = ORC_CSR; }
    if (bitmap < min) { min = bitmap; scheme = ORC_BITMAP; }
    if (!is_binary) { /* test.cpp:70-74, msmf.h:28-33 */
        uint64_t dense = (ss - z) * b;
        if (dense < min) { min = dense; scheme = ORC_DENSE; }
    }
    if (bits) *bits = min;
    return scheme;
}

/* ---- LSD radix sort of (key, payload) u64 pairs, stable ------------------ */
static void radix_sort_u64(uint64_t *key, uint64_t *pay, uint64_t n, int key_bits)
{
    if (n < 2) return;
    uint64_t *k2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *p2 = pay ? (uint64_t *)malloc(n * sizeof(uint64_t)) : NULL;
    uint64_t *ka = key, *kb = k2, *pa = pay, *pb = p2;
    for (int shift = 0; shift < key_bits; shift += 11) {
        uint64_t cnt[2048];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; i++) cnt[(ka[i] >> shift) & 2047]++;
        uint64_t sum = 0;
        for (int d = 0; d < 2048; d++) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (uint64_t i = 0; i < n; i++) {
            uint64_t d = (ka[i] >> shift) & 2047;
            uint64_t o = cnt[d]++;
            kb[o] = ka[i];
            if (pa) pb[o] = pa[i];
        }
        uint64_t *t = ka; ka = kb; kb = t;
        t = pa; pa = pb; pb = t;
    }
    if (ka != key) {
        memcpy(key, ka, n * sizeof(uint64_t));
        if (pay) memcpy(pay, pa, n * sizeof(uint64_t));
    }
    free(k2);
    free(p2);
}

static int bits_for(uint64_t maxval)
{
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) b++;
    return b;
}

/* Block key (I,J) with I = row >> kr, J = col >> ks, ordered by I then J
 * (block_stats.cpp:69-74, preprocessing/test.cpp:243-252). */
static uint64_t *block_keys(uint64_t nnz, const uint32_t *rows, const uint32_t *cols, int kr, int ks,
                            int *key_bits)
{
    uint64_t *key = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    uint32_t maxr = 0, maxc = 0;
    for (uint64_t i = 0; i < nnz; i++) {
        if (rows[i] > maxr) maxr = rows[i];
        if (cols[i] > maxc) maxc = cols[i];
    }
    int cb = bits_for((uint64_t)(maxc >> ks));
    int rb = bits_for((uint64_t)(maxr >> kr));
    for (uint64_t i = 0; i < nnz; i++)
        key[i] = ((uint64_t)(rows[i] >> kr) << cb) | (uint64_t)(cols[i] >> ks);
    *key_bits = cb + rb;
    return key;
}

uint64_t orc_smf_preproc(uint64_t m, uint64_t n, uint64_t nnz, const uint32_t *rows,
                         const uint32_t *cols, uint64_t s, uint64_t b, int is_binary,
                         uint64_t nblk_by_scheme[4])
{
    /* preprocessing/test.cpp:39-53 */
    uint64_t M = m / s + (m % s ? 1 : 0);
    uint64_t N = n / s + (n % s ? 1 : 0);
    uint64_t hdr = orc_ceil_log2(M) + orc_ceil_log2(N) + 2;
    int ks = (int)orc_ceil_log2(s);
    int kb;
    uint64_t *key = block_keys(nnz, rows, cols, ks, ks, &kb);
    radix_sort_u64(key, NULL, nnz, kb);
    uint64_t total = 0;
    if (nblk_by_scheme) memset(nblk_by_scheme, 0, 4 * sizeof(uint64_t));
    /* test.cpp:346-361: one processor call per maximal run of equal (I,J) */
    uint64_t i = 0;
    while (i < nnz) {
        uint64_t j = i + 1;
        while (j < nnz && key[j] == key[i]) j++;
        uint64_t bits;
        int sc = orc_choose(ORC_POLICY_PREPROC, s, j - i, b, is_binary, &bits);
        total += hdr + bits;
        if (nblk_by_scheme) nblk_by_scheme[sc]++;
        i = j;
    }
    free(key);
    return total;
}

uint64_t orc_block_hist(uint64_t nnz, const uint32_t *rows, const uint32_t *cols, int kr, int ks,
                        uint64_t *z_out, uint64_t *count_out, uint64_t cap)
{
    int kb;
    uint64_t *key = block_keys(nnz, rows, cols, kr, ks, &kb);
    radix_sort_u64(key, NULL, nnz, kb);
    /* run lengths (block_stats.cpp:88-99) accumulated into a z -> count map */
    uint64_t maxz = 0;
    uint64_t i = 0;
    while (i < nnz) {
        uint64_t j = i + 1;
        while (j < nnz && key[j] == key[i]) j++;
        if (j - i > maxz) maxz = j - i;
        i = j;
    }
    uint64_t *h = (uint64_t *)calloc(maxz + 1, sizeof(uint64_t));
    i = 0;
    while (i < nnz) {
        uint64_t j = i + 1;
        while (j < nnz && key[j] == key[i]) j++;
        h[j - i]++;
        i = j;
    }
    uint64_t nd = 0;
    for (uint64_t z = 1; z <= maxz; z++) {
        if (!h[z]) continue;
        if (nd < cap) { z_out[nd] = z; count_out[nd] = h[z]; }
        nd++;
    }
    free(h);
    free(key);
    return nd;
}

static uint64_t ceil_div_u64(uint64_t a, uint64_t b) { return a / b + (a % b ? 1 : 0); }

/* include/abhsf/msmf.h:9-63 for an r x s block */
static uint64_t msmf_coo(uint64_t r, uint64_t s, uint64_t z) { return orc_ceil_log2(r * s) + z * (orc_ceil_log2(r) + orc_ceil_log2(s)); }
static uint64_t msmf_csr(uint64_t r, uint64_t s, uint64_t z) { return r * orc_ceil_log2(r * s) + z * orc_ceil_log2(s); }
static uint64_t msmf_min(uint64_t r, uint64_t s, uint64_t z, uint64_t b, int bin, int with_csr)
{
    uint64_t m = msmf_coo(r, s, z);
    if (with_csr) { uint64_t c = msmf_csr(r, s, z); if (c < m) m = c; }
    if (r * s < m) m = r * s;
    if (!bin) { uint64_t d = (r * s - z) * b; if (d < m) m = d; }
    return m;
}

void orc_mmf_row(uint64_t m, uint64_t n, uint64_t nnz, int is_binary, uint64_t bits_single, int mult,
                 uint64_t r, uint64_t s, uint64_t nh, const uint64_t *hz, const uint64_t *hc,
                 uint64_t out[8])
{
    /* block_statistics/mmf.cpp:53-102 */
    uint64_t coo32 = nnz * 32 * 2;
    uint64_t csr32 = (m + 1 + nnz) * 32;
    uint64_t M = ceil_div_u64(m, r), N = ceil_div_u64(n, s);
    uint64_t NNZ = 0;
    for (uint64_t i = 0; i < nh; i++) NNZ += hc[i];
    uint64_t csr0 = M * orc_ceil_log2(N + 1) + NNZ * orc_ceil_log2(N);
    uint64_t b = bits_single * (uint64_t)mult;
    uint64_t coo = csr0, csr = csr0, bm = csr0, dense = csr0, ab = csr0, ab134 = csr0;
    if (is_binary) dense = UINT64_MAX;
    for (uint64_t i = 0; i < nh; i++) {
        uint64_t z = hz[i], c = hc[i];
        coo += msmf_coo(r, s, z) * c;
        csr += msmf_csr(r, s, z) * c;
        bm += r * s * c;
        if (!is_binary) dense += (r * s - z) * b * c;
        ab += (msmf_min(r, s, z, b, is_binary, 1) + 2) * c;
        ab134 += (msmf_min(r, s, z, b, is_binary, 0) + 2) * c;
    }
    /* mmf.cpp:119-138 */
    uint64_t valf = nnz * b;
    out[0] = coo32 + valf; out[1] = csr32 + valf; out[2] = coo + valf; out[3] = csr + valf;
    out[4] = bm + valf; out[5] = dense + valf; out[6] = ab + valf; out[7] = ab134 + valf;
}

double orc_rows_stdev(uint64_t m, const uint64_t *rows_nnz)
{
    /* rows-nnz-stdev.cpp:11-30: population standard deviation */
    uint64_t nnz = 0;
    for (uint64_t k = 0; k < m; k++) nnz += rows_nnz[k];
    double avg = (double)nnz / (double)m, acc = 0.0;
    for (uint64_t k = 0; k < m; k++) { double d = (double)rows_nnz[k] - avg; acc += d * d; }
    return sqrt(acc / (double)m);
}

/* ---- ABHSF storage builder ------------------------------------------------ */

static void *xcalloc(uint64_t n, uint64_t sz) { return calloc(n ? n : 1, sz); }

void orc_free(orc_abhsf *A)
{
    for (int x = 0; x < 4; x++) {
        free(A->sptr[x]); free(A->eptr[x]); free(A->bcol[x]); free(A->z1[x]); free(A->val[x]);
        A->sptr[x] = A->eptr[x] = NULL; A->bcol[x] = A->val[x] = NULL; A->z1[x] = NULL;
    }
    free(A->coo_idx); free(A->csr_rowptr); free(A->csr_col); free(A->bm_words);
    A->coo_idx = A->csr_rowptr = NULL; A->csr_col = NULL; A->bm_words = NULL;
}

static void put_uint(void *base, int bytes, uint64_t i, uint64_t v)
{
    if (bytes == 1) ((uint8_t *)base)[i] = (uint8_t)v;
    else if (bytes == 2) ((uint16_t *)base)[i] = (uint16_t)v;
    else ((uint32_t *)base)[i] = (uint32_t)v;
}
static uint64_t get_uint(const void *base, int bytes, uint64_t i)
{
    if (bytes == 1) return ((const uint8_t *)base)[i];
    if (bytes == 2) return ((const uint16_t *)base)[i];
    return ((const uint32_t *)base)[i];
}
static void put_val(void *base, int bytes, uint64_t i, double v)
{
    if (bytes == 8) ((double *)base)[i] = v;
    else ((float *)base)[i] = (float)v;
}
static double get_val(const void *base, int bytes, uint64_t i)
{
    return bytes == 8 ? ((const double *)base)[i] : (double)((const float *)base)[i];
}

int orc_build(orc_abhsf *A, int64_t m, int64_t n, const int64_t *rowptr, const int32_t *col,
              const double *val, int k, int policy, int val_bytes, int is_binary)
{
    memset(A, 0, sizeof(*A));
    if (k < 1 || k > 8 || m < 1 || n < 1 || m > INT32_MAX || n > INT32_MAX) return 1;
    if (val_bytes != 8 && val_bytes != 4) return 1;
    if (policy != ORC_POLICY_PREPROC && policy != ORC_POLICY_MSMF) return 1;
    const int64_t s = (int64_t)1 << k;
    uint64_t nnz = (uint64_t)rowptr[m];
    if (rowptr[0] != 0 || nnz > UINT32_MAX) return 1;
    for (int64_t r = 0; r < m; r++) {
        if (rowptr[r + 1] < rowptr[r]) return 1;
        for (int64_t e = rowptr[r]; e < rowptr[r + 1]; e++) {
            if (col[e] < 0 || col[e] >= n) return 1;
            if (e > rowptr[r] && col[e] <= col[e - 1]) return 1;
        }
    }
    A->m = m; A->n = n; A->s = (int32_t)s; A->k = k;
    A->M = (m + s - 1) / s; A->N = (n + s - 1) / s;
    A->policy = policy; A->is_binary = is_binary; A->val_bytes = val_bytes;
    A->cost_b = is_binary ? 0 : (uint64_t)val_bytes * 8;
    A->bcol_bytes = A->N <= 65536 ? 2 : 4;
    A->coo_idx_bytes = s <= 16 ? 1 : 2;
    A->csr_rp_bytes = s <= 8 ? 1 : 2;
    A->bm_nw = (int32_t)(s * s / 64 > 0 ? s * s / 64 : 1);
    A->nnz = nnz;
    const int64_t M = A->M;
    const uint64_t ss = (uint64_t)(s * s);

    /* Pass 1: order each strip's entries by (J, lrow, lcol).  A stable sort by J
     * of the strip's CSR-ordered entries gives exactly that order
     * (spmv/orig/spmv.cpp:85-110 is the full comparator). */
    uint64_t *perm = (uint64_t *)xcalloc(nnz, sizeof(uint64_t)); /* element ids in block order */
    uint64_t nblk_cap = 1024, nblk = 0;
    uint32_t *bJ = (uint32_t *)malloc(nblk_cap * 4);
    uint32_t *bZ = (uint32_t *)malloc(nblk_cap * 4);
    uint64_t *bStart = (uint64_t *)malloc(nblk_cap * 8);
    uint8_t *bS = (uint8_t *)malloc(nblk_cap);
    int64_t *stripBlk = (int64_t *)xcalloc((uint64_t)M + 1, sizeof(int64_t));
    uint64_t cap = 0;
    uint64_t *key = NULL, *pay = NULL;
    for (int64_t I = 0; I < M; I++) {
        int64_t r0 = I * s, r1 = r0 + s < m ? r0 + s : m;
        uint64_t e0 = (uint64_t)rowptr[r0], e1 = (uint64_t)rowptr[r1], cnt = e1 - e0;
        if (cnt > cap) {
            cap = cnt;
            free(key); free(pay);
            key = (uint64_t *)malloc(cap * 8);
            pay = (uint64_t *)malloc(cap * 8);
        }
        for (uint64_t e = e0; e < e1; e++) {
            key[e - e0] = (uint64_t)(col[e] >> k);
            pay[e - e0] = e;
        }
        radix_sort_u64(key, pay, cnt, bits_for((uint64_t)(A->N)));
        stripBlk[I] = (int64_t)nblk;
        uint64_t i = 0;
        while (i < cnt) {
            uint64_t j = i + 1;
            while (j < cnt && key[j] == key[i]) j++;
            if (nblk == nblk_cap) {
                nblk_cap *= 2;
                bJ = (uint32_t *)realloc(bJ, nblk_cap * 4);
                bZ = (uint32_t *)realloc(bZ, nblk_cap * 4);
                bStart = (uint64_t *)realloc(bStart, nblk_cap * 8);
                bS = (uint8_t *)realloc(bS, nblk_cap);
            }
            bJ[nblk] = (uint32_t)key[i];
            bZ[nblk] = (uint32_t)(j - i);
            bStart[nblk] = e0 + i;
            bS[nblk] = (uint8_t)orc_choose(policy, (uint64_t)s, j - i, A->cost_b, is_binary, NULL);
            nblk++;
            i = j;
        }
        memcpy(perm + e0, pay, cnt * sizeof(uint64_t));
    }
    stripBlk[M] = (int64_t)nblk;
    free(key); free(pay);

    /* Pass 2: per-scheme counts and strip offsets */
    for (int x = 0; x < 4; x++) {
        A->sptr[x] = (uint32_t *)xcalloc((uint64_t)M + 1, 4);
        A->eptr[x] = (uint32_t *)xcalloc((uint64_t)M + 1, 4);
    }
    uint64_t nb[4] = {0, 0, 0, 0}, ne[4] = {0, 0, 0, 0};
    for (int64_t I = 0; I < M; I++) {
        for (int x = 0; x < 4; x++) { A->sptr[x][I] = (uint32_t)nb[x]; A->eptr[x][I] = (uint32_t)ne[x]; }
        for (int64_t b = stripBlk[I]; b < stripBlk[I + 1]; b++) {
            int x = bS[b];
            nb[x]++;
            ne[x] += (x == ORC_DENSE) ? ss : bZ[b];
        }
    }
    for (int x = 0; x < 4; x++) {
        A->sptr[x][M] = (uint32_t)nb[x];
        A->eptr[x][M] = (uint32_t)ne[x];
        A->nb[x] = nb[x];
        A->ne[x] = ne[x];
        if (ne[x] > UINT32_MAX) { orc_free(A); return 1; }
        A->bcol[x] = xcalloc(nb[x], (uint64_t)A->bcol_bytes);
        A->z1[x] = (uint16_t *)xcalloc(nb[x], 2);
        A->val[x] = xcalloc(ne[x], (uint64_t)val_bytes);
    }
    A->coo_idx = xcalloc(ne[ORC_COO], (uint64_t)A->coo_idx_bytes);
    A->csr_rowptr = xcalloc(nb[ORC_CSR] * (uint64_t)(s + 1), (uint64_t)A->csr_rp_bytes);
    A->csr_col = (uint8_t *)xcalloc(ne[ORC_CSR], 1);
    A->bm_words = (uint64_t *)xcalloc(nb[ORC_BITMAP] * (uint64_t)A->bm_nw, 8);

    /* Pass 3: emit (88.cpp:103-130 layouts; bit lrow*s+lcol LSB-first) */
    uint64_t ob[4] = {0, 0, 0, 0}, oe[4] = {0, 0, 0, 0};
    uint64_t smf = 0;
    uint64_t hdr = orc_ceil_log2((uint64_t)A->M) + orc_ceil_log2((uint64_t)A->N) + 2;
    uint32_t *rowcnt = (uint32_t *)malloc((size_t)(s + 1) * 4);
    /* row of every element: rowof[e] = r for rowptr[r] <= e < rowptr[r + 1] */
    uint32_t *rowof = (uint32_t *)xcalloc(nnz > 0 ? nnz : 1, 4);
    for (int64_t r = 0; r < m; r++)
        for (int64_t e = rowptr[r]; e < rowptr[r + 1]; e++) rowof[e] = (uint32_t)r;
    for (uint64_t b = 0; b < nblk; b++) {
        int x = bS[b];
        uint64_t z = bZ[b], o = ob[x]++, e = oe[x];
        uint64_t bits;
        orc_choose(policy, (uint64_t)s, z, A->cost_b, is_binary, &bits);
        smf += (policy == ORC_POLICY_PREPROC) ? hdr + bits : bits + 2;
        put_uint(A->bcol[x], A->bcol_bytes, o, bJ[b]);
        A->z1[x][o] = (uint16_t)(z - 1);
        if (x == ORC_CSR) memset(rowcnt, 0, (size_t)(s + 1) * 4);
        for (uint64_t q = 0; q < z; q++) {
            uint64_t el = perm[bStart[b] + q];
            uint64_t lrow = (uint64_t)rowof[el] & (uint64_t)(s - 1), lcol = (uint64_t)col[el] & (uint64_t)(s - 1);
            double v = val ? val[el] : 1.0;
            switch (x) {
            case ORC_COO:
                put_uint(A->coo_idx, A->coo_idx_bytes, e + q, (lrow << k) | lcol);
                put_val(A->val[x], val_bytes, e + q, v);
                break;
            case ORC_CSR:
                rowcnt[lrow + 1]++;
                A->csr_col[e + q] = (uint8_t)lcol;
                put_val(A->val[x], val_bytes, e + q, v);
                break;
            case ORC_BITMAP: {
                uint64_t p = lrow * (uint64_t)s + lcol;
                A->bm_words[o * (uint64_t)A->bm_nw + (p >> 6)] |= (uint64_t)1 << (p & 63);
                put_val(A->val[x], val_bytes, e + q, v);
                break;
            }
            default:
                put_val(A->val[x], val_bytes, e + lrow * (uint64_t)s + lcol, v);
            }
        }
        if (x == ORC_CSR) {
            uint64_t acc = 0;
            for (int64_t r = 0; r <= s; r++) {
                acc += rowcnt[r];
                put_uint(A->csr_rowptr, A->csr_rp_bytes, o * (uint64_t)(s + 1) + (uint64_t)r, acc);
            }
        }
        oe[x] += (x == ORC_DENSE) ? ss : z;
    }
    if (policy == ORC_POLICY_MSMF)
        smf += (uint64_t)A->M * orc_ceil_log2((uint64_t)A->N + 1) + nblk * orc_ceil_log2((uint64_t)A->N);
    A->smf_bits = smf;
    free(rowcnt);
    free(rowof);
    free(perm); free(bJ); free(bZ); free(bStart); free(bS); free(stripBlk);
    return 0;
}

void orc_spmm(const orc_abhsf *A, int nrhs, double alpha, const double *X, int64_t ldx, double beta,
              double *Y, int64_t ldy)
{
    const int64_t s = A->s, M = A->M;
    const int k = A->k, vb = A->val_bytes, W = A->bm_nw;
#pragma omp parallel
    {
        double *acc = (double *)malloc((size_t)(s * nrhs) * sizeof(double));
#pragma omp for schedule(dynamic, 16)
        for (int64_t I = 0; I < M; I++) {
            memset(acc, 0, (size_t)(s * nrhs) * sizeof(double));
            /* COO group: y[lrow] += v * x[bcol*s + lcol]  (88.cpp:214-235) */
            for (uint32_t b = A->sptr[0][I], e = A->eptr[0][I]; b < A->sptr[0][I + 1]; b++) {
                uint64_t J = get_uint(A->bcol[0], A->bcol_bytes, b), z = (uint64_t)A->z1[0][b] + 1;
                for (uint64_t q = 0; q < z; q++, e++) {
                    uint64_t id = get_uint(A->coo_idx, A->coo_idx_bytes, e);
                    uint64_t lr = id >> k, lc = id & (uint64_t)(s - 1);
                    double v = get_val(A->val[0], vb, e);
                    const double *xr = X + (int64_t)(J * (uint64_t)s + lc) * ldx;
                    for (int j = 0; j < nrhs; j++) acc[lr * nrhs + j] += v * xr[j];
                }
            }
            /* CSR group */
            for (uint32_t b = A->sptr[1][I], e = A->eptr[1][I]; b < A->sptr[1][I + 1]; b++) {
                uint64_t J = get_uint(A->bcol[1], A->bcol_bytes, b);
                for (int64_t r = 0; r < s; r++) {
                    uint64_t p0 = get_uint(A->csr_rowptr, A->csr_rp_bytes, (uint64_t)b * (s + 1) + r);
                    uint64_t p1 = get_uint(A->csr_rowptr, A->csr_rp_bytes, (uint64_t)b * (s + 1) + r + 1);
                    for (uint64_t q = p0; q < p1; q++) {
                        uint64_t lc = A->csr_col[e + q];
                        double v = get_val(A->val[1], vb, e + q);
                        const double *xr = X + (int64_t)(J * (uint64_t)s + lc) * ldx;
                        for (int j = 0; j < nrhs; j++) acc[r * nrhs + j] += v * xr[j];
                    }
                }
                e += A->z1[1][b] + 1u;
            }
            /* bitmap group: set bits in order, values packed (88_bitmap.cpp:183-207) */
            for (uint32_t b = A->sptr[2][I], e = A->eptr[2][I]; b < A->sptr[2][I + 1]; b++) {
                uint64_t J = get_uint(A->bcol[2], A->bcol_bytes, b);
                const uint64_t *w = A->bm_words + (uint64_t)b * W;
                for (uint64_t p = 0; p < (uint64_t)(s * s); p++) {
                    if (!((w[p >> 6] >> (p & 63)) & 1)) continue;
                    uint64_t lr = p >> k, lc = p & (uint64_t)(s - 1);
                    double v = get_val(A->val[2], vb, e++);
                    const double *xr = X + (int64_t)(J * (uint64_t)s + lc) * ldx;
                    for (int j = 0; j < nrhs; j++) acc[lr * nrhs + j] += v * xr[j];
                }
            }
            /* dense group: row-major s x s GEMV (88.cpp:587-604) */
            for (uint32_t b = A->sptr[3][I]; b < A->sptr[3][I + 1]; b++) {
                uint64_t J = get_uint(A->bcol[3], A->bcol_bytes, b);
                uint64_t e = (uint64_t)b * (uint64_t)(s * s);
                for (int64_t r = 0; r < s; r++)
                    for (int64_t c = 0; c < s; c++) {
                        int64_t gc = (int64_t)J * s + c;
                        if (gc >= A->n) continue;
                        double v = get_val(A->val[3], vb, e + (uint64_t)(r * s + c));
                        const double *xr = X + gc * ldx;
                        for (int j = 0; j < nrhs; j++) acc[r * nrhs + j] += v * xr[j];
                    }
            }
            for (int64_t r = 0; r < s; r++) {
                int64_t gr = I * s + r;
                if (gr >= A->m) break;
                for (int j = 0; j < nrhs; j++) {
                    double *yp = Y + gr * ldy + j;
                    *yp = (beta == 0.0) ? alpha * acc[r * nrhs + j] : alpha * acc[r * nrhs + j] + beta * *yp;
                }
            }
        }
        free(acc);
    }
}

void orc_csr_spmv(int64_t m, const int64_t *rowptr, const int32_t *col, const double *val,
                  double alpha, const double *x, double beta, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < m; r++) {
        double acc = 0.0;
        for (int64_t e = rowptr[r]; e < rowptr[r + 1]; e++) acc += val[e] * x[col[e]];
        y[r] = (beta == 0.0) ? alpha * acc : alpha * acc + beta * y[r];
    }
}

uint64_t orc_decode(const orc_abhsf *A, int64_t *rows, int64_t *cols, double *vals, uint64_t cap)
{
    const int64_t s = A->s;
    const int k = A->k, vb = A->val_bytes, W = A->bm_nw;
    uint64_t o = 0;
#define EMIT(R, C, V)                                                                          \
    do {                                                                                       \
        if (o < cap) { rows[o] = (R); cols[o] = (C); vals[o] = (V); }                          \
        o++;                                                                                   \
    } while (0)
    for (int64_t I = 0; I < A->M; I++) {
        /* merge the four scheme groups back into (J) order */
        uint32_t b[4], e[4];
        for (int x = 0; x < 4; x++) { b[x] = A->sptr[x][I]; e[x] = A->eptr[x][I]; }
        for (;;) {
            int best = -1;
            uint64_t bestJ = 0;
            for (int x = 0; x < 4; x++) {
                if (b[x] >= A->sptr[x][I + 1]) continue;
                uint64_t J = get_uint(A->bcol[x], A->bcol_bytes, b[x]);
                if (best < 0 || J < bestJ) { best = x; bestJ = J; }
            }
            if (best < 0) break;
            int x = best;
            uint64_t z = (uint64_t)A->z1[x][b[x]] + 1;
            int64_t R0 = I * s, C0 = (int64_t)bestJ * s;
            if (x == ORC_COO) {
                for (uint64_t q = 0; q < z; q++) {
                    uint64_t id = get_uint(A->coo_idx, A->coo_idx_bytes, e[x] + q);
                    EMIT(R0 + (int64_t)(id >> k), C0 + (int64_t)(id & (uint64_t)(s - 1)), get_val(A->val[x], vb, e[x] + q));
                }
                e[x] += (uint32_t)z;
            } else if (x == ORC_CSR) {
                for (int64_t r = 0; r < s; r++) {
                    uint64_t p0 = get_uint(A->csr_rowptr, A->csr_rp_bytes, (uint64_t)b[x] * (s + 1) + r);
                    uint64_t p1 = get_uint(A->csr_rowptr, A->csr_rp_bytes, (uint64_t)b[x] * (s + 1) + r + 1);
                    for (uint64_t q = p0; q < p1; q++)
                        EMIT(R0 + r, C0 + A->csr_col[e[x] + q], get_val(A->val[x], vb, e[x] + q));
                }
                e[x] += (uint32_t)z;
            } else if (x == ORC_BITMAP) {
                const uint64_t *w = A->bm_words + (uint64_t)b[x] * W;
                for (uint64_t p = 0; p < (uint64_t)(s * s); p++)
                    if ((w[p >> 6] >> (p & 63)) & 1) {
                        EMIT(R0 + (int64_t)(p >> k), C0 + (int64_t)(p & (uint64_t)(s - 1)), get_val(A->val[x], vb, e[x]));
                        e[x]++;
                    }
            } else {
                uint64_t base = (uint64_t)b[x] * (uint64_t)(s * s);
                for (int64_t r = 0; r < s; r++)
                    for (int64_t c = 0; c < s; c++) {
                        double v = get_val(A->val[x], vb, base + (uint64_t)(r * s + c));
                        if (v != 0.0 && R0 + r < A->m && C0 + c < A->n) EMIT(R0 + r, C0 + c, v);
                    }
            }
            b[x]++;
        }
    }
#undef EMIT
    return o;
}


/* ---- tagged varint delta codec (spmv/unified/spmv.cpp:859-896, 923-944) ---- */

uint64_t orc_varint_encode(uint64_t nnz, const uint64_t *keys, uint8_t *out)
{
    uint64_t pos = 0;
    for (uint64_t k = 1; k < nnz; k++) {
        const uint64_t delta = keys[k] - keys[k - 1];
        uint64_t word;
        int bytes;
        if (delta < (1ull << 6)) { word = delta << 2; bytes = 1; }
        else if (delta < (1ull << 14)) { word = (delta << 2) | 1u; bytes = 2; }
        else if (delta < (1ull << 30)) { word = (delta << 2) | 2u; bytes = 4; }
        else { word = (delta << 2) | 3u; bytes = 8; }
        if (out)
            for (int b = 0; b < bytes; b++) out[pos + (uint64_t)b] = (uint8_t)(word >> (8 * b));
        pos += (uint64_t)bytes;
    }
    return pos;
}

uint64_t orc_varint_decode(uint64_t nnz, uint64_t first, const uint8_t *in, uint64_t *keys)
{
    uint64_t pos = 0;
    if (nnz == 0) return 0;
    keys[0] = first;
    for (uint64_t k = 1; k < nnz; k++) {
        const int tag = in[pos] & 3;
        const int bytes = 1 << tag;
        uint64_t word = 0;
        for (int b = 0; b < bytes; b++) word |= (uint64_t)in[pos + (uint64_t)b] << (8 * b);
        keys[k] = keys[k - 1] + (word >> 2);
        pos += (uint64_t)bytes;
    }
    return pos;
}


/* ---- u32 delta codecs (spmv/orig/delta.cpp:124-151; spmv/unified/spmv.cpp:636-689) ---- */

static uint64_t d32_key(int layout, uint64_t row, uint64_t col)
{
    if (layout == 1) return (row << 31) + col;
    return ((row >> 3) << 31) + ((col >> 3) << 6) + ((row & 7) << 3) + (col & 7);
}

static void d32_rowcol(int layout, uint64_t key, int64_t *row, int64_t *col)
{
    if (layout == 1) {
        *row = (int64_t)(key >> 31);
        *col = (int64_t)(key & ((1ull << 31) - 1));
    } else {
        *row = (int64_t)(((key >> 31) << 3) | ((key >> 3) & 7));
        *col = (int64_t)((((key >> 6) & ((1ull << 25) - 1)) << 3) | (key & 7));
    }
}

uint64_t orc_delta32_encode(int layout, int64_t m, const int64_t *rowptr, const int32_t *col, uint64_t chunk,
                            uint32_t *da, uint64_t *perm, uint64_t *base, uint32_t *xstart, uint64_t *exc)
{
    const uint64_t nnz = (uint64_t)rowptr[m];
    uint64_t *key = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    uint64_t *pay = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    for (int64_t r = 0; r < m; r++)
        for (int64_t e = rowptr[r]; e < rowptr[r + 1]; e++) {
            key[e] = d32_key(layout, (uint64_t)r, (uint64_t)col[e]);
            pay[e] = (uint64_t)e;
        }
    /* layout 1 keys are already ascending in CSR order; layout 2 is sorted by
       key, which orders (row/8, col/8, row%8, col%8) lexicographically */
    if (layout != 1) radix_sort_u64(key, pay, nnz, 64);
    uint64_t prev = 0, nx = 0;
    for (uint64_t e = 0; e < nnz; e++) {
        const uint64_t d = key[e] - prev;
        if (e % chunk == 0) {
            /* a chunk's first element never escapes: its base absorbs a
               delta too wide for 32 bits */
            base[e / chunk] = d < (1ull << 32) ? prev : key[e];
            da[e] = d < (1ull << 32) ? (uint32_t)d : 0u;
            xstart[e / chunk] = (uint32_t)nx;
        } else if (d >= (1ull << 32)) {
            da[e] = 0u;
            exc[nx++] = key[e];
        } else {
            da[e] = (uint32_t)d;
        }
        perm[e] = pay[e];
        prev = key[e];
    }
    xstart[(nnz + chunk - 1) / chunk] = (uint32_t)nx;
    free(key);
    free(pay);
    return nx;
}

void orc_delta32_decode(int layout, uint64_t nnz, uint64_t chunk, const uint32_t *da, const uint64_t *base,
                        const uint32_t *xstart, const uint64_t *exc, int64_t *rows, int64_t *cols)
{
    uint64_t key = 0, x = 0;
    for (uint64_t e = 0; e < nnz; e++) {
        if (e % chunk == 0) {
            key = base[e / chunk];
            x = xstart[e / chunk];
        }
        key = (da[e] || e % chunk == 0) ? key + da[e] : exc[x++];
        d32_rowcol(layout, key, &rows[e], &cols[e]);
    }
}
