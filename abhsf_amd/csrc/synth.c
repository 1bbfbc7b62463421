/*
 * synth.c -- seeded synthetic matrix generators (host, OpenMP), CSR output.
 *
 * Every random draw is a counter-based hash of (seed, stream, index), so the
 * output is identical for any thread count.  The reference's own generators
 * (include/abhsf/utils/synthetic.h:8-91, spmv/unified/spmv.cpp:1078-1156) use
 * unseeded std::random_device; these families are the seeded workloads the
 * benchmarks are quoted on.
 *
 * Two-call protocol: call with rowptr/col/val == NULL to get nnz (rowptr may be
 * non-NULL to receive row pointers), then with arrays of that size.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014) used as a counter-based hash */
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(seed ^ mix64(stream * 0x632be59bd9b4e019ull ^ mix64(i)));
}
static inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }
/* U[-1, 1) */
static inline double uval(uint64_t seed, uint64_t i) { return 2.0 * u01(rnd(seed, 7, i)) - 1.0; }
/* N(0,1) by Box-Muller on two hashed uniforms */
static inline double nval(uint64_t seed, uint64_t i)
{
    double u1 = u01(rnd(seed, 11, i)), u2 = u01(rnd(seed, 13, i));
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void scan_rowptr(int64_t *rp, int64_t m)
{
    int64_t acc = 0;
    for (int64_t r = 0; r <= m; r++) { int64_t c = rp[r]; rp[r] = acc; acc += c; }
}

/* 2D 5-point Laplacian on a g x g grid, natural ordering, 4 / -1. */
int64_t synth_laplacian2d(int64_t g, int64_t *rowptr, int32_t *col, double *val)
{
    const int64_t n = g * g;
    if (!col) return 5 * n - 4 * g;
    int64_t e = 0;
    for (int64_t p = 0; p < n; p++) {
        int64_t i = p / g, j = p % g;
        rowptr[p] = e;
        if (i > 0) { col[e] = (int32_t)(p - g); val[e++] = -1.0; }
        if (j > 0) { col[e] = (int32_t)(p - 1); val[e++] = -1.0; }
        col[e] = (int32_t)p; val[e++] = 4.0;
        if (j < g - 1) { col[e] = (int32_t)(p + 1); val[e++] = -1.0; }
        if (i < g - 1) { col[e] = (int32_t)(p + g); val[e++] = -1.0; }
    }
    rowptr[n] = e;
    return e;
}

/* 3D 27-point stencil on a g^3 grid, natural ordering, 26 / -1. */
int64_t synth_stencil27(int64_t g, int64_t *rowptr, int32_t *col, double *val)
{
    const int64_t n = g * g * g;
    const int64_t nnz = (3 * g - 2) * (3 * g - 2) * (3 * g - 2);
    if (!col) return nnz;
    /* row pointers: count neighbours in each dimension */
    for (int64_t p = 0; p < n; p++) {
        int64_t i = p / (g * g), j = (p / g) % g, k = p % g;
        int64_t ci = 3 - (i == 0) - (i == g - 1), cj = 3 - (j == 0) - (j == g - 1), ck = 3 - (k == 0) - (k == g - 1);
        rowptr[p] = ci * cj * ck;
    }
    scan_rowptr(rowptr, n);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) {
        int64_t i = p / (g * g), j = (p / g) % g, k = p % g;
        int64_t e = rowptr[p];
        for (int64_t di = -1; di <= 1; di++)
            for (int64_t dj = -1; dj <= 1; dj++)
                for (int64_t dk = -1; dk <= 1; dk++) {
                    int64_t a = i + di, b = j + dj, c = k + dk;
                    if (a < 0 || a >= g || b < 0 || b >= g || c < 0 || c >= g) continue;
                    int64_t q = (a * g + b) * g + c;
                    col[e] = (int32_t)q;
                    val[e++] = (q == p) ? 26.0 : -1.0;
                }
    }
    return nnz;
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Block-structured mix: 32-row strips; type per strip with probabilities
 * p_dense / p_band / rest:
 *   dense   3 full 32x32 blocks: the diagonal block and 2 distinct uniform
 *           block columns
 *   banded  block-tridiagonal over 8x8 sub-blocks (24 entries per row)
 *   scatter 12 distinct uniform columns per row
 * values N(0,1). */
int64_t synth_block_mix(int64_t n, double p_dense, double p_band, uint64_t seed, int64_t *rowptr, int32_t *col,
                        double *val)
{
    const int64_t S = 32, nstrip = (n + S - 1) / S, NB = (n + S - 1) / S, nsub = (n + 7) / 8;
    if (rowptr) {
        for (int64_t r = 0; r < n; r++) {
            int64_t I = r / S;
            double u = u01(rnd(seed, 1, (uint64_t)I));
            int64_t c;
            if (u < p_dense) {
                c = 0;
                /* three distinct block columns; edge block may be partial */
                int64_t bc[3];
                bc[0] = I;
                int nb = 1;
                for (uint64_t t = 0; nb < 3 && t < 1000; t++) {
                    int64_t J = (int64_t)(rnd(seed, 2, (uint64_t)I * 1000 + t) % (uint64_t)NB);
                    int dup = 0;
                    for (int q = 0; q < nb; q++) dup |= (bc[q] == J);
                    if (!dup) bc[nb++] = J;
                }
                for (int q = 0; q < nb; q++) c += (bc[q] * S + S <= n) ? S : n - bc[q] * S;
            } else if (u < p_dense + p_band) {
                int64_t R = r / 8;
                c = 0;
                for (int64_t Q = R - 1; Q <= R + 1; Q++)
                    if (Q >= 0 && Q < nsub) c += (Q * 8 + 8 <= n) ? 8 : n - Q * 8;
            } else {
                c = n < 12 ? n : 12;
            }
            rowptr[r] = c;
        }
        scan_rowptr(rowptr, n);
    }
    if (!col) return rowptr ? rowptr[n] : -1;
    (void)nstrip;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < n; r++) {
        int64_t I = r / S;
        double u = u01(rnd(seed, 1, (uint64_t)I));
        int64_t e = rowptr[r];
        const int64_t e0 = e;
        if (u < p_dense) {
            int64_t bc[3];
            bc[0] = I;
            int nb = 1;
            for (uint64_t t = 0; nb < 3 && t < 1000; t++) {
                int64_t J = (int64_t)(rnd(seed, 2, (uint64_t)I * 1000 + t) % (uint64_t)NB);
                int dup = 0;
                for (int q = 0; q < nb; q++) dup |= (bc[q] == J);
                if (!dup) bc[nb++] = J;
            }
            for (int a = 1; a < nb; a++)
                for (int b = a; b > 0 && bc[b] < bc[b - 1]; b--) { int64_t t = bc[b]; bc[b] = bc[b - 1]; bc[b - 1] = t; }
            for (int q = 0; q < nb; q++)
                for (int64_t c = bc[q] * S; c < bc[q] * S + S && c < n; c++) col[e++] = (int32_t)c;
        } else if (u < p_dense + p_band) {
            int64_t R = r / 8;
            for (int64_t Q = R - 1; Q <= R + 1; Q++)
                if (Q >= 0 && Q < nsub)
                    for (int64_t c = Q * 8; c < Q * 8 + 8 && c < n; c++) col[e++] = (int32_t)c;
        } else {
            int want = n < 12 ? (int)n : 12, got = 0;
            int32_t cs[12];
            for (uint64_t t = 0; got < want; t++) {
                int32_t c = (int32_t)(rnd(seed, 3, (uint64_t)r * 4096 + t) % (uint64_t)n);
                int dup = 0;
                for (int q = 0; q < got; q++) dup |= (cs[q] == c);
                if (!dup) cs[got++] = c;
            }
            qsort(cs, (size_t)got, sizeof(int32_t), cmp_i32);
            for (int q = 0; q < got; q++) col[e++] = cs[q];
        }
        for (int64_t q = e0; q < e; q++) val[q] = nval(seed, (uint64_t)q);
    }
    return rowptr[n];
}

/* Banded: row i holds columns [i-bw, i+bw] within [0, n); values N(0,1). */
int64_t synth_banded(int64_t n, int64_t bw, uint64_t seed, int64_t *rowptr, int32_t *col, double *val)
{
    if (rowptr) {
        for (int64_t r = 0; r < n; r++) {
            int64_t lo = r - bw < 0 ? 0 : r - bw, hi = r + bw >= n ? n - 1 : r + bw;
            rowptr[r] = hi - lo + 1;
        }
        scan_rowptr(rowptr, n);
    }
    if (!col) return rowptr ? rowptr[n] : -1;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) {
        int64_t lo = r - bw < 0 ? 0 : r - bw, hi = r + bw >= n ? n - 1 : r + bw;
        int64_t e = rowptr[r];
        for (int64_t c = lo; c <= hi; c++, e++) { col[e] = (int32_t)c; val[e] = nval(seed, (uint64_t)e); }
    }
    return rowptr[n];
}

/* Random blocks: bs x bs blocks; each block row holds nbr distinct block
 * columns (the diagonal one and nbr-1 uniform ones); each position of a chosen
 * block is present with probability density.  Values N(0,1). */
int64_t synth_random_blocks(int64_t n, int64_t bs, int64_t nbr, double density, uint64_t seed, int64_t *rowptr,
                            int32_t *col, double *val)
{
    const int64_t NB = (n + bs - 1) / bs;
    if (nbr > NB) nbr = NB;
    if (nbr > 64) nbr = 64;
    /* entry (r, c) of chosen block J is present iff hash < density */
    const uint64_t thr = (uint64_t)(density * 18446744073709551615.0);
#define BLOCK_COLS(I, bc, nb)                                                                       \
    do {                                                                                            \
        bc[0] = (I);                                                                                \
        nb = 1;                                                                                     \
        for (uint64_t t = 0; nb < nbr && t < 100000; t++) {                                         \
            int64_t J_ = (int64_t)(rnd(seed, 21, (uint64_t)(I) * 100000 + t) % (uint64_t)NB);       \
            int d_ = 0;                                                                             \
            for (int q_ = 0; q_ < nb; q_++) d_ |= (bc[q_] == J_);                                   \
            if (!d_) bc[nb++] = J_;                                                                 \
        }                                                                                           \
        for (int a_ = 1; a_ < nb; a_++)                                                             \
            for (int b_ = a_; b_ > 0 && bc[b_] < bc[b_ - 1]; b_--) {                                \
                int64_t t_ = bc[b_]; bc[b_] = bc[b_ - 1]; bc[b_ - 1] = t_;                          \
            }                                                                                       \
    } while (0)
    if (rowptr) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t r = 0; r < n; r++) {
            int64_t bc[64];
            int nb;
            BLOCK_COLS(r / bs, bc, nb);
            int64_t c = 0;
            for (int q = 0; q < nb; q++)
                for (int64_t cc = bc[q] * bs; cc < bc[q] * bs + bs && cc < n; cc++)
                    c += (cc == r) || (rnd(seed, 22, (uint64_t)r * (uint64_t)n + (uint64_t)cc) <= thr);
            rowptr[r] = c;
        }
        scan_rowptr(rowptr, n);
    }
    if (!col) return rowptr ? rowptr[n] : -1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < n; r++) {
        int64_t bc[64];
        int nb;
        BLOCK_COLS(r / bs, bc, nb);
        int64_t e = rowptr[r];
        for (int q = 0; q < nb; q++)
            for (int64_t cc = bc[q] * bs; cc < bc[q] * bs + bs && cc < n; cc++)
                if ((cc == r) || (rnd(seed, 22, (uint64_t)r * (uint64_t)n + (uint64_t)cc) <= thr)) {
                    col[e] = (int32_t)cc;
                    val[e] = nval(seed, (uint64_t)e);
                    e++;
                }
    }
#undef BLOCK_COLS
    return rowptr[n];
}

/* R-MAT (Chakrabarti, Zhan, Faloutsos, SDM 2004): ef * 2^scale edges, each
 * edge descends scale levels choosing quadrant a/b/c/d; no permutation or
 * noise; duplicates merged; values U[-1,1].  Returns nnz after merging.
 * Work arrays are allocated internally; call once with col == NULL to get the
 * merged count (rowptr receives row pointers), then with arrays. */
int64_t synth_rmat(int scale, int64_t ef, double a, double b, double c, uint64_t seed, int64_t *rowptr,
                   int32_t *col, double *val)
{
    const int64_t n = (int64_t)1 << scale, ne = ef * n;
    const uint64_t ta = (uint64_t)(a * 18446744073709551615.0);
    const uint64_t tb = (uint64_t)((a + b) * 18446744073709551615.0);
    const uint64_t tc = (uint64_t)((a + b + c) * 18446744073709551615.0);
    uint32_t *er = (uint32_t *)malloc((size_t)ne * 4), *ec = (uint32_t *)malloc((size_t)ne * 4);
    if (!er || !ec) { free(er); free(ec); return -1; }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; e++) {
        uint32_t r = 0, cc = 0;
        for (int l = 0; l < scale; l++) {
            uint64_t u = rnd(seed, 31 + (uint64_t)l, (uint64_t)e);
            uint32_t rb = (u >= tb), cb = (u >= ta && u < tb) || (u >= tc);
            r = (r << 1) | rb;
            cc = (cc << 1) | cb;
        }
        er[e] = r;
        ec[e] = cc;
    }
    /* bucket edges by row (counting sort), then sort + merge each row */
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < ne; e++) cnt[er[e] + 1]++;
    for (int64_t r = 0; r < n; r++) cnt[r + 1] += cnt[r];
    int32_t *bycol = (int32_t *)malloc((size_t)ne * 4);
    int64_t *pos = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    memcpy(pos, cnt, (size_t)n * sizeof(int64_t));
    for (int64_t e = 0; e < ne; e++) bycol[pos[er[e]]++] = (int32_t)ec[e];
    free(pos);
    free(er);
    free(ec);
    int64_t *uniq = (int64_t *)malloc((size_t)n * sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; r++) {
        int32_t *p = bycol + cnt[r];
        int64_t k = cnt[r + 1] - cnt[r];
        qsort(p, (size_t)k, sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t i = 0; i < k; i++)
            if (i == 0 || p[i] != p[i - 1]) p[u++] = p[i];
        uniq[r] = u;
    }
    int64_t tot = 0;
    for (int64_t r = 0; r < n; r++) {
        if (rowptr) rowptr[r] = tot;
        tot += uniq[r];
    }
    if (rowptr) rowptr[n] = tot;
    if (col) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < n; r++) {
            int64_t o = rowptr[r];
            for (int64_t i = 0; i < uniq[r]; i++) {
                col[o + i] = bycol[cnt[r] + i];
                val[o + i] = uval(seed, (uint64_t)(o + i));
            }
        }
    }
    free(uniq);
    free(bycol);
    free(cnt);
    return tot;
}

int synth_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Uniform block matrix of the reference's micro-benchmarks
 * (spmv/orig/88.cpp:43-141, 150-202): nb x nb blocks of size bs, every block a
 * copy of one random pattern of bnnz distinct positions drawn from the seeded
 * stream; all values 1.  With x = 1 every row sums to the pattern's row
 * count, so sum(y) = bnnz * nb^2 per SpMV (88.cpp:828-832). */
int64_t synth_uniform_blocks(int64_t nb, int64_t bs, int64_t bnnz, uint64_t seed, int64_t *rowptr, int32_t *col,
                             double *val)
{
    if (bnnz < 1 || bnnz > bs * bs || bs > 64) return -1;
    uint8_t pat[64 * 64];
    memset(pat, 0, sizeof(pat));
    for (int64_t got = 0, t = 0; got < bnnz; t++) {
        uint64_t p = rnd(seed, 41, (uint64_t)t) % (uint64_t)(bs * bs);
        if (!pat[p]) { pat[p] = 1; got++; }
    }
    int64_t rowcnt[64];
    for (int64_t r = 0; r < bs; r++) {
        rowcnt[r] = 0;
        for (int64_t c = 0; c < bs; c++) rowcnt[r] += pat[r * bs + c];
    }
    const int64_t n = nb * bs;
    if (rowptr) {
        for (int64_t r = 0; r < n; r++) rowptr[r] = rowcnt[r % bs] * nb;
        scan_rowptr(rowptr, n);
    }
    if (!col) return rowptr ? rowptr[n] : bnnz * nb * nb;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) {
        int64_t e = rowptr[r];
        const int64_t lr = r % bs;
        for (int64_t J = 0; J < nb; J++)
            for (int64_t c = 0; c < bs; c++)
                if (pat[lr * bs + c]) { col[e] = (int32_t)(J * bs + c); val[e++] = 1.0; }
    }
    return rowptr[n];
}
