// spmv.hip -- ABHSF SpMV / SpMM kernels for gfx950 (wave64).
//
// One wavefront owns one strip (block row of s rows) and accumulates its
// s x R output tile in LDS; the four scheme groups of the strip are walked in
// turn (COO, CSR, bitmap, dense), each with a decode specialised to its layout:
//   COO / CSR  element-parallel: E = 64/R entries per wave step, R right-hand
//              sides per entry; the entry's block is found by a binary search
//              over a 64-block chunk's inclusive nnz prefix (wave scan in LDS);
//              CSR rows come from a binary search over the block's row pointers.
//   bitmap     position-parallel: lane i of a 64-bit word tests bit i, the value
//              offset is popcount(word & lanemask_lt) (the word is wave-uniform);
//   dense      the bitmap walk with every bit set.
// Products of the same output row are summed with xor-shuffles inside the wave
// where rows are known to be contiguous (bitmap, dense) and with LDS atomic
// adds where they are not (COO, CSR).
//
// Reference semantics: y[lrow] += v * x[bcol*s + lcol] per scheme
// (spmv/orig/88.cpp:214-235 COO, 88_bitmap.cpp:183-207 bitmap, 88.cpp:587-604
// dense); the whole-matrix call is y = alpha*A*x + beta*y, which with
// alpha = beta = 1 is the accumulating csr_matrix::spmv of
// spmv/unified/spmv.cpp:140-159.
//
// COMP (FP32 value storage): every product is split exactly with an FMA
// (TwoProduct) and every LDS accumulation is made error-free with TwoSum: the
// atomic returns the old value a, so the rounding error of fl(a + p) is
// recovered exactly and added to a second accumulator (Ogita, Rump, Oishi,
// "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005, Dot2).
#include "spmv.hpp"

namespace abhsf {
namespace {

template <typename T>
__device__ inline T ld(const void *p, uint64_t i) { return static_cast<const T *>(p)[i]; }

struct Acc {
    double *hi;  // [s * R]
    double *lo;  // [s * R] (COMP only)
};

// Atomic accumulate of p into acc slot i.
template <bool COMP>
__device__ inline void acc_atomic(const Acc &a, uint32_t i, double p, double perr)
{
    if constexpr (COMP) {
        double old = atomicAdd(&a.hi[i], p);
        double s, e;
        two_sum(old, p, s, e);
        atomicAdd(&a.lo[i], e + perr);
    } else {
        atomicAdd(&a.hi[i], p);
    }
}

// Product v*x, with its exact rounding error when COMP.
template <bool COMP>
__device__ inline double prod(double v, double x, double &err)
{
    double p = v * x;
    if constexpr (COMP) err = __builtin_fma(v, x, -p); else err = 0.0;
    return p;
}

template <typename VT, typename BT, typename IT, typename RT, int R, bool COMP>
__global__ __launch_bounds__(256) void k_spmm_strip(DevMat A, const double *__restrict__ X, int64_t ldx,
                                                    double *__restrict__ Y, int64_t ldy, int nrhs,
                                                    double alpha, double beta, int wpb, int64_t nstrip_total)
{
    constexpr int E = kWave / R;  // entries per wave step
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = lane / R, j = lane % R;
    const bool jok = j < nrhs;
    const int s = A.s, k = A.k;
    const int64_t I = (int64_t)blockIdx.x * wpb + wave;
    if (I >= nstrip_total) return;  // whole wave leaves; no workgroup barriers below

    const size_t per_wave = (size_t)s * R * sizeof(double) * (COMP ? 2 : 1) + 2 * kWave * sizeof(uint32_t);
    unsigned char *base = smem + per_wave * wave;
    Acc acc;
    acc.hi = reinterpret_cast<double *>(base);
    acc.lo = COMP ? acc.hi + (size_t)s * R : nullptr;
    uint32_t *s_end = reinterpret_cast<uint32_t *>(base + (size_t)s * R * sizeof(double) * (COMP ? 2 : 1));
    uint32_t *s_bc = s_end + kWave;

    for (int i = lane; i < s * R; i += kWave) {
        acc.hi[i] = 0.0;
        if constexpr (COMP) acc.lo[i] = 0.0;
    }
    wave_sync();

    const uint32_t smask = (uint32_t)s - 1;

    // ---- COO and CSR groups: element-parallel over 64-block chunks ----------
#pragma unroll 1
    for (int sc = COO; sc <= CSR; sc++) {
        const uint32_t b0 = A.sptr[sc][I], b1 = A.sptr[sc][I + 1];
        uint32_t e = A.eptr[sc][I];
        const VT *val = static_cast<const VT *>(A.val[sc]);
#pragma unroll 1
        for (uint32_t cb = b0; cb < b1; cb += kWave) {
            const uint32_t jb = cb + lane;
            const bool ok = jb < b1;
            const uint32_t z = ok ? (uint32_t)A.z1[sc][jb] + 1u : 0u;
            const uint32_t incl = wave_incl_scan_u32(z);
            const uint32_t T = __shfl(incl, kWave - 1, kWave);
            s_end[lane] = incl;
            s_bc[lane] = ok ? (uint32_t)ld<BT>(A.bcol[sc], jb) : 0u;
            wave_sync();
            const int nbk = (int)min<uint32_t>(kWave, b1 - cb);
#pragma unroll 1
            for (uint32_t q = 0; q < T; q += E) {
                const uint32_t t = q + slot;
                if (t < T) {
                    int lo = 0, hi = nbk - 1;  // first chunk block with end > t
                    while (lo < hi) {
                        int mid = (lo + hi) >> 1;
                        if (s_end[mid] > t) hi = mid; else lo = mid + 1;
                    }
                    uint32_t lr, lc;
                    if (sc == COO) {
                        const uint32_t id = ld<IT>(A.coo_idx, e + t);
                        lr = id >> k;
                        lc = id & smask;
                    } else {
                        const uint32_t qb = t - (lo ? s_end[lo - 1] : 0u);  // entry within block
                        const RT *rp = static_cast<const RT *>(A.csr_rowptr) + (uint64_t)(cb + lo) * (s + 1);
                        int rl = 0, rh = s - 1;  // last row with rp[row] <= qb
                        while (rl < rh) {
                            int mid = (rl + rh + 1) >> 1;
                            if ((uint32_t)rp[mid] <= qb) rl = mid; else rh = mid - 1;
                        }
                        lr = (uint32_t)rl;
                        lc = A.csr_col[e + t];
                    }
                    if (jok) {
                        const double v = (double)val[e + t];
                        const double xv = X[((int64_t)s_bc[lo] * s + lc) * ldx + j];
                        double pe;
                        const double p = prod<COMP>(v, xv, pe);
                        acc_atomic<COMP>(acc, lr * R + j, p, pe);
                    }
                }
            }
            e += T;
            wave_sync();
        }
    }

    // ---- bitmap and dense groups: position-parallel over 64-bit words --------
    // Within a word, positions t*E + slot (t < R) share a row in aligned groups
    // of G = min(E, s) slots, so products are summed with xor-shuffles over the
    // slot bits and the group leader updates the row (no two leaders of one
    // step hit the same row and rhs, so no atomics are needed).
    const int G = E < s ? E : s;
    const int W = A.W;
    const uint32_t ssq = (uint32_t)s * (uint32_t)s;
#pragma unroll 1
    for (int sc = BITMAP; sc <= DENSE; sc++) {
        const uint32_t b0 = A.sptr[sc][I], b1 = A.sptr[sc][I + 1];
        uint64_t vbase = (sc == BITMAP) ? (uint64_t)A.eptr[BITMAP][I] : 0;
        const VT *val = static_cast<const VT *>(A.val[sc]);
#pragma unroll 1
        for (uint32_t b = b0; b < b1; b++) {
            const int64_t c0 = (int64_t)ld<BT>(A.bcol[sc], b) * s;
            const uint64_t *words = A.bm_words + (uint64_t)b * W;
            if (sc == DENSE) vbase = (uint64_t)b * ssq;
#pragma unroll 1
            for (int w = 0; w < W; w++) {
                const uint64_t word = (sc == BITMAP) ? words[w] : ~0ull;
#pragma unroll
                for (int t = 0; t < R; t++) {
                    const uint32_t i = (uint32_t)(t * E + slot);  // bit within the word
                    const uint32_t pos = (uint32_t)w * 64u + i;    // position in the block
                    const bool inb = pos < ssq;
                    const bool bit = inb && ((word >> i) & 1ull);
                    const uint32_t lr = pos >> k, lc = pos & smask;
                    double p = 0.0, pe = 0.0;
                    if (bit && jok) {
                        const uint64_t vi = (sc == BITMAP) ? vbase + __popcll(word & ((1ull << i) - 1ull))
                                                           : vbase + pos;
                        const int64_t gc = c0 + lc;
                        const double xv = (gc < A.n) ? X[gc * ldx + j] : 0.0;
                        p = prod<COMP>((double)val[vi], xv, pe);
                    }
                    for (int o = (G >> 1) * R; o >= R; o >>= 1) {
                        const double q = __shfl_xor(p, o, kWave);
                        if constexpr (COMP) {
                            const double qe = __shfl_xor(pe, o, kWave);
                            double sum, err;
                            two_sum(p, q, sum, err);
                            p = sum;
                            pe += qe + err;
                        } else {
                            p += q;
                        }
                    }
                    if ((slot & (G - 1)) == 0 && inb && jok) {
                        const uint32_t ai = lr * R + j;
                        if constexpr (COMP) {
                            double sum, err;
                            two_sum(acc.hi[ai], p, sum, err);
                            acc.hi[ai] = sum;
                            acc.lo[ai] += err + pe;
                        } else {
                            acc.hi[ai] += p;
                        }
                    }
                }
                wave_sync();
                if (sc == BITMAP) vbase += __popcll(word);  // dense positions are absolute
            }
        }
    }

    wave_sync();
    // ---- write the strip's rows ---------------------------------------------
    for (int i = lane; i < s * R; i += kWave) {
        const int r = i / R, jj = i % R;
        const int64_t gr = I * s + r;
        if (gr >= A.m || jj >= nrhs) continue;
        double a = acc.hi[i];
        if constexpr (COMP) a += acc.lo[i];
        double *yp = Y + gr * ldy + jj;
        *yp = (beta == 0.0) ? alpha * a : alpha * a + beta * *yp;
    }
}

template <typename VT, typename BT, typename IT, typename RT, int R, bool COMP>
void launch_t(const DevMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy, int nrhs, double alpha,
              double beta, hipStream_t st)
{
    const size_t per_wave = (size_t)A.s * R * sizeof(double) * (COMP ? 2 : 1) + 2 * kWave * sizeof(uint32_t);
    int wpb = (int)(65536 / per_wave);
    if (wpb > 4) wpb = 4;
    ABHSF_REQUIRE(wpb >= 1, ABHSF_ERR_UNSUPPORTED, "output tile does not fit in LDS");
    const int64_t nblk = (A.M + wpb - 1) / wpb;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "too many strips");
    if (A.M == 0) return;
    hipLaunchKernelGGL((k_spmm_strip<VT, BT, IT, RT, R, COMP>), dim3((unsigned)nblk), dim3(64 * wpb),
                       per_wave * wpb, st, A, X, ldx, Y, ldy, nrhs, alpha, beta, wpb, A.M);
    ABHSF_HIP(hipGetLastError());
}

template <typename VT, typename BT, typename IT, typename RT, bool COMP>
void launch_r(int R, const DevMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy, int nrhs,
              double alpha, double beta, hipStream_t st)
{
    switch (R) {
    case 1: launch_t<VT, BT, IT, RT, 1, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st); break;
    case 2: launch_t<VT, BT, IT, RT, 2, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st); break;
    case 4: launch_t<VT, BT, IT, RT, 4, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st); break;
    case 8: launch_t<VT, BT, IT, RT, 8, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st); break;
    case 16: launch_t<VT, BT, IT, RT, 16, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st); break;
    case 32: launch_t<VT, BT, IT, RT, 32, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st); break;
    default: throw Error(ABHSF_ERR_UNSUPPORTED, "unsupported rhs tile");
    }
}

template <typename VT, bool COMP>
void launch_v(const SpmvLaunch &L, const DevMat &A, int R, const double *X, int64_t ldx, double *Y, int64_t ldy,
              int nrhs, double alpha, double beta, hipStream_t st)
{
    const bool b16 = L.bcol_bytes == 2;
    const int s = A.s;
    if (s <= 8) {
        if (b16) launch_r<VT, uint16_t, uint8_t, uint8_t, COMP>(R, A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        else launch_r<VT, uint32_t, uint8_t, uint8_t, COMP>(R, A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
    } else if (s <= 16) {
        if (b16) launch_r<VT, uint16_t, uint8_t, uint16_t, COMP>(R, A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        else launch_r<VT, uint32_t, uint8_t, uint16_t, COMP>(R, A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
    } else {
        if (b16) launch_r<VT, uint16_t, uint16_t, uint16_t, COMP>(R, A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        else launch_r<VT, uint32_t, uint16_t, uint16_t, COMP>(R, A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
    }
}

}  // namespace

int max_rhs_tile(int s, bool comp)
{
    int R = 32;
    while (R > 1 && (size_t)s * R * sizeof(double) * (comp ? 2 : 1) + 2 * kWave * sizeof(uint32_t) > 65536) R >>= 1;
    return R;
}

void spmm(const SpmvLaunch &L, const DevMat &A, int nrhs, double alpha, const double *X, int64_t ldx, double beta,
          double *Y, int64_t ldy, hipStream_t st)
{
    const bool comp = L.val_bytes == 4;
    const int rmax = max_rhs_tile(A.s, comp);
    for (int j0 = 0; j0 < nrhs; j0 += rmax) {
        const int nr = nrhs - j0 < rmax ? nrhs - j0 : rmax;
        int R = 1;
        while (R < nr) R <<= 1;
        if (comp) launch_v<float, true>(L, A, R, X + j0, ldx, Y + j0, ldy, nr, alpha, beta, st);
        else launch_v<double, false>(L, A, R, X + j0, ldx, Y + j0, ldy, nr, alpha, beta, st);
    }
}

}  // namespace abhsf
