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

#include <algorithm>
#include <type_traits>

namespace abhsf {
namespace {

template <typename T>
__device__ inline T ld(const void *p, uint64_t i) { return static_cast<const T *>(p)[i]; }

// x may be any double pointer (x + 1 of a caller's array included); the
// paired loads of the dense paths take it as double2 only when it is 16-byte
// aligned.
__device__ inline bool x_aligned16(const double *x) { return (reinterpret_cast<uintptr_t>(x) & 15u) == 0; }

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

// ===========================================================================
// Single right-hand side (SpMV).
// ===========================================================================

// Accumulate p (with product error pe) into a lane-private compensated pair.
template <bool COMP>
__device__ inline void acc_add(double &hi, double &lo, double v, double xv)
{
    if constexpr (COMP) {
        double pe;
        const double p = prod<true>(v, xv, pe);
        double s, e;
        two_sum(hi, p, s, e);
        hi = s;
        lo += e + pe;
    } else {
        hi = __builtin_fma(v, xv, hi);
    }
}

// Sum (hi, lo) over aligned groups of P lanes (P in {1, 2, 4, 8}).
template <bool COMP, int P>
__device__ inline void group_reduce_p(double &hi, double &lo)
{
    if constexpr (!COMP) {
        hi = group_sum_dpp<P>(hi);
    } else {
        auto step = [&](auto ctrl_tag) {
            constexpr int CTRL = decltype(ctrl_tag)::value;
            const double th = dpp_f64<CTRL>(hi), tl = dpp_f64<CTRL>(lo);
            double s, e;
            two_sum(hi, th, s, e);
            hi = s;
            lo += tl + e;
        };
        if constexpr (P >= 2) step(std::integral_constant<int, DPP_QUAD_XOR1>{});
        if constexpr (P >= 4) step(std::integral_constant<int, DPP_QUAD_XOR2>{});
        if constexpr (P >= 8) step(std::integral_constant<int, DPP_ROW_HALF_MIRROR>{});
    }
}

template <bool COMP>
__device__ inline void group_reduce(int P, double &hi, double &lo)
{
    switch (P) {
    case 1: break;
    case 2: group_reduce_p<COMP, 2>(hi, lo); break;
    case 4: group_reduce_p<COMP, 4>(hi, lo); break;
    default: group_reduce_p<COMP, 8>(hi, lo); break;
    }
}

// CSR group, blocks [b0, b1) whose first value is at offset e.  Lanes own
// block rows: for s <= 64 a wave step covers 64/s blocks (lane l -> block l/s,
// row l%s), for s > 64 one block in s/64 row passes.  A lane reads its row's
// bounds from the block's row pointers, sums the row's products in registers
// and adds the sum to the strip accumulator: no search, one LDS update per
// row.  s_st holds the chunk's exclusive value offsets, s_bc its block columns.
template <typename VT, typename BT, typename RT, bool COMP>
__device__ inline void csr_rows_r1(const DevMat &A, uint32_t b0, uint32_t b1, uint32_t e,
                                   const double *__restrict__ x, const Acc &acc, uint32_t *s_st, uint32_t *s_bc,
                                   int lane)
{
    if (b0 >= b1) return;
    const int s = A.s;
    const int sw = s < kWave ? s : kWave;          // lanes per block
    const int G = kWave / sw;                      // blocks per wave step
    const int passes = s / sw;                     // row passes per block
    const VT *__restrict__ val = static_cast<const VT *>(A.val[CSR]);
    const uint8_t *__restrict__ ccol = A.csr_col;
    const RT *__restrict__ rowptr = static_cast<const RT *>(A.csr_rowptr);
#pragma unroll 1
    for (uint32_t cb = b0; cb < b1; cb += kWave) {
        const uint32_t jb = cb + lane;
        const bool ok = jb < b1;
        const uint32_t z = ok ? (uint32_t)A.z1[CSR][jb] + 1u : 0u;
        const uint32_t incl = wave_incl_scan_dpp(z);
        s_st[lane] = e + incl - z;
        s_bc[lane] = ok ? (uint32_t)ld<BT>(A.bcol[CSR], jb) : 0u;
        e += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        wave_sync();
        const int nbk = (int)min<uint32_t>(kWave, b1 - cb);
        const int lb = lane / sw, lr = lane % sw;
#pragma unroll 1
        for (int j0 = 0; j0 < nbk; j0 += G) {
            const int j = j0 + lb;
            if (j < nbk) {
                const RT *rp = rowptr + (uint64_t)(cb + (uint32_t)j) * (uint64_t)(s + 1);
                const uint32_t st = s_st[j];
                const int64_t xc = (int64_t)s_bc[j] * s;
#pragma unroll 1
                for (int ps = 0; ps < passes; ps++) {
                    const int r = ps * sw + lr;
                    const uint32_t q0 = rp[r], q1 = rp[r + 1];
                    double hi = 0.0, lo = 0.0;
                    uint32_t q = q0;
                    for (; q + 1 < q1; q += 2) {
                        const double v0 = (double)val[st + q], v1 = (double)val[st + q + 1];
                        const double x0 = x[xc + ccol[st + q]], x1 = x[xc + ccol[st + q + 1]];
                        acc_add<COMP>(hi, lo, v0, x0);
                        acc_add<COMP>(hi, lo, v1, x1);
                    }
                    if (q < q1) acc_add<COMP>(hi, lo, (double)val[st + q], x[xc + ccol[st + q]]);
                    if (q1 > q0) acc_atomic<COMP>(acc, (uint32_t)r, hi, lo);
                }
            }
        }
        wave_sync();
    }
}

// COO group, blocks [b0, b1) with first value at offset e.  Chunks of 64
// blocks, lane j <-> block j.  Phase A: every lane handles the first ZS
// entries of its own block (independent loads, no lookup); phase B, only when
// some block of the chunk holds more than ZS entries: the remaining entries
// are spread one per lane and located by a binary search over the chunk's
// prefix sums in LDS.  A block's z1 holds z - 1 in its low A.zb bits and, in
// a tile, the block's strip t within the tile above them: t << 2k is added to
// its in-block indices, so (index >> k) = (t << k) + lrow is the tile row and
// the column bits are unchanged.
template <typename VT, typename BT, typename IT, bool COMP, int ZS, int CH, bool TILE = false>
__device__ inline void coo_r1(const DevMat &A, uint32_t b0, uint32_t b1, uint32_t e, const double *__restrict__ x,
                              const Acc &acc, uint32_t *s_end, uint32_t *s_bc, uint32_t *s_st, uint32_t *s_to,
                              int lane)
{
    if (b0 >= b1) return;
    const int s = A.s, k = A.k, zb = TILE ? A.zb : 16;
    const uint32_t smask = (uint32_t)s - 1, zmask = (1u << zb) - 1u;
    const VT *__restrict__ val = static_cast<const VT *>(A.val[COO]);
    const IT *__restrict__ idx = static_cast<const IT *>(A.coo_idx);
    const uint16_t *__restrict__ z1 = A.z1[COO];
    const BT *__restrict__ bcol = static_cast<const BT *>(A.bcol[COO]);
    // CH sub-chunks of 64 blocks per step: all their loads are issued before
    // any of them is consumed, so several memory round trips overlap.
#pragma unroll 1
    for (uint32_t cb = b0; cb < b1; cb += (uint32_t)(kWave * CH)) {
        // zt = z | (strip within the tile << 17)  (z <= 2^16)
        uint32_t zt[CH], bc[CH], st[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const uint32_t jb = cb + (uint32_t)(c * kWave) + (uint32_t)lane;
            const bool ok = jb < b1;
            const uint32_t r = ok ? (uint32_t)z1[jb] : 0u;
            if constexpr (TILE) zt[c] = ok ? ((r & zmask) + 1u) | ((r >> zb) << 17) : 0u;
            else zt[c] = ok ? r + 1u : 0u;
            bc[c] = ok ? (uint32_t)bcol[jb] : 0u;
        }
        auto zof = [&](int c) { return TILE ? zt[c] & 0x1FFFFu : zt[c]; };
        auto tzof = [&](int c) { return TILE ? (zt[c] >> 17) << (2 * k) : 0u; };
        const uint32_t e_step = e;
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const uint32_t incl = wave_incl_scan_dpp(zof(c));
            st[c] = e + incl - zof(c);
            e += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        }
        // blocks of 16 entries or more on average (a uniform decision): phase
        // A would keep a few lanes busy for a whole memory round trip, so
        // phase B takes every entry
        const uint32_t nbs = min<uint32_t>(b1 - cb, (uint32_t)(kWave * CH));
        const uint32_t za = (e - e_step >= 16u * nbs) ? 0u : (uint32_t)ZS;
        if (za) {
            uint32_t id[CH][ZS];
            double v[CH][ZS], xv[CH][ZS];
#pragma unroll
            for (int c = 0; c < CH; c++)
#pragma unroll
                for (int i = 0; i < ZS; i++) {
                    const bool h = (uint32_t)i < zof(c);
                    id[c][i] = h ? (uint32_t)idx[st[c] + i] + tzof(c) : 0u;
                    v[c][i] = h ? (double)val[st[c] + i] : 0.0;
                }
#pragma unroll
            for (int c = 0; c < CH; c++)
#pragma unroll
                for (int i = 0; i < ZS; i++)
                    xv[c][i] = ((uint32_t)i < zof(c)) ? x[(int64_t)bc[c] * s + (id[c][i] & smask)] : 0.0;
#pragma unroll
            for (int c = 0; c < CH; c++)
#pragma unroll
                for (int i = 0; i < ZS; i++)
                    if ((uint32_t)i < zof(c)) {
                        double pe;
                        const double p = prod<COMP>(v[c][i], xv[c][i], pe);
                        acc_atomic<COMP>(acc, id[c][i] >> k, p, pe);
                    }
        }
        // phase B: entries beyond the first za of a block, one per lane
#pragma unroll 1
        for (int c = 0; c < CH; c++) {
            const uint32_t rem = zof(c) > za ? zof(c) - za : 0u;
            if (__ballot(rem != 0) == 0) continue;
            const uint32_t rincl = wave_incl_scan_dpp(rem);
            const uint32_t RT = (uint32_t)__builtin_amdgcn_readlane((int)rincl, kWave - 1);
            s_end[lane] = rincl;
            s_bc[lane] = bc[c];
            s_st[lane] = st[c] + za;
            if constexpr (TILE) s_to[lane] = tzof(c);
            wave_sync();
            const uint32_t cs = cb + (uint32_t)(c * kWave);
            const int nbk = (int)min<uint32_t>(kWave, b1 - cs);
            // two windows of 64 entries per step: both searches, then all
            // loads, then the updates
#pragma unroll 1
            for (uint32_t q = 0; q < RT; q += 2 * kWave) {
                uint32_t pos[2], xb[2], ide[2], rb[2];
                double ve[2], xe[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const uint32_t t = q + (uint32_t)(u * kWave + lane);
                    int lo = 0, hi = nbk - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_end[mid] > t) hi = mid; else lo = mid + 1;
                    }
                    pos[u] = s_st[lo] + t - (lo ? s_end[lo - 1] : 0u);
                    xb[u] = s_bc[lo];
                    rb[u] = TILE ? s_to[lo] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const bool h = q + (uint32_t)(u * kWave + lane) < RT;
                    ide[u] = h ? (uint32_t)idx[pos[u]] + rb[u] : 0u;
                    ve[u] = h ? (double)val[pos[u]] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const bool h = q + (uint32_t)(u * kWave + lane) < RT;
                    xe[u] = h ? x[(int64_t)xb[u] * s + (ide[u] & smask)] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 2; u++)
                    if (q + (uint32_t)(u * kWave + lane) < RT) {
                        double pe;
                        const double p = prod<COMP>(ve[u], xe[u], pe);
                        acc_atomic<COMP>(acc, ide[u] >> k, p, pe);
                    }
            }
            wave_sync();
        }
    }
}

// Bitmap or dense group of strip I.  Each lane owns cw consecutive row-major
// positions of the block (P = s / cw lanes per row); it accumulates its
// products in registers, the P lanes of a row are summed with DPP, and the
// first lane of the row adds the sum to the strip accumulator.
template <typename VT, typename BT, bool COMP>
__device__ inline void bmdn_r1(const DevMat &A, int sc, uint32_t b0, uint32_t b1, uint64_t vb,
                               const double *__restrict__ x, const Acc &acc, int lane)
{
    if (b0 >= b1) return;
    const int s = A.s, k = A.k;
    const uint32_t ss = (uint32_t)s * (uint32_t)s;
    const uint32_t cw = s >= 64 ? (uint32_t)s : (ss >= 64 ? ss / 64 : 1u);
    const int npass = (int)((ss + 64u * cw - 1) / (64u * cw));
    const int P = (int)((uint32_t)s / cw);
    const int64_t n = A.n;
    const VT *__restrict__ val = static_cast<const VT *>(A.val[sc]);
#pragma unroll 1
    for (uint32_t b = b0; b < b1; b++) {
        const int64_t c0 = (int64_t)ld<BT>(A.bcol[sc], b) * s;
        const uint64_t *__restrict__ words = A.bm_words + (uint64_t)b * A.W;
        const uint64_t db = (uint64_t)b * ss;
#pragma unroll 1
        for (int q = 0; q < npass; q++) {
            const uint32_t pbase = ((uint32_t)q * 64u + (uint32_t)lane) * cw;
            const bool valid = pbase < ss;
            const uint32_t row = pbase >> k, col0 = pbase & ((uint32_t)s - 1);
            const int64_t gc0 = c0 + col0;
            double hi = 0.0, lo = 0.0;
            if (sc == DENSE) {
                if (valid) {
                    const VT *vp = val + db + pbase;
                    // double2 loads of x need a 16-byte aligned x (a caller may pass
                    // x + 1 of an array): otherwise the scalar walk below
                    if (gc0 + (int64_t)cw <= n && (cw & 1) == 0 && x_aligned16(x)) {
#pragma unroll 4
                        for (uint32_t c = 0; c < cw; c += 2) {
                            const double2 xv = *reinterpret_cast<const double2 *>(x + gc0 + c);
                            double v0, v1;
                            if constexpr (sizeof(VT) == 8) {
                                const double2 vv = *reinterpret_cast<const double2 *>(vp + c);
                                v0 = vv.x;
                                v1 = vv.y;
                            } else {
                                const float2 vv = *reinterpret_cast<const float2 *>(vp + c);
                                v0 = vv.x;
                                v1 = vv.y;
                            }
                            acc_add<COMP>(hi, lo, v0, xv.x);
                            acc_add<COMP>(hi, lo, v1, xv.y);
                        }
                    } else {
                        for (uint32_t c = 0; c < cw; c++)
                            if (gc0 + c < n) acc_add<COMP>(hi, lo, (double)vp[c], x[gc0 + c]);
                    }
                }
            } else {
                // bitmap: popcount of my positions -> value offset by a wave scan
                const uint32_t nw = cw > 64 ? cw / 64 : 1;
                const uint32_t w0 = pbase >> 6, sh = pbase & 63;
                const uint64_t m = cw >= 64 ? ~0ull : ((1ull << cw) - 1ull);
                uint32_t cnt = 0;
                if (valid)
                    for (uint32_t w = 0; w < nw; w++) cnt += __popcll((words[w0 + w] >> sh) & m);
                const uint32_t incl = wave_incl_scan_dpp(cnt);
                const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
                uint64_t off = vb + incl - cnt;
                if (valid) {
                    for (uint32_t w = 0; w < nw; w++) {
                        uint64_t bits = (words[w0 + w] >> sh) & m;
                        const int64_t xb = gc0 + (int64_t)w * 64;
                        // four set bits per step so their loads are in flight together
                        while (bits) {
                            const uint32_t c0 = (uint32_t)__builtin_ctzll(bits);
                            bits &= bits - 1;
                            const bool h1 = bits != 0;
                            const uint32_t c1 = h1 ? (uint32_t)__builtin_ctzll(bits) : c0;
                            bits &= bits - 1;
                            const bool h2 = bits != 0;
                            const uint32_t c2 = h2 ? (uint32_t)__builtin_ctzll(bits) : c0;
                            bits &= bits - 1;
                            const bool h3 = bits != 0;
                            const uint32_t c3 = h3 ? (uint32_t)__builtin_ctzll(bits) : c0;
                            bits &= bits - 1;
                            const double v0 = (double)val[off];
                            const double v1 = (double)val[h1 ? off + 1 : off];
                            const double v2 = (double)val[h2 ? off + 2 : off];
                            const double v3 = (double)val[h3 ? off + 3 : off];
                            const double x0 = x[xb + c0], x1 = x[xb + c1], x2 = x[xb + c2], x3 = x[xb + c3];
                            acc_add<COMP>(hi, lo, v0, x0);
                            if (h1) acc_add<COMP>(hi, lo, v1, x1);
                            if (h2) acc_add<COMP>(hi, lo, v2, x2);
                            if (h3) acc_add<COMP>(hi, lo, v3, x3);
                            off += 1u + h1 + h2 + h3;
                        }
                    }
                }
                vb += tot;
            }
            group_reduce<COMP>(P, hi, lo);
            if (valid && (lane & (P - 1)) == 0) {
                if constexpr (COMP) {
                    double s2, e2;
                    two_sum(acc.hi[row], hi, s2, e2);
                    acc.hi[row] = s2;
                    acc.lo[row] += lo + e2;
                } else {
                    acc.hi[row] += hi;
                }
            }
        }
    }
}

// Dense group of an S-row strip (S = 16 or 32), FP64 values, coalesced: wave
// step q reads positions q*128 + 2l, +1 of the block (1 KB contiguous); lane
// l always has columns 2(l % L), +1 (L = S/2 lanes per row) and rows
// q*(128/S) + l/L, so its x pair is read once per block and it keeps one
// accumulator per step over all the group's blocks; the L lanes of a row are
// summed with DPP at the end.  (Lanes owning S*S/64 consecutive positions
// made a load touch up to 64 lines: S = 32 streamed 3.8 TB/s, now 5.6.)
template <typename BT, int S>
__device__ inline void dense_coal_r1(const DevMat &A, uint32_t b0, uint32_t b1, const double *__restrict__ x,
                                     const Acc &acc, int lane)
{
    static_assert(S == 16 || S == 32, "S x S blocks of whole 128-position steps");
    constexpr int L = S / 2, NQ = S * S / 128, RQ = 128 / S;
    if (b0 >= b1) return;
    const double *__restrict__ val = static_cast<const double *>(A.val[DENSE]);
    const int64_t n = A.n;
    const int cl = 2 * (lane % L);
    const bool xal = x_aligned16(x);  // wave-uniform: x pairs as double2 loads or two scalar loads
    double a[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) a[q] = 0.0;
#pragma unroll 1
    for (uint32_t b = b0; b < b1; b++) {
        const int64_t c0 = (int64_t)ld<BT>(A.bcol[DENSE], b) * S + cl;
        double x0 = 0.0, x1 = 0.0;  // columns past n hold stored zeros; x is not read there
        if (c0 + 1 < n && !xal) {
            x0 = x[c0];
            x1 = x[c0 + 1];
        } else if (c0 + 1 < n) {
            const double2 xv = *reinterpret_cast<const double2 *>(x + c0);
            x0 = xv.x;
            x1 = xv.y;
        } else if (c0 < n) {
            x0 = x[c0];
        }
        const double2 *__restrict__ vp = reinterpret_cast<const double2 *>(val + (uint64_t)b * (S * S)) + lane;
#pragma unroll
        for (int h = 0; h < NQ; h += 4) {  // at most four loads in flight (register budget)
            constexpr int H = NQ < 4 ? NQ : 4;
            double2 v[H];
#pragma unroll
            for (int q = 0; q < H; q++) v[q] = vp[(h + q) * 64];
#pragma unroll
            for (int q = 0; q < H; q++) a[h + q] = __builtin_fma(v[q].y, x1, __builtin_fma(v[q].x, x0, a[h + q]));
            __builtin_amdgcn_sched_barrier(0);  // keep the next loads behind these
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const double r = group_sum_dpp<L>(a[q]);
        if (lane % L == 0) acc.hi[q * RQ + lane / L] += r;
    }
}

// Dense group of strip I: the coalesced path for FP64 16- and 32-row blocks,
// the bitmap/dense walk otherwise.  The index types tell the block-size class
// at compile time (u8 indices with u16 row pointers: s = 16; u16 indices:
// s >= 32), so each kernel carries only the path it can take.
template <typename VT, typename BT, typename IT, typename RT, bool COMP>
__device__ inline void dense_r1(const DevMat &A, uint32_t b0, uint32_t b1, const double *__restrict__ x,
                                const Acc &acc, int lane)
{
    if constexpr (sizeof(VT) == 8 && !COMP && sizeof(IT) == 2) {
        if (A.s == 32) {
            dense_coal_r1<BT, 32>(A, b0, b1, x, acc, lane);
            return;
        }
    }
    if constexpr (sizeof(VT) == 8 && !COMP && sizeof(IT) == 1 && sizeof(RT) == 2) {
        if (A.s == 16) {
            dense_coal_r1<BT, 16>(A, b0, b1, x, acc, lane);
            return;
        }
    }
    bmdn_r1<VT, BT, COMP>(A, DENSE, b0, b1, 0, x, acc, lane);
}

// Work items: items [0, n_light) are light items -- whole strips light[p] for
// p in [pack[item], pack[item + 1]) (p = item without pack; strip p without
// light) -- whose rows are written to y; items n_light + u are
// units u of heavy strips: one scheme group's block range, whose partial sums
// go to part[u] and are combined by k_spmv_fixup.
// Workgroups are dispatched to the 8 XCDs round-robin.  Renumber them so that
// each XCD (and its L2) receives runs of kRun consecutive work items (the x
// windows of neighbouring strips stay in one L2) while the runs rotate over
// the XCDs (a skewed range, e.g. the units of heavy strips, is still spread
// over all of them).  A tail shorter than 8 runs keeps the dispatch order.
__device__ inline int64_t xcd_block_id()
{
    constexpr int64_t kXcd = 8, kRun = 32;
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t full = (nb / (kXcd * kRun)) * (kXcd * kRun);
    if (b >= full) return b;
    const int64_t x = b % kXcd, i = b / kXcd;
    return ((i / kRun) * kXcd + x) * kRun + (i % kRun);
}

// One work item of one matrix (see Balance): the wave accumulates in acc_buf
// (s doubles, 2s when compensated) and writes y rows or a unit's partials.
// TL: 0 the matrix is not tiled, 1 it is, 2 either (read from B.tiled); the
// single-matrix kernels are compiled for 0 or 1 so that each carries only one
// of the light-item paths (scalar-register pressure).
template <typename VT, typename BT, typename IT, typename RT, bool COMP, int ZS = 2, int CH = 4, int TL = 2>
__device__ inline void spmv_item(const DevMat &A, const Balance &B, int64_t item, const double *__restrict__ x,
                                 double *__restrict__ y, double alpha, double beta, double *acc_buf,
                                 uint32_t *s_end, uint32_t *s_bc, uint32_t *s_st, uint32_t *s_to, int lane)
{
    const int s = A.s;
    Acc acc;
    acc.hi = acc_buf;
    acc.lo = COMP ? acc_buf + 256 : nullptr;
    for (int i = lane; i < s; i += kWave) {
        acc.hi[i] = 0.0;
        if constexpr (COMP) acc.lo[i] = 0.0;
    }
    wave_sync();
    if (item < B.n_light) {
        // light item: light strips at positions [p0, p1) of the light list
        uint32_t p0 = (uint32_t)item, p1 = (uint32_t)item + 1u;
        if (B.pack) {
            p0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)B.pack[item]);
            p1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)B.pack[item + 1]);
        }
        const bool tiled = TL == 1 || (TL == 2 && B.tiled);
        if (tiled) {
            // the item's strips light[p0 .. p1): one pass over their COO
            // blocks (rows (position << k) + lrow), then each strip's CSR /
            // bitmap / dense groups, then all rows
            const uint32_t *R = B.coo_rng + (uint64_t)item * 3;
            const uint32_t cb0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)R[0]);
            const uint32_t cb1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)R[1]);
            const uint32_t ce0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)R[2]);
            const int T = (int)(p1 - p0);
            const int rows = T * s;
            for (int i = s + lane; i < rows; i += kWave) {
                acc.hi[i] = 0.0;
                if constexpr (COMP) acc.lo[i] = 0.0;
            }
            wave_sync();
            if (T > 1)
                coo_r1<VT, BT, IT, COMP, ZS, CH, true>(A, cb0, cb1, ce0, x, acc, s_end, s_bc, s_st, s_to, lane);
            else
                coo_r1<VT, BT, IT, COMP, ZS, CH>(A, cb0, cb1, ce0, x, acc, s_end, s_bc, s_st, s_to, lane);
#pragma unroll 1
            for (int t = 0; t < T; t++) {
                const int64_t I = (int64_t)__builtin_amdgcn_readfirstlane((int)B.light[p0 + (uint32_t)t]);
                Acc at;
                at.hi = acc.hi + t * s;
                at.lo = COMP ? acc.lo + t * s : nullptr;
                csr_rows_r1<VT, BT, RT, COMP>(A, A.sptr[CSR][I], A.sptr[CSR][I + 1], A.eptr[CSR][I], x, at, s_st,
                                              s_bc, lane);
                wave_sync();
                bmdn_r1<VT, BT, COMP>(A, BITMAP, A.sptr[BITMAP][I], A.sptr[BITMAP][I + 1], A.eptr[BITMAP][I], x,
                                      at, lane);
                dense_r1<VT, BT, IT, RT, COMP>(A, A.sptr[DENSE][I], A.sptr[DENSE][I + 1], x, at, lane);
                wave_sync();
            }
            for (int r = lane; r < rows; r += kWave) {
                const int64_t gr = (int64_t)B.light[p0 + (uint32_t)(r >> A.k)] * s + (r & (s - 1));
                if (gr < A.m) {
                    double a = acc.hi[r];
                    if constexpr (COMP) a += acc.lo[r];
                    y[gr] = (beta == 0.0) ? alpha * a : alpha * a + beta * y[gr];
                }
            }
            wave_sync();
            return;
        }
#pragma unroll 1
        for (uint32_t p = p0; p < p1; p++) {
            const int64_t I = B.light ? (int64_t)__builtin_amdgcn_readfirstlane((int)B.light[p]) : (int64_t)p;
            if (p != p0) {  // the previous strip's rows are written; clear for this one
                for (int i = lane; i < s; i += kWave) {
                    acc.hi[i] = 0.0;
                    if constexpr (COMP) acc.lo[i] = 0.0;
                }
                wave_sync();
            }
            coo_r1<VT, BT, IT, COMP, ZS, CH>(A, A.sptr[COO][I], A.sptr[COO][I + 1], A.eptr[COO][I], x, acc, s_end, s_bc,
                                             s_st, s_to, lane);
            csr_rows_r1<VT, BT, RT, COMP>(A, A.sptr[CSR][I], A.sptr[CSR][I + 1], A.eptr[CSR][I], x, acc, s_st, s_bc,
                                          lane);
            wave_sync();
            bmdn_r1<VT, BT, COMP>(A, BITMAP, A.sptr[BITMAP][I], A.sptr[BITMAP][I + 1], A.eptr[BITMAP][I], x, acc,
                                  lane);
            dense_r1<VT, BT, IT, RT, COMP>(A, A.sptr[DENSE][I], A.sptr[DENSE][I + 1], x, acc, lane);
            wave_sync();
            for (int r = lane; r < s; r += kWave) {
                const int64_t gr = I * s + r;
                if (gr >= A.m) break;
                double a = acc.hi[r];
                if constexpr (COMP) a += acc.lo[r];
                y[gr] = (beta == 0.0) ? alpha * a : alpha * a + beta * y[gr];
            }
            wave_sync();
        }
    } else {
        const int64_t u = item - B.n_light;
        const uint32_t *U = B.units + u * 4;
        const uint32_t sc = U[0] & 3u, b0 = U[1], b1 = U[2], e0 = U[3];
        // constant scheme arguments: the descriptor's per-scheme arrays are
        // then indexed statically (a runtime index would force a local copy
        // of the descriptor, as in the batched kernel, into scratch memory)
        // the tiled kernel's units take one entry per block in phase A (its
        // tiles keep ZS): R-MAT scale 22 s = 64 / 128 / 256 0.623 / 0.590 /
        // 0.536 -> 0.611 / 0.584 / 0.528 ms, same registers
        if (sc == COO)
            coo_r1<VT, BT, IT, COMP, (TL == 1 ? 1 : ZS), CH>(A, b0, b1, e0, x, acc, s_end, s_bc, s_st, s_to, lane);
        else if (sc == CSR)
            csr_rows_r1<VT, BT, RT, COMP>(A, b0, b1, e0, x, acc, s_st, s_bc, lane);
        else if (sc == BITMAP)
            bmdn_r1<VT, BT, COMP>(A, BITMAP, b0, b1, e0, x, acc, lane);
        else
            dense_r1<VT, BT, IT, RT, COMP>(A, b0, b1, x, acc, lane);
        wave_sync();
        double *part = B.part + u * (size_t)s * (COMP ? 2 : 1);
        for (int r = lane; r < s; r += kWave) {
            part[r] = acc.hi[r];
            if constexpr (COMP) part[s + r] = acc.lo[r];
        }
    }
    wave_sync();  // the next item reuses the LDS buffers
}

// Work items: items [0, n_light) are light items -- whole strips light[p] for
// p in [pack[item], pack[item + 1]) (p = item without pack; strip p without
// light) -- whose rows are written to y; items n_light + u are
// units u of heavy strips: one scheme group's block range, whose partial sums
// go to part[u] and are combined by k_spmv_fixup.
template <typename VT, typename BT, typename IT, typename RT, bool COMP, int ZS, int CH, bool TILED>
__global__ __launch_bounds__(64 * kStripWaves) void k_spmv_strip(DevMat A, const double *__restrict__ x,
                                                                 double *__restrict__ y, double alpha, double beta,
                                                                 Balance B)
{
    constexpr int kWaves = kStripWaves;
    // s rows, or a tile's (kTileRows; kTileRowsComp compensated, lo half at +256)
    static_assert(kTileRowsComp == 256 && kTileRows >= 256, "accumulator layout");
    __shared__ double s_acc[kWaves][COMP ? 512 : kTileRows];
    __shared__ uint32_t s_end[kWaves][kWave];
    __shared__ uint32_t s_bc[kWaves][kWave];
    __shared__ uint32_t s_st[kWaves][kWave];
    __shared__ uint32_t s_to[kWaves][kWave];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    int64_t item = xcd_block_id() * kWaves + wave;
    if (B.order) {  // a slot table (column panels: items per XCD; tiles: launch order)
        const uint32_t o = B.order[(int64_t)blockIdx.x * kWaves + wave];
        if (o == kNoItem) return;
        item = (int64_t)__builtin_amdgcn_readfirstlane((int)o);
    }
    if (item >= B.n_light + B.n_units) return;  // whole wave leaves; no workgroup barriers below
    spmv_item<VT, BT, IT, RT, COMP, ZS, CH, TILED ? 1 : 0>(A, B, item, x, y, alpha, beta, s_acc[wave], s_end[wave],
                                                           s_bc[wave], s_st[wave], s_to[wave], lane);
}

// ===========================================================================
// Symmetric storage (one triangle T stored): the main pass computes
// y = alpha*T*x + beta*y, this pass adds scale * T_off^T * x (scale = alpha,
// or -alpha for skew-symmetric; T_off = T without its diagonal): for every
// stored entry (i, j), i != j, y[j] += scale * a_ij * x[i], by global FP64
// atomics (entries of a strip scatter over the column range).  FP64 values.
// ===========================================================================
template <typename BT, typename IT, typename RT>
__device__ inline void trans_group(const DevMat &A, int sc, uint32_t b0, uint32_t b1, uint64_t e, int64_t I,
                                   const double *__restrict__ x, int64_t ldx, double *y, int64_t ldy, double scale,
                                   uint32_t *s_end, uint32_t *s_bc, int lane)
{
    if (b0 >= b1) return;
    const int s = A.s, k = A.k;
    const uint32_t smask = (uint32_t)s - 1, ss = (uint32_t)s * (uint32_t)s;
    const double *__restrict__ val = static_cast<const double *>(A.val[sc]);
    const int64_t r0 = I * s;
    auto put = [&](uint32_t lr, int64_t gc, double v) {
        const int64_t gr = r0 + lr;
        if (gr != gc) atomicAdd(&y[gc * ldy], scale * v * x[gr * ldx]);
    };
    if (sc == COO || sc == CSR) {
        // element-parallel over chunks of 64 blocks; an entry's block by binary
        // search over the chunk's prefix sums, its CSR row over the block's row
        // pointers
#pragma unroll 1
        for (uint32_t cb = b0; cb < b1; cb += kWave) {
            const uint32_t jb = cb + lane;
            const bool ok = jb < b1;
            const uint32_t z = ok ? (uint32_t)A.z1[sc][jb] + 1u : 0u;
            const uint32_t incl = wave_incl_scan_dpp(z);
            const uint32_t T = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
            s_end[lane] = incl;
            s_bc[lane] = ok ? (uint32_t)ld<BT>(A.bcol[sc], jb) : 0u;
            wave_sync();
            const int nbk = (int)min<uint32_t>(kWave, b1 - cb);
#pragma unroll 1
            for (uint32_t q = 0; q < T; q += kWave) {
                const uint32_t t = q + lane;
                if (t < T) {
                    int lo = 0, hi = nbk - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_end[mid] > t) hi = mid; else lo = mid + 1;
                    }
                    uint32_t lr, lc;
                    if (sc == COO) {
                        const uint32_t id = ld<IT>(A.coo_idx, e + t);
                        lr = id >> k;
                        lc = id & smask;
                    } else {
                        const uint32_t qb = t - (lo ? s_end[lo - 1] : 0u);
                        const RT *rp = static_cast<const RT *>(A.csr_rowptr) + (uint64_t)(cb + lo) * (s + 1);
                        int rl = 0, rh = s - 1;
                        while (rl < rh) {
                            const int mid = (rl + rh + 1) >> 1;
                            if ((uint32_t)rp[mid] <= qb) rl = mid; else rh = mid - 1;
                        }
                        lr = (uint32_t)rl;
                        lc = A.csr_col[e + t];
                    }
                    put(lr, (int64_t)s_bc[lo] * s + lc, val[e + t]);
                }
            }
            e += T;
            wave_sync();
        }
        return;
    }
    // bitmap / dense: one block at a time, lanes over its positions
#pragma unroll 1
    for (uint32_t b = b0; b < b1; b++) {
        const int64_t c0 = (int64_t)ld<BT>(A.bcol[sc], b) * s;
        if (sc == DENSE) {
            const uint64_t db = (uint64_t)b * ss;
            for (uint32_t pos = (uint32_t)lane; pos < ss; pos += kWave) {
                const double v = val[db + pos];
                if (v != 0.0) put(pos >> k, c0 + (pos & smask), v);
            }
        } else {
            const uint64_t *__restrict__ words = A.bm_words + (uint64_t)b * A.W;
#pragma unroll 1
            for (int w = 0; w < A.W; w++) {
                const uint64_t word = words[w];
                const uint32_t pos = (uint32_t)w * 64u + (uint32_t)lane;
                if ((word >> lane) & 1ull)
                    put(pos >> k, c0 + (pos & smask), val[e + __popcll(word & ((1ull << lane) - 1ull))]);
                e += (uint64_t)__popcll(word);
            }
        }
    }
}

template <typename BT, typename IT, typename RT>
__global__ __launch_bounds__(256) void k_spmv_trans(DevMat A, Balance B, const double *__restrict__ x, int64_t ldx,
                                                    double *y, int64_t ldy, double scale)
{
    constexpr int kWaves = 4;
    __shared__ uint32_t s_end[kWaves][kWave];
    __shared__ uint32_t s_bc[kWaves][kWave];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * kWaves + wave;
    if (item >= B.n_light + B.n_units) return;
    uint32_t *se = s_end[wave], *sb = s_bc[wave];
    if (item < B.n_light) {
        uint32_t p0 = (uint32_t)item, p1 = (uint32_t)item + 1u;
        if (B.pack) {
            p0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)B.pack[item]);
            p1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)B.pack[item + 1]);
        }
#pragma unroll 1
        for (uint32_t p = p0; p < p1; p++) {
            const int64_t I = B.light ? (int64_t)__builtin_amdgcn_readfirstlane((int)B.light[p]) : (int64_t)p;
#pragma unroll 1
            for (int sc = COO; sc <= DENSE; sc++)
                trans_group<BT, IT, RT>(A, sc, A.sptr[sc][I], A.sptr[sc][I + 1],
                                        sc == DENSE ? 0 : A.eptr[sc][I], I, x, ldx, y, ldy, scale, se, sb, lane);
        }
    } else {
        const uint32_t *U = B.units + (item - B.n_light) * 4;
        const uint32_t u0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)U[0]);
        trans_group<BT, IT, RT>(A, (int)(u0 & 3u), U[1], U[2], U[3], (int64_t)(u0 >> 2), x, ldx, y, ldy, scale, se,
                                sb, lane);
    }
}

// A wave-uniform table field, read through the constant address space (the
// table does not change during a launch): a scalar load into SGPRs.
template <typename T>
__device__ inline T ld_const(const T *p)
{
    return *(const __attribute__((address_space(4))) T *)p;
}

// Copy of a batch table entry, field by field (a word-wise copy through a
// cast pointer kept the whole entry in per-lane scratch memory).
__device__ inline BatchEntry load_entry(const BatchEntry *p)
{
    BatchEntry e;
    e.A.m = ld_const(&p->A.m);
    e.A.n = ld_const(&p->A.n);
    e.A.M = ld_const(&p->A.M);
    e.A.N = ld_const(&p->A.N);
    e.A.s = ld_const(&p->A.s);
    e.A.k = ld_const(&p->A.k);
    e.A.W = ld_const(&p->A.W);
    e.A.zb = ld_const(&p->A.zb);
#pragma unroll
    for (int x = 0; x < 4; x++) {
        e.A.sptr[x] = ld_const(&p->A.sptr[x]);
        e.A.eptr[x] = ld_const(&p->A.eptr[x]);
        e.A.bcol[x] = ld_const(&p->A.bcol[x]);
        e.A.z1[x] = ld_const(&p->A.z1[x]);
        e.A.val[x] = ld_const(&p->A.val[x]);
    }
    e.A.coo_idx = ld_const(&p->A.coo_idx);
    e.A.csr_rowptr = ld_const(&p->A.csr_rowptr);
    e.A.csr_col = ld_const(&p->A.csr_col);
    e.A.bm_words = ld_const(&p->A.bm_words);
    e.B.light = ld_const(&p->B.light);
    e.B.pack = ld_const(&p->B.pack);
    e.B.units = ld_const(&p->B.units);
    e.B.heavy = ld_const(&p->B.heavy);
    e.B.part = ld_const(&p->B.part);
    e.B.order = ld_const(&p->B.order);
    e.B.n_light = ld_const(&p->B.n_light);
    e.B.n_units = ld_const(&p->B.n_units);
    e.B.n_heavy = ld_const(&p->B.n_heavy);
    e.B.n_slots = ld_const(&p->B.n_slots);
    e.B.tiled = ld_const(&p->B.tiled);
    e.B.coo_rng = ld_const(&p->B.coo_rng);
    e.x = ld_const(&p->x);
    e.y = ld_const(&p->y);
    e.item0 = ld_const(&p->item0);
    e.code = ld_const(&p->code);
    e.comp = ld_const(&p->comp);
    return e;
}

// Many matrices in one launch: the waves take the global items [0, G), wave w
// items w, w + W, ... (W = all waves of the grid); item_mat[g] names the
// matrix of item g, whose index-type combination (batch_code) selects the
// decode (a uniform branch), so one launch serves every member.  FP64
// values.  TIMED: each wave also adds the shader-clock cycles of its item to
// cyc[matrix] (the per-matrix cost inside a full launch).
template <bool TIMED, bool COMP>
// (6 waves per SIMD asked of the register allocator: 86 VGPRs allowed 5; the
// 2000-matrix batch of tuned picks 6.11 -> 5.96 ms, plain s = 64 8.98 ->
// 9.00 ms)
__global__ __launch_bounds__(256, 6) void k_batch_spmv(const BatchEntry *__restrict__ E,
                                                    const uint32_t *__restrict__ item_mat, int64_t g0, int64_t G,
                                                    double alpha, double beta, unsigned long long *cyc)
{
    using VT = typename std::conditional<COMP, float, double>::type;
    constexpr int kWaves = 4;
    static_assert(kTileRows >= 2 * kTileRowsComp, "compensated accumulators: hi and lo halves");
    __shared__ double s_acc[kWaves][kTileRows];
    __shared__ uint32_t s_end[kWaves][kWave];
    __shared__ uint32_t s_bc[kWaves][kWave];
    __shared__ uint32_t s_st[kWaves][kWave];
    __shared__ uint32_t s_to[kWaves][kWave];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * kWaves;
#pragma unroll 1
    for (int64_t g = g0 + (int64_t)blockIdx.x * kWaves + wave; g < G; g += nw) {
        const int mat = __builtin_amdgcn_readfirstlane((int)item_mat[g]);
        // the table is read-only during the launch: its fields are read through
        // the constant address space (scalar loads into SGPRs, as the
        // single-matrix kernel gets them from its arguments)
        const BatchEntry e = load_entry(E + mat);
        long long t0 = 0;
        if constexpr (TIMED) t0 = clock64();
        const int64_t it = g - e.item0;
        double *sa = s_acc[wave];
        uint32_t *se = s_end[wave], *sb = s_bc[wave], *ss = s_st[wave], *so = s_to[wave];
        switch (e.code) {  // batch_code: (bcol u32 ? 3 : 0) + (s <= 8 ? 0 : s <= 16 ? 1 : 2)
        case 0:
            spmv_item<VT, uint16_t, uint8_t, uint8_t, COMP>(e.A, e.B, it, e.x, e.y, alpha, beta, sa, se, sb,
                                                                 ss, so, lane);
            break;
        case 1:
            spmv_item<VT, uint16_t, uint8_t, uint16_t, COMP>(e.A, e.B, it, e.x, e.y, alpha, beta, sa, se, sb,
                                                                  ss, so, lane);
            break;
        case 2:
            spmv_item<VT, uint16_t, uint16_t, uint16_t, COMP>(e.A, e.B, it, e.x, e.y, alpha, beta, sa, se,
                                                                   sb, ss, so, lane);
            break;
        case 3:
            spmv_item<VT, uint32_t, uint8_t, uint8_t, COMP>(e.A, e.B, it, e.x, e.y, alpha, beta, sa, se, sb,
                                                                 ss, so, lane);
            break;
        case 4:
            spmv_item<VT, uint32_t, uint8_t, uint16_t, COMP>(e.A, e.B, it, e.x, e.y, alpha, beta, sa, se, sb,
                                                                  ss, so, lane);
            break;
        default:
            spmv_item<VT, uint32_t, uint16_t, uint16_t, COMP>(e.A, e.B, it, e.x, e.y, alpha, beta, sa, se,
                                                                   sb, ss, so, lane);
            break;
        }
        if constexpr (TIMED) {
            const long long t1 = clock64();
            if (lane == 0) atomicAdd(&cyc[mat], (unsigned long long)(t1 - t0));
        }
    }
}

// Rows [r0, r0 + rs) of one heavy strip I: y = alpha * (sum of the partial
// rows of units [u0, u0 + nu)) + beta * y, by a workgroup of kFixThreads
// threads.  The L = kFixThreads / rs threads of row r take units l, l + L, l + 2L, ... in
// kFixChains interleaved chains (that many loads in flight per thread); the
// chains and then the L threads are combined in a fixed order, so the sum
// does not depend on timing.  Partials of unit u: part[u * s * (COMP ? 2 : 1)
// + r] (+ s: the compensation terms).
// (R-MAT scale 22 at s = 256: the heaviest strip holds 584 units and sets the
// pass's length.  SpMV 0.545 ms with 512 threads x 4 chains, 0.532 ms with
// 256 x 16; 256 x 8 and 512 x 16 are in between, 32 chains spill: 0.61+ ms.
// 1024 threads x 4 chains was slower than 512 x 4.)  With compensated
// partials (two loads and a TwoSum per chain step) 16 chains do not stay in
// registers and the loads serialise: fix-up 285 us with 16 chains, 77 us with
// 8, 119 us with 4 (FP64 partials: 38 / 58 / 82 us).
#ifndef ABHSF_FIX_THREADS
#define ABHSF_FIX_THREADS 256
#endif
constexpr int kFixThreads = ABHSF_FIX_THREADS;
static_assert(kFixThreads >= 256, "at least one thread per row of a 256-row strip");
#ifndef ABHSF_FIX_CHAINS
#define ABHSF_FIX_CHAINS 16
#endif
template <bool COMP>
constexpr int kFixChains = COMP ? ABHSF_FIX_CHAINS / 2 : ABHSF_FIX_CHAINS;
static_assert(kFixChains<true> >= 4, "a quarter round of the fix-up takes one chain at least");
// Rows per fix-up workgroup: a strip of s > kFixRows rows is summed by s /
// kFixRows workgroups, kFixThreads / kFixRows threads per row.  R-MAT scale
// 22 at s = 256 (heaviest strip 585 units), fix-up with 256 / 128 / 64-row
// slices: FP64 39 / 31.5 / 36 us, compensated 91 / 48 / 40 us.
template <bool COMP>
constexpr int kFixRows = COMP ? 64 : 128;
static_assert(kFixThreads % 128 == 0, "whole threads per row");
// A heavy strip whose units would leave more than kFixUnits of them to each
// thread gets narrower slices (rows halved down to kFixRowsMin, threads per
// row doubled): the strips with the most units set the pass's length.
#ifndef ABHSF_FIX_UNITS
#define ABHSF_FIX_UNITS 32
#endif
constexpr int kFixUnits = ABHSF_FIX_UNITS;
constexpr int kFixRowsMin = 16;

template <bool COMP>
__device__ inline void fix_add(double &hi, double &lo, double v, double vlo)
{
    if constexpr (COMP) {
        double s2, e2;
        two_sum(hi, v, s2, e2);
        hi = s2;
        lo += vlo + e2;
    } else {
        hi += v;
    }
}

// One round of the fix-up's chains 0 .. W - 1 (units u + j * L; a chain past
// nu reads unit l and adds 0).
template <bool COMP, int W>
__device__ inline void fix_round(const double *__restrict__ part, size_t stride, uint32_t u0, uint32_t u, uint32_t l,
                                 uint32_t L, uint32_t nu, int s, int r, double *hi, double *lo)
{
#pragma unroll
    for (int j = 0; j < W; j++) {
        const uint32_t uj = u + (uint32_t)j * L;
        const bool ok = uj < nu;
        const double *p = part + (size_t)(u0 + (ok ? uj : l)) * stride;
        const double v = p[r], vlo = COMP ? p[s + r] : 0.0;
        fix_add<COMP>(hi[j], lo[j], ok ? v : 0.0, ok ? vlo : 0.0);
    }
}

template <bool COMP>
__device__ inline void fixup_strip(int s, int rs, int r0, int64_t m, int64_t I, const double *__restrict__ part,
                                   uint32_t u0, uint32_t nu, double *__restrict__ y, double alpha, double beta,
                                   double *sh)
{
    const uint32_t L = (uint32_t)(kFixThreads / rs);
    // consecutive threads take consecutive rows of one unit (coalesced loads)
    const int r = r0 + (int)(threadIdx.x % (uint32_t)rs);
    const uint32_t l = threadIdx.x / (uint32_t)rs;
    const size_t stride = (size_t)s * (COMP ? 2 : 1);
    double hi[kFixChains<COMP>], lo[kFixChains<COMP>];
#pragma unroll
    for (int j = 0; j < kFixChains<COMP>; j++) hi[j] = lo[j] = 0.0;
    // chain j takes units u + j * L: full rounds of kFixChains loads, then
    // one round of kFixChains, half or a quarter of them, whichever covers
    // the units left.  Its chains past nu read unit l again and add 0: no
    // branch, so a round's loads are all in flight together (a branch per
    // chain serialised them: R-MAT scale 22 s = 256 with 128-row slices,
    // SpMV 504 -> 584 us).
    constexpr int C = kFixChains<COMP>;
    uint32_t u = l;
    for (; u + (uint32_t)(C - 1) * L < nu; u += (uint32_t)C * L)
        fix_round<COMP, C>(part, stride, u0, u, l, L, nu, s, r, hi, lo);
    if (u < nu) {
        const uint32_t left = (nu - u + L - 1) / L;  // 1 .. C - 1
        if (left > (uint32_t)C / 2) fix_round<COMP, C>(part, stride, u0, u, l, L, nu, s, r, hi, lo);
        else if (left > (uint32_t)C / 4) fix_round<COMP, C / 2>(part, stride, u0, u, l, L, nu, s, r, hi, lo);
        else fix_round<COMP, C / 4>(part, stride, u0, u, l, L, nu, s, r, hi, lo);
    }
#pragma unroll
    for (int j = 1; j < kFixChains<COMP>; j++) fix_add<COMP>(hi[0], lo[0], hi[j], lo[j]);
    sh[threadIdx.x] = hi[0];
    if constexpr (COMP) sh[kFixThreads + threadIdx.x] = lo[0];
    __syncthreads();
    if (l == 0) {
        const uint32_t lr = (uint32_t)(r - r0);
        double H = 0.0, Lo = 0.0;
        for (uint32_t t = 0; t < L; t++)
            fix_add<COMP>(H, Lo, sh[t * (uint32_t)rs + lr], COMP ? sh[kFixThreads + t * (uint32_t)rs + lr] : 0.0);
        const int64_t gr = I * s + r;
        if (gr < m) {
            const double a = H + Lo;
            y[gr] = (beta == 0.0) ? alpha * a : alpha * a + beta * y[gr];
        }
    }
}

// Heavy strips of every matrix: hv[2h] = matrix, hv[2h+1] = heavy index.
// TIMED: thread 0 adds the workgroup's cycles times its wave count.
template <bool TIMED>
__global__ __launch_bounds__(kFixThreads) void k_batch_fixup(const BatchEntry *__restrict__ E,
                                                             const uint32_t *__restrict__ hv, double alpha,
                                                             double beta, unsigned long long *cyc)
{
    __shared__ double sh[2 * kFixThreads];
    long long t0 = 0;
    if constexpr (TIMED) t0 = clock64();
    const uint32_t mat = hv[2 * blockIdx.x];
    const BatchEntry &e = E[mat];
    const uint32_t h = hv[2 * blockIdx.x + 1];
    if (e.comp)  // workgroup-uniform
        fixup_strip<true>(e.A.s, e.A.s, 0, e.A.m, e.B.heavy[h * 3], e.B.part, e.B.heavy[h * 3 + 1],
                          e.B.heavy[h * 3 + 2], e.y, alpha, beta, sh);
    else
        fixup_strip<false>(e.A.s, e.A.s, 0, e.A.m, e.B.heavy[h * 3], e.B.part, e.B.heavy[h * 3 + 1],
                           e.B.heavy[h * 3 + 2], e.y, alpha, beta, sh);
    if constexpr (TIMED) {
        __syncthreads();
        if (threadIdx.x == 0)
            atomicAdd(&cyc[mat], (unsigned long long)(clock64() - t0) * (unsigned long long)(kFixThreads / kWave));
    }
}

// Rows of heavy strips (one workgroup per heavy strip).
template <bool COMP>
__global__ __launch_bounds__(kFixThreads) void k_spmv_fixup(int s, int64_t m, double *__restrict__ y, double alpha,
                                                            double beta, Balance B)
{
    __shared__ double sh[kFixThreads * (COMP ? 2 : 1)];
    // workgroup b sums rows [r0, r0 + rs) of heavy strip h (fix_rows)
    const uint32_t h = B.fix[2 * (size_t)blockIdx.x], w = B.fix[2 * (size_t)blockIdx.x + 1];
    const int r0 = (int)(w & 0xFFFFu), rs = (int)(w >> 16);
    fixup_strip<COMP>(s, rs, r0, m, B.heavy[h * 3], B.part, B.heavy[h * 3 + 1], B.heavy[h * 3 + 2], y, alpha, beta,
                      sh);
}

// LDS bytes per wave of the SpMM kernel: the s x R accumulator (twice when
// compensated), the decoded COO window (value, row, column) and the chunk's
// prefix sums and block columns.
__host__ __device__ inline size_t spmm_wave_lds(int s, int R, bool comp)
{
    return (size_t)s * R * sizeof(double) * (comp ? 2 : 1) + kWave * (sizeof(double) + 4 * sizeof(uint32_t));
}

// ===========================================================================
// Several right-hand sides (SpMM), R lanes per entry.
// ===========================================================================
// Dense s x s block times an s x R slice of X on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: A 16x4 with lane l holding A[l%16][l/16], B 4x16
// with lane l holding B[l/16][l%16], D 16x16 with register r of lane l
// holding D[4r + l/16][l%16], r = 0..3; layout measured on gfx950).  The 16 x 16 product tiles are added to the
// strip accumulator (one lane per element, so no atomics).
template <typename VT, int R>
__device__ inline void dense_block_mfma(const VT *__restrict__ D, int s, int64_t c0, int64_t n,
                                        const double *__restrict__ X, int64_t ldx, int nrhs, double *acc, int lane)
{
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll 1
    for (int r0 = 0; r0 < s; r0 += 16) {
#pragma unroll
        for (int j0 = 0; j0 < R; j0 += 16) {
            typedef double v4d __attribute__((ext_vector_type(4)));
            v4d c = {0.0, 0.0, 0.0, 0.0};
            const int j = j0 + li;
#pragma unroll 4
            for (int k0 = 0; k0 < s; k0 += 4) {
                const double a = (double)D[(int64_t)(r0 + li) * s + k0 + lk];
                const int64_t gc = c0 + k0 + lk;
                const double bx = (gc < n && j < nrhs) ? X[gc * ldx + j] : 0.0;
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bx, c, 0, 0, 0);
            }
            // v_mfma_f64_16x16x4f64 result: register r of lane l holds row 4r + l/16, column l%16
            acc[(r0 + lk) * R + j] += c[0];
            acc[(r0 + 4 + lk) * R + j] += c[1];
            acc[(r0 + 8 + lk) * R + j] += c[2];
            acc[(r0 + 12 + lk) * R + j] += c[3];
        }
    }
}

// CSR group of the SpMM kernel, strip I.  Lanes own block rows as in the
// single-vector CSR walk (for s <= 64, lane l takes row l % s of block l / s
// of each step; for s > 64, one block per step in s / 64 row passes) and hold
// the row's R right-hand sides in registers: each entry's column and value
// are loaded once and its R values of X are read back to back (16-byte loads
// where X allows).  Row sums carry over the strip's blocks; the G = 64 / s
// lanes that share a row (s < 64) are summed with xor-shuffles, and the row's
// owner adds the sums to the accumulator (one writer per element: no atomics).
// X holds the wave's R columns (nrhs of them valid) at stride ldx.
// 27-point stencil 160^3, s = 64, FP64, k = 8 / k = 32: 1.10 / 4.07 ms, against
// 1.39 / 4.94 ms for a walk of 8 rows of R lanes per step (s * R / 64 chains
// of dependent loads per block).  More loads in flight per lane were slower
// throughout: two blocks per lane 1.43 / 5.87 ms, 2 or 4 entries per chain
// 1.38 / 5.10 and 1.62 / 6.04 ms, the next block's row pointers prefetched
// 1.22 / 4.52 ms -- the X reads queue in the memory pipeline (address unit
// busy 65 % of the time), not in each wave's dependency chain.
template <typename VT, typename BT, typename RT, int R, bool COMP>
__device__ inline void spmm_csr_rows(const DevMat &A, int64_t I, const double *__restrict__ X, int64_t ldx,
                                     int nrhs, const Acc &acc, uint32_t *s_st, uint32_t *s_bc, int lane)
{
    const uint32_t b0 = A.sptr[CSR][I], b1 = A.sptr[CSR][I + 1];
    if (b0 >= b1) return;
    uint32_t e = A.eptr[CSR][I];
    const int s = A.s;
    const int sw = s < kWave ? s : kWave;  // lanes per block
    const int G = kWave / sw;              // blocks per wave step
    const int passes = s / sw;             // row passes per block
    const int lb = lane / sw, lr = lane % sw;
    const int nr = nrhs < R ? nrhs : R;
    // whole tile, X rows 16-byte aligned: R / 2 double2 loads per entry
    const bool vec = nr == R && ((reinterpret_cast<uintptr_t>(X) & 15u) == 0) && ((ldx & 1) == 0);
    const VT *__restrict__ val = static_cast<const VT *>(A.val[CSR]);
    const uint8_t *__restrict__ ccol = A.csr_col;
    const RT *__restrict__ rowptr = static_cast<const RT *>(A.csr_rowptr);
#pragma unroll 1
    for (uint32_t cb = b0; cb < b1; cb += kWave) {
        const uint32_t jb = cb + lane;
        const bool ok = jb < b1;
        const uint32_t z = ok ? (uint32_t)A.z1[CSR][jb] + 1u : 0u;
        const uint32_t incl = wave_incl_scan_dpp(z);
        s_st[lane] = e + incl - z;
        s_bc[lane] = ok ? (uint32_t)ld<BT>(A.bcol[CSR], jb) : 0u;
        e += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        wave_sync();
        const int nbk = (int)min<uint32_t>(kWave, b1 - cb);
#pragma unroll 1
        for (int ps = 0; ps < passes; ps++) {
            const int r = ps * sw + lr;
            double hi[R], lo[R];
#pragma unroll
            for (int jj = 0; jj < R; jj++) hi[jj] = lo[jj] = 0.0;
#pragma unroll 1
            for (int j0 = 0; j0 < nbk; j0 += G) {
                const int jblk = j0 + lb;
                const bool h = jblk < nbk;
                const RT *rp = rowptr + (uint64_t)(cb + (uint32_t)(h ? jblk : 0)) * (uint64_t)(s + 1);
                const uint32_t ra = h ? (uint32_t)rp[r] : 0u, rc = h ? (uint32_t)rp[r + 1] : 0u;
                const uint32_t q0 = h ? s_st[jblk] + ra : 0u, ql = rc - ra;
                const int64_t xc = h ? (int64_t)s_bc[jblk] * s : 0;
                // (the loads below carry the t < ql predicate although it holds
                // inside the loop: that form measured 5 % faster, k = 32 4.07
                // against 4.27 ms)
                const uint32_t len = max(0u, ql);
#pragma unroll 1
                for (uint32_t t = 0; t < len; t++) {
                    double v, xv[R];
                    int64_t xr;
                    const bool hh = t < ql;
                    xr = hh ? (xc + ccol[q0 + t]) * ldx : 0;
                    v = hh ? (double)val[q0 + t] : 0.0;
                    bool done = false;
                    if constexpr (R >= 2) {
                        if (vec) {  // 16-byte loads: half the address work per value
                            const double2 *xp = reinterpret_cast<const double2 *>(X + xr);
#pragma unroll
                            for (int jj = 0; jj < R; jj += 2) {
                                double2 w = {0.0, 0.0};
                                if (hh) w = xp[jj >> 1];
                                xv[jj] = w.x;
                                xv[jj + 1] = w.y;
                            }
                            done = true;
                        }
                    }
                    if (!done) {
#pragma unroll
                        for (int jj = 0; jj < R; jj++) xv[jj] = (hh && jj < nr) ? X[xr + jj] : 0.0;
                    }
                    if (hh) {
#pragma unroll
                        for (int jj = 0; jj < R; jj++) acc_add<COMP>(hi[jj], lo[jj], v, xv[jj]);
                    }
                }
            }
            for (int o = sw; o < kWave; o <<= 1) {
#pragma unroll
                for (int jj = 0; jj < R; jj++) {
                    const double q = __shfl_xor(hi[jj], o, kWave);
                    if constexpr (COMP) {
                        const double qe = __shfl_xor(lo[jj], o, kWave);
                        double sum, err;
                        two_sum(hi[jj], q, sum, err);
                        hi[jj] = sum;
                        lo[jj] += qe + err;
                    } else {
                        hi[jj] += q;
                    }
                }
            }
            if (lb == 0) {
#pragma unroll
                for (int jj = 0; jj < R; jj++) {
                    if (jj >= nr) break;
                    const uint32_t ai = (uint32_t)r * R + jj;
                    if constexpr (COMP) {
                        double sum, err;
                        two_sum(acc.hi[ai], hi[jj], sum, err);
                        acc.hi[ai] = sum;
                        acc.lo[ai] += err + lo[jj];
                    } else {
                        acc.hi[ai] += hi[jj];
                    }
                }
            }
        }
        wave_sync();
    }
}

// Products of a decoded window of nq <= 64 entries (rows d_row, X rows
// d_col, values d_val, in LDS) with the wave's R columns of X, added to the
// strip accumulator with LDS atomics.  With a whole, 16-byte aligned tile
// (vec) H = R / 2 lanes take an entry, a double2 of X each, so H steps of
// 64 / H entries cover the window and all of its X loads are issued before
// its adds (one memory round trip per window; R-lane steps of one column per
// lane, below, wait for each step's loads in turn).
template <int R, bool COMP>
__device__ inline void spmm_window(uint32_t nq, const uint32_t *d_row, const uint32_t *d_col, const double *d_val,
                                   const double *__restrict__ X, int64_t ldx, int nrhs, bool vec, const Acc &acc,
                                   int lane)
{
    if constexpr (R >= 2) {
        if (vec) {
            constexpr int H = R / 2, E2 = kWave / H;
            const int slot2 = lane / H, c2 = lane % H;
            double2 xw[H];
#pragma unroll
            for (int g = 0; g < H; g++) {
                const uint32_t ent = (uint32_t)(g * E2 + slot2);
                xw[g] = double2{0.0, 0.0};
                if (ent < nq) xw[g] = reinterpret_cast<const double2 *>(X + (int64_t)d_col[ent] * ldx)[c2];
            }
            if constexpr (COMP) {
                // every returning add of the window first, then the TwoSums and
                // the compensation adds: the hi and lo arrays each receive their
                // adds in the same order as one product at a time, so the sums
                // are the same, but the returns are awaited once
                double p[2 * H], pe[2 * H], old[2 * H];
                uint32_t ai[2 * H];
#pragma unroll
                for (int g = 0; g < H; g++) {
                    const uint32_t ent = (uint32_t)(g * E2 + slot2);
                    const bool ok = ent < nq;
                    const double v = ok ? d_val[ent] : 0.0;
                    ai[2 * g] = ok ? d_row[ent] * R + 2 * c2 : 0u;
                    ai[2 * g + 1] = ai[2 * g] + 1;
                    p[2 * g] = prod<true>(v, xw[g].x, pe[2 * g]);
                    p[2 * g + 1] = prod<true>(v, xw[g].y, pe[2 * g + 1]);
                }
#pragma unroll
                for (int i = 0; i < 2 * H; i++)
                    if ((uint32_t)((i >> 1) * E2 + slot2) < nq) old[i] = atomicAdd(&acc.hi[ai[i]], p[i]);
#pragma unroll
                for (int i = 0; i < 2 * H; i++)
                    if ((uint32_t)((i >> 1) * E2 + slot2) < nq) {
                        double sum, err;
                        two_sum(old[i], p[i], sum, err);
                        atomicAdd(&acc.lo[ai[i]], err + pe[i]);
                    }
                return;
            }
#pragma unroll
            for (int g = 0; g < H; g++) {
                const uint32_t ent = (uint32_t)(g * E2 + slot2);
                if (ent < nq) {
                    const double v = d_val[ent];
                    const uint32_t ai = d_row[ent] * R + 2 * c2;
                    double pe;
                    double p = prod<COMP>(v, xw[g].x, pe);
                    acc_atomic<COMP>(acc, ai, p, pe);
                    p = prod<COMP>(v, xw[g].y, pe);
                    acc_atomic<COMP>(acc, ai + 1, p, pe);
                }
            }
            return;
        }
    }
    constexpr int E = kWave / R;  // entries per step
    const int slot = lane / R, j = lane % R;
#pragma unroll 1
    for (uint32_t g = 0; g < nq; g += E) {
        const uint32_t ent = g + slot;
        if (ent < nq && j < nrhs) {
            const double xv = X[(int64_t)d_col[ent] * ldx + j];
            double pe;
            const double p = prod<COMP>(d_val[ent], xv, pe);
            acc_atomic<COMP>(acc, d_row[ent] * R + j, p, pe);
        }
    }
}

// Two launch shapes: sib == 0, wave w of workgroup b takes strip b * wpb + w
// for the R right-hand sides at X, Y; sib > 0, the workgroup's sib waves all
// take strip b, wave w the R-column tile w of the sib * R columns (nrhs in
// all): the siblings read the same matrix entries together, so the matrix
// streams from HBM once for all tiles and the tiles' X values share cache
// lines in the CU's L1.
template <typename VT, typename BT, typename IT, typename RT, int R, bool COMP>
__global__ __launch_bounds__(256) void k_spmm_strip(DevMat A, const double *__restrict__ X, int64_t ldx,
                                                    double *__restrict__ Y, int64_t ldy, int nrhs,
                                                    double alpha, double beta, int wpb, int64_t nstrip_total,
                                                    int sib)
{
    constexpr int E = kWave / R;  // entries per wave step
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = lane / R, j = lane % R;
    const int s = A.s, k = A.k;
    int64_t I = (int64_t)blockIdx.x * wpb + wave;
    if (sib) {
        I = blockIdx.x;
        X += (int64_t)wave * R;
        Y += (int64_t)wave * R;
        nrhs -= wave * R;
        if (nrhs <= 0) return;  // whole wave leaves; no workgroup barriers below
    }
    const bool jok = j < nrhs;
    if (I >= nstrip_total) return;  // whole wave leaves; no workgroup barriers below
    // whole tile, X rows 16-byte aligned: the decoded windows read X as double2
    const bool vec = nrhs >= R && ((reinterpret_cast<uintptr_t>(X) & 15u) == 0) && ((ldx & 1) == 0);

    const size_t per_wave = spmm_wave_lds(s, R, COMP);
    unsigned char *base = smem + per_wave * wave;
    Acc acc;
    acc.hi = reinterpret_cast<double *>(base);
    acc.lo = COMP ? acc.hi + (size_t)s * R : nullptr;
    double *d_val = reinterpret_cast<double *>(base + (size_t)s * R * sizeof(double) * (COMP ? 2 : 1));
    uint32_t *s_end = reinterpret_cast<uint32_t *>(d_val + kWave);
    uint32_t *s_bc = s_end + kWave;
    uint32_t *d_row = s_bc + kWave;
    uint32_t *d_col = d_row + kWave;

    for (int i = lane; i < s * R; i += kWave) {
        acc.hi[i] = 0.0;
        if constexpr (COMP) acc.lo[i] = 0.0;
    }
    wave_sync();

    const uint32_t smask = (uint32_t)s - 1;

    // ---- COO group: windows of 64 entries are decoded once (one entry per
    // lane: block by binary search over the chunk's prefix sums, local row and
    // global column, value) into LDS, then E = 64/R entries per step times R
    // right-hand sides read the decoded window -------------------------------
    {
        const uint32_t b0 = A.sptr[COO][I], b1 = A.sptr[COO][I + 1];
        uint32_t e = A.eptr[COO][I];
        const VT *val = static_cast<const VT *>(A.val[COO]);
#pragma unroll 1
        for (uint32_t cb = b0; cb < b1; cb += kWave) {
            const uint32_t jb = cb + lane;
            const bool ok = jb < b1;
            const uint32_t z = ok ? (uint32_t)A.z1[COO][jb] + 1u : 0u;
            const uint32_t incl = wave_incl_scan_dpp(z);
            const uint32_t T = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
            s_end[lane] = incl;
            s_bc[lane] = ok ? (uint32_t)ld<BT>(A.bcol[COO], jb) : 0u;
            wave_sync();
            const int nbk = (int)min<uint32_t>(kWave, b1 - cb);
#pragma unroll 1
            for (uint32_t q = 0; q < T; q += kWave) {
                const uint32_t t = q + lane;
                if (t < T) {
                    int lo = 0, hi = nbk - 1;  // first chunk block with end > t
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (s_end[mid] > t) hi = mid; else lo = mid + 1;
                    }
                    const uint32_t id = ld<IT>(A.coo_idx, e + t);
                    d_row[lane] = id >> k;
                    d_col[lane] = s_bc[lo] * (uint32_t)s + (id & smask);
                    d_val[lane] = (double)val[e + t];
                }
                wave_sync();
                spmm_window<R, COMP>(min<uint32_t>(kWave, T - q), d_row, d_col, d_val, X, ldx, nrhs, vec, acc, lane);
                wave_sync();
            }
            e += T;
        }
    }

    // ---- CSR group: lanes own block rows, R right-hand sides per lane in
    // registers across the strip's blocks
    spmm_csr_rows<VT, BT, RT, R, COMP>(A, I, X, ldx, nrhs, acc, s_end, s_bc, lane);
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
            if constexpr (R >= 16 && !COMP) {
                if (sc == DENSE && s >= 16) {
                    dense_block_mfma<VT, R>(val + vbase, s, c0, A.n, X, ldx, nrhs, acc.hi, lane);
                    continue;
                }
            }
            if constexpr (R >= 2) {
                if (sc == BITMAP) {
                    // decode each word's set bits (in value order) into the
                    // window, then E entries x R right-hand sides per step:
                    // no step is spent on empty positions
#pragma unroll 1
                    for (int w = 0; w < W; w++) {
                        const uint64_t word = words[w];
                        const uint32_t nq = (uint32_t)__popcll(word);
                        if ((word >> lane) & 1ull) {
                            const uint32_t ix = (uint32_t)__popcll(word & ((1ull << lane) - 1ull));
                            const uint32_t pos = (uint32_t)w * 64u + (uint32_t)lane;
                            d_row[ix] = pos >> k;
                            d_col[ix] = (uint32_t)(c0 + (pos & smask));
                            d_val[ix] = (double)val[vbase + ix];
                        }
                        wave_sync();
                        spmm_window<R, COMP>(nq, d_row, d_col, d_val, X, ldx, nrhs, vec, acc, lane);
                        wave_sync();
                        vbase += nq;
                    }
                    continue;
                }
            }
#pragma unroll 1
            for (int w = 0; w < W; w++) {
                const uint64_t word = (sc == BITMAP) ? words[w] : ~0ull;
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

// nrhs <= R: one tile, wpb strips per workgroup; nrhs > R (at most
// kSpmmSiblings * R): sibling waves, one strip per workgroup, one R-column
// tile per wave.
constexpr int kSpmmSiblings = 4;

template <typename VT, typename BT, typename IT, typename RT, int R, bool COMP>
void launch_t(const DevMat &A, const double *X, int64_t ldx, double *Y, int64_t ldy, int nrhs, double alpha,
              double beta, hipStream_t st)
{
    const size_t per_wave = spmm_wave_lds(A.s, R, COMP);
    const int sib = nrhs > R ? (nrhs + R - 1) / R : 0;
    ABHSF_REQUIRE(sib <= kSpmmSiblings, ABHSF_ERR_UNSUPPORTED, "more column tiles than sibling waves");
    int wpb = sib ? sib : (int)(65536 / per_wave);
    if (wpb > 4) wpb = 4;
    ABHSF_REQUIRE(wpb >= 1 && per_wave * wpb <= 65536, ABHSF_ERR_UNSUPPORTED, "output tile does not fit in LDS");
    const int64_t nblk = sib ? A.M : (A.M + wpb - 1) / wpb;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "too many strips");
    if (A.M == 0) return;
    hipLaunchKernelGGL((k_spmm_strip<VT, BT, IT, RT, R, COMP>), dim3((unsigned)nblk), dim3(64 * wpb),
                       per_wave * wpb, st, A, X, ldx, Y, ldy, nrhs, alpha, beta, wpb, A.M, sib);
    ABHSF_HIP(hipGetLastError());
}

template <typename VT, typename BT, typename IT, typename RT, bool COMP, bool TILED>
void launch_strip(const DevMat &A, const Balance &b, int64_t nblk, bool coo_single, const double *x, double *y,
                  double alpha, double beta, hipStream_t st)
{
    // COO blocks of (almost) one entry: phase A takes one entry from each of
    // 8 x 64 blocks per step (more gathers in flight, no unused entry slot);
    // otherwise two entries from each of 4 x 64 blocks.  Tiles always take
    // the 4 x 64 steps: a tile's COO stream advances through x in smaller
    // column ranges per step (block_mix 0.480 -> 0.474 ms, uniform scatter
    // 0.624 -> 0.606 ms).
    if constexpr (!TILED) {
        if (coo_single) {
            hipLaunchKernelGGL((k_spmv_strip<VT, BT, IT, RT, COMP, 1, 8, false>), dim3((unsigned)nblk),
                               dim3(64 * kStripWaves), 0, st, A, x, y, alpha, beta, b);
            return;
        }
    }
    if constexpr (TILED) {
        if (b.tile_zs1) {
            hipLaunchKernelGGL((k_spmv_strip<VT, BT, IT, RT, COMP, 1, 4, true>), dim3((unsigned)nblk),
                               dim3(64 * kStripWaves), 0, st, A, x, y, alpha, beta, b);
            return;
        }
    }
    hipLaunchKernelGGL((k_spmv_strip<VT, BT, IT, RT, COMP, 2, 4, TILED>), dim3((unsigned)nblk), dim3(64 * kStripWaves),
                       0, st, A, x, y, alpha, beta, b);
}

template <typename VT, typename BT, typename IT, typename RT, bool COMP>
void launch_spmv(const DevMat &A, const Balance &B, bool coo_single, const double *x, double *y, double alpha,
                 double beta, hipStream_t st)
{
    if (A.M == 0) return;
    Balance b = B;
    if (!b.light && b.n_units == 0) b.n_light = A.M;  // one item per strip
    const int64_t nitems = b.n_light + b.n_units;
    const int64_t nblk = b.order ? b.n_slots / kStripWaves : (nitems + kStripWaves - 1) / kStripWaves;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "too many strips");
    // tiled matrices have their own kernels (each carries one light-item
    // path); the choice is made once, here: a tiled storage read by an untiled
    // kernel would take its z1 strip bits for entry counts
    if (b.tiled)
        launch_strip<VT, BT, IT, RT, COMP, true>(A, b, nblk, coo_single, x, y, alpha, beta, st);
    else
        launch_strip<VT, BT, IT, RT, COMP, false>(A, b, nblk, coo_single, x, y, alpha, beta, st);
    ABHSF_HIP(hipGetLastError());
    if (b.n_heavy) {
        ABHSF_REQUIRE(b.n_fix > 0 && b.n_fix < (1ll << 31), ABHSF_ERR_STATE, "fix-up table missing");
        hipLaunchKernelGGL((k_spmv_fixup<COMP>), dim3((unsigned)b.n_fix), dim3(kFixThreads), 0, st, A.s, A.m, y,
                           alpha, beta, b);
        ABHSF_HIP(hipGetLastError());
    }
}

template <typename VT, typename BT, typename IT, typename RT, bool COMP>
void launch_r(int R, const DevMat &A, const Balance &Bal, bool coo_single, const double *X, int64_t ldx, double *Y,
              int64_t ldy, int nrhs, double alpha, double beta, hipStream_t st)
{
    switch (R) {
    case 1:
        if (nrhs == 1 && ldx == 1 && ldy == 1)
            launch_spmv<VT, BT, IT, RT, COMP>(A, Bal, coo_single, X, Y, alpha, beta, st);
        else launch_t<VT, BT, IT, RT, 1, COMP>(A, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        break;
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
    const Balance &Bal = L.bal;
    const bool cs = L.coo_single;
    const bool b16 = L.bcol_bytes == 2;
    const int s = A.s;
    if (s <= 8) {
        if (b16) launch_r<VT, uint16_t, uint8_t, uint8_t, COMP>(R, A, Bal, cs, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        else launch_r<VT, uint32_t, uint8_t, uint8_t, COMP>(R, A, Bal, cs, X, ldx, Y, ldy, nrhs, alpha, beta, st);
    } else if (s <= 16) {
        if (b16) launch_r<VT, uint16_t, uint8_t, uint16_t, COMP>(R, A, Bal, cs, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        else launch_r<VT, uint32_t, uint8_t, uint16_t, COMP>(R, A, Bal, cs, X, ldx, Y, ldy, nrhs, alpha, beta, st);
    } else {
        if (b16) launch_r<VT, uint16_t, uint16_t, uint16_t, COMP>(R, A, Bal, cs, X, ldx, Y, ldy, nrhs, alpha, beta, st);
        else launch_r<VT, uint32_t, uint16_t, uint16_t, COMP>(R, A, Bal, cs, X, ldx, Y, ldy, nrhs, alpha, beta, st);
    }
}

}  // namespace

int fix_rows(int s, uint32_t nu, bool comp)
{
    int rs = std::min(s, comp ? kFixRows<true> : kFixRows<false>);
    while (rs > kFixRowsMin && (uint64_t)nu * (uint64_t)rs > (uint64_t)kFixUnits * (uint64_t)kFixThreads) rs /= 2;
    return rs;
}

int batch_code(int bcol_bytes, int s) { return (bcol_bytes == 4 ? 3 : 0) + (s <= 8 ? 0 : (s <= 16 ? 1 : 2)); }

void batch_spmv(const BatchEntry *dE, const uint32_t *item_mat, int64_t G64, int64_t G, const uint32_t *dheavy,
                int64_t nheavy, double alpha, double beta, hipStream_t st, unsigned long long *cycles)
{
    auto launch = [&](auto comp_tag, int64_t g0, int64_t g1) {
        constexpr bool C = decltype(comp_tag)::value;
        if (g1 <= g0) return;
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((g1 - g0 + 3) / 4, 8192));
        if (cycles)
            hipLaunchKernelGGL((k_batch_spmv<true, C>), dim3((unsigned)grid), dim3(256), 0, st, dE, item_mat, g0, g1,
                               alpha, beta, cycles);
        else
            hipLaunchKernelGGL((k_batch_spmv<false, C>), dim3((unsigned)grid), dim3(256), 0, st, dE, item_mat, g0,
                               g1, alpha, beta, cycles);
        ABHSF_HIP(hipGetLastError());
    };
    launch(std::integral_constant<bool, false>{}, 0, G64);
    launch(std::integral_constant<bool, true>{}, G64, G);
    if (nheavy > 0) {
        if (cycles)
            hipLaunchKernelGGL(k_batch_fixup<true>, dim3((unsigned)nheavy), dim3(kFixThreads), 0, st, dE, dheavy,
                               alpha, beta, cycles);
        else
            hipLaunchKernelGGL(k_batch_fixup<false>, dim3((unsigned)nheavy), dim3(kFixThreads), 0, st, dE, dheavy,
                               alpha, beta, cycles);
        ABHSF_HIP(hipGetLastError());
    }
}

void spmv_transposed(const SpmvLaunch &L, const DevMat &A, int nrhs, double scale, const double *X, int64_t ldx,
                     double *Y, int64_t ldy, hipStream_t st)
{
    ABHSF_REQUIRE(L.val_bytes == 8, ABHSF_ERR_UNSUPPORTED, "symmetric storage needs FP64 values");
    if (A.M == 0) return;
    Balance b = L.bal;
    if (!b.light && b.n_units == 0) b.n_light = A.M;
    const int64_t nblk = (b.n_light + b.n_units + 3) / 4;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "too many strips");
    const bool b16 = L.bcol_bytes == 2;
    for (int j = 0; j < nrhs; j++) {
        const double *x = X + j;
        double *y = Y + j;
#define ABHSF_TRANS(BT_, IT_, RT_)                                                                             \
    hipLaunchKernelGGL((k_spmv_trans<BT_, IT_, RT_>), dim3((unsigned)nblk), dim3(256), 0, st, A, b, x, ldx, y, \
                       ldy, scale)
        if (A.s <= 8) {
            if (b16) ABHSF_TRANS(uint16_t, uint8_t, uint8_t); else ABHSF_TRANS(uint32_t, uint8_t, uint8_t);
        } else if (A.s <= 16) {
            if (b16) ABHSF_TRANS(uint16_t, uint8_t, uint16_t); else ABHSF_TRANS(uint32_t, uint8_t, uint16_t);
        } else {
            if (b16) ABHSF_TRANS(uint16_t, uint16_t, uint16_t); else ABHSF_TRANS(uint32_t, uint16_t, uint16_t);
        }
#undef ABHSF_TRANS
        ABHSF_HIP(hipGetLastError());
    }
}

namespace {
// Column blocks between a row-major n x ld array and contiguous columns
// (column c of the block at dst + c * rows), through a transposed LDS tile so
// that both sides are read and written in whole rows / runs: a workgroup
// takes 256 rows of at most kColBlock columns.
constexpr int kColBlock = 8;

__global__ __launch_bounds__(256) void k_gather_cols(const double *__restrict__ src, int64_t lds, int64_t rows,
                                                     int cols, double *__restrict__ dst)
{
    __shared__ double t[kColBlock][257];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    for (int e = threadIdx.x; e < 256 * cols; e += 256) {
        const int rr = e / cols, c = e % cols;
        if (r0 + rr < rows) t[c][rr] = src[(r0 + rr) * lds + c];
    }
    __syncthreads();
    const int64_t r = r0 + threadIdx.x;
    if (r < rows)
        for (int c = 0; c < cols; c++) dst[(int64_t)c * rows + r] = t[c][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_scatter_cols(const double *__restrict__ src, int64_t rows, int cols,
                                                      double *__restrict__ dst, int64_t ldd)
{
    __shared__ double t[kColBlock][257];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int64_t r = r0 + threadIdx.x;
    if (r < rows)
        for (int c = 0; c < cols; c++) t[c][threadIdx.x] = src[(int64_t)c * rows + r];
    __syncthreads();
    for (int e = threadIdx.x; e < 256 * cols; e += 256) {
        const int rr = e / cols, c = e % cols;
        if (r0 + rr < rows) dst[(r0 + rr) * ldd + c] = t[c][rr];
    }
}
}  // namespace

int col_block() { return kColBlock; }

void gather_cols(const double *src, int64_t lds, int64_t rows, int cols, double *dst, hipStream_t st)
{
    ABHSF_REQUIRE(cols >= 1 && cols <= kColBlock, ABHSF_ERR_INVALID, "column block too wide");
    if (rows <= 0) return;
    const int64_t nblk = (rows + 255) / 256;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "vector too long");
    hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)nblk), dim3(256), 0, st, src, lds, rows, cols, dst);
    ABHSF_HIP(hipGetLastError());
}

void scatter_cols(const double *src, int64_t rows, int cols, double *dst, int64_t ldd, hipStream_t st)
{
    ABHSF_REQUIRE(cols >= 1 && cols <= kColBlock, ABHSF_ERR_INVALID, "column block too wide");
    if (rows <= 0) return;
    const int64_t nblk = (rows + 255) / 256;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "vector too long");
    hipLaunchKernelGGL(k_scatter_cols, dim3((unsigned)nblk), dim3(256), 0, st, src, rows, cols, dst, ldd);
    ABHSF_HIP(hipGetLastError());
}

int max_rhs_tile(int s, bool comp)
{
    int R = 32;
    while (R > 1 && spmm_wave_lds(s, R, comp) > 65536) R >>= 1;
    return R;
}

void spmm(const SpmvLaunch &L, const DevMat &A, int nrhs, double alpha, const double *X, int64_t ldx, double beta,
          double *Y, int64_t ldy, hipStream_t st)
{
    const bool comp = L.val_bytes == 4;
    // column tiles of at most 8 right-hand sides (16 with dense blocks, whose
    // tiles go through the FP64 matrix cores from R = 16): wider tiles hold a
    // larger s x R accumulator in LDS per wave and lose more occupancy than
    // the extra pass over the matrix costs (stencil27, s = 64: k = 32 as one
    // tile 12.2 ms, as 2 x 16 4.3 ms each, as 4 x 8 1.5 ms each)
    // up to kSpmmSiblings tiles per launch go to sibling waves of one strip
    // (the matrix is read once for all of them)
    const int rmax = std::min(max_rhs_tile(A.s, comp), L.has_dense && !comp ? 16 : 8);
    const int sibmax = std::max(1, std::min(kSpmmSiblings, (int)(65536 / spmm_wave_lds(A.s, rmax, comp))));
    const int span = rmax * sibmax;
    for (int j0 = 0; j0 < nrhs; j0 += span) {
        const int nr = nrhs - j0 < span ? nrhs - j0 : span;
        int R = 1;
        while (R < nr && R < rmax) R <<= 1;
        if (comp) launch_v<float, true>(L, A, R, X + j0, ldx, Y + j0, ldy, nr, alpha, beta, st);
        else launch_v<double, false>(L, A, R, X + j0, ldx, Y + j0, ldy, nr, alpha, beta, st);
    }
}

}  // namespace abhsf
