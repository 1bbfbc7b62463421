// common.hpp -- shared definitions of the MI355X ABHSF library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/abhsf_c.h"

namespace abhsf {

constexpr int kWave = 64;

enum Scheme : int { COO = 0, CSR = 1, BITMAP = 2, DENSE = 3 };

// Error carried from the implementation to the C ABI boundary.
struct Error : std::runtime_error {
    abhsf_status code;
    Error(abhsf_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ABHSF_HIP(call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::abhsf::Error(ABHSF_ERR_HIP, std::string(#call " failed: ") +               \
                                                    hipGetErrorString(e_));                    \
    } while (0)

#define ABHSF_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                       \
        if (!(cond)) throw ::abhsf::Error((code), (msg));                                      \
    } while (0)

// ceil(log2 x) with log2(1) = 0 (include/abhsf/utils/math.h:17-31).
__host__ __device__ inline uint32_t ceil_log2_u64(uint64_t x)
{
    uint32_t e = 0;
    while (e < 64 && (x >> e) != 0) e++;
    if (x != 0 && (x & (x - 1)) == 0) e--;
    return e;
}

// Scheme choice for an s x s block with z entries.  PREPROC follows
// preprocessing/test.cpp:55-77, MSMF include/abhsf/msmf.h:9-48; ties keep the
// first scheme in the order COO, CSR, bitmap, dense.
__host__ __device__ inline int choose_scheme(int policy, uint64_t s, uint64_t z, uint64_t b,
                                             int is_binary, uint64_t *bits)
{
    const uint64_t ss = s * s;
    const uint64_t lg_s = ceil_log2_u64(s), lg_ss = ceil_log2_u64(ss);
    uint64_t coo, csr;
    if (policy == ABHSF_POLICY_PREPROC) {
        coo = 2 * lg_s * z + lg_ss;
        csr = lg_s * z + (s + 1) * lg_ss;
    } else {
        coo = lg_ss + z * (2 * lg_s);
        csr = s * lg_ss + z * lg_s;
    }
    int sc = COO;
    uint64_t mn = coo;
    if (csr < mn) { mn = csr; sc = CSR; }
    if (ss < mn) { mn = ss; sc = BITMAP; }
    if (!is_binary) {
        uint64_t dense = (ss - z) * b;
        if (dense < mn) { mn = dense; sc = DENSE; }
    }
    if (bits) *bits = mn;
    return sc;
}

// ---- wave-level helpers (64 lanes) ------------------------------------------

// Order LDS accesses of one wavefront (no cross-wave synchronisation).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int lane_id() { return __lane_id(); }

__device__ inline uint32_t wave_incl_scan_u32(uint32_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

template <int G>
__device__ inline double group_sum(double v)
{
    // Sum over aligned groups of G lanes; every lane of the group gets the sum.
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// ---- DPP cross-lane moves (gfx9 encodings; rows are 16 lanes) ---------------
enum : int {
    DPP_QUAD_XOR1 = 0xB1,   // quad_perm [1,0,3,2]
    DPP_QUAD_XOR2 = 0x4E,   // quad_perm [2,3,0,1]
    DPP_ROW_SHR1 = 0x111,
    DPP_ROW_SHR2 = 0x112,
    DPP_ROW_SHR4 = 0x114,
    DPP_ROW_SHR8 = 0x118,
    DPP_ROW_MIRROR = 0x140,       // lane i <- 15 - i within a row
    DPP_ROW_HALF_MIRROR = 0x141,  // lane i <- 7 - i within 8 lanes
    DPP_ROW_BCAST15 = 0x142,      // lane 15 of row r -> row r+1
    DPP_ROW_BCAST31 = 0x143       // lane 31 -> rows 2 and 3
};

// Lanes whose source is masked off or out of the row read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_f64(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>((uint32_t)b), hi = dpp_u32<CTRL, ROW_MASK>((uint32_t)(b >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Inclusive prefix sum over the 64 lanes (Hillis-Steele within rows, then
// row_bcast15 / row_bcast31 carry the row totals up the wave).
__device__ inline uint32_t wave_incl_scan_dpp(uint32_t v)
{
    v += dpp_u32<DPP_ROW_SHR1>(v);
    v += dpp_u32<DPP_ROW_SHR2>(v);
    v += dpp_u32<DPP_ROW_SHR4>(v);
    v += dpp_u32<DPP_ROW_SHR8>(v);
    v += dpp_u32<DPP_ROW_BCAST15, 0xa>(v);
    v += dpp_u32<DPP_ROW_BCAST31, 0xc>(v);
    return v;
}

// Sum over aligned groups of G consecutive lanes (G in 1..16); every lane of
// a group receives the group sum.
template <int G>
__device__ inline double group_sum_dpp(double v)
{
    if constexpr (G >= 2) v += dpp_f64<DPP_QUAD_XOR1>(v);
    if constexpr (G >= 4) v += dpp_f64<DPP_QUAD_XOR2>(v);
    if constexpr (G >= 8) v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
    if constexpr (G >= 16) v += dpp_f64<DPP_ROW_MIRROR>(v);
    return v;
}

__device__ inline uint64_t lanemask_lt() { return (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id())); }

// Error-free transformation a + b = s + e (Knuth TwoSum, TAOCP vol. 2 4.2.2).
__device__ inline void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Device view of one ABHSF matrix (all pointers into the handle's storage).
struct DevMat {
    int64_t m, n, M, N;
    int32_t s, k, W;  // W = u64 words per bitmap block
    int32_t zb;       // COO z1: z - 1 in the low zb bits, strip within the tile above (16: untiled)
    const uint32_t *sptr[4];
    const uint32_t *eptr[4];
    const void *bcol[4];
    const uint16_t *z1[4];
    const void *val[4];
    const void *coo_idx;
    const void *csr_rowptr;
    const uint8_t *csr_col;
    const uint64_t *bm_words;
};

}  // namespace abhsf
