// prims.hip -- reduce-then-scan exclusive scans and a stable LSD radix sort.
//
// Both are deterministic (integer only).  They serve the CSR -> ABHSF
// conversion, which replaces the reference's parallel sort + linear block scan
// (preprocessing/test.cpp:237-368).
#include "prims.hpp"

namespace abhsf {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;  // 2048 elements per workgroup

template <typename T>
__device__ inline T shfl_up_t(T v, int o) { return __shfl_up(v, o, kWave); }

// Workgroup exclusive scan of one value per thread; also returns the total.
template <typename T>
__device__ inline T block_excl_scan(T v, T *lds, T &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T t = shfl_up_t(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == kWave - 1) lds[wave] = incl;
    __syncthreads();
    T woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / kWave; w++) {
        T c = lds[w];
        if (w < wave) woff += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return woff + incl - v;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_tile_reduce(const T *__restrict__ in, uint64_t n,
                                                          T *__restrict__ sums)
{
    __shared__ T lds[kThreads / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        uint64_t g = base + (uint64_t)i * kThreads + threadIdx.x;
        if (g < n) acc += in[g];
    }
    T tot;
    (void)block_excl_scan(acc, lds, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of one tile; each thread owns kItems consecutive elements.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_tile_scan(const T *in, T *out, uint64_t n,
                                                        const T *__restrict__ tile_off,
                                                        T *__restrict__ total)
{
    __shared__ T lds[kThreads / kWave];
    __shared__ T stage[kTile];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    // coalesced load into LDS, then each thread scans its kItems run
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        uint64_t g = base + (uint64_t)i * kThreads + threadIdx.x;
        stage[i * kThreads + threadIdx.x] = (g < n) ? in[g] : T(0);
    }
    __syncthreads();
    T v[kItems];
    T acc = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        v[i] = stage[threadIdx.x * kItems + i];
        acc += v[i];
    }
    T tot;
    T off = block_excl_scan(acc, lds, tot);
    if (tile_off) off += tile_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        stage[threadIdx.x * kItems + i] = off;
        off += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        uint64_t g = base + (uint64_t)i * kThreads + threadIdx.x;
        if (g < n) out[g] = stage[i * kThreads + threadIdx.x];
    }
    if (total && threadIdx.x == 0 && gridDim.x == 1) *total = tot;
}

template <typename T>
void exclusive_scan_impl(const T *in, T *out, uint64_t n, T *total, hipStream_t st, std::vector<DevBuf> &keep)
{
    if (n == 0) {
        if (total) ABHSF_HIP(hipMemsetAsync(total, 0, sizeof(T), st));
        return;
    }
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    if (ntiles == 1) {
        k_tile_scan<T><<<1, kThreads, 0, st>>>(in, out, n, nullptr, total);
        ABHSF_HIP(hipGetLastError());
        return;
    }
    keep.emplace_back(ntiles * sizeof(T));
    T *sums = keep.back().get<T>();
    k_tile_reduce<T><<<(unsigned)ntiles, kThreads, 0, st>>>(in, n, sums);
    ABHSF_HIP(hipGetLastError());
    exclusive_scan_impl<T>(sums, sums, ntiles, total, st, keep);
    k_tile_scan<T><<<(unsigned)ntiles, kThreads, 0, st>>>(in, out, n, sums, nullptr);
    ABHSF_HIP(hipGetLastError());
}

// ---- radix sort --------------------------------------------------------------

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint64_t *__restrict__ keys, uint64_t n,
                                                         int shift, uint32_t *__restrict__ hist,
                                                         uint32_t ntiles)
{
    __shared__ uint32_t cnt[kRadix];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        uint64_t g = base + (uint64_t)i * kThreads + threadIdx.x;
        if (g < n) atomicAdd(&cnt[(keys[g] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(const uint64_t *__restrict__ kin,
                                                            const uint32_t *__restrict__ vin,
                                                            uint64_t *__restrict__ kout,
                                                            uint32_t *__restrict__ vout, uint64_t n,
                                                            int shift, const uint32_t *__restrict__ offs,
                                                            uint32_t ntiles)
{
    static_assert(kThreads == kRadix, "one thread per digit");
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wc[kThreads / kWave][kRadix];
    const int wave = threadIdx.x >> 6;
    run[threadIdx.x] = offs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
#pragma unroll
        for (int w = 0; w < kThreads / kWave; w++) wc[w][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t g = base + (uint64_t)r * kThreads + threadIdx.x;
        const bool valid = g < n;
        const uint64_t key = valid ? kin[g] : 0;
        const uint32_t val = valid ? vin[g] : 0;
        const uint32_t d = (uint32_t)(key >> shift) & (kRadix - 1);
        // lanes of this wave holding the same digit
        uint64_t mask = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            mask &= bit ? bal : ~bal;
        }
        const uint32_t rank = __popcll(mask & lanemask_lt());
        if (valid && rank == 0) wc[wave][d] = __popcll(mask);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int w = 0; w < wave; w++) pos += wc[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kThreads / kWave; w++) add += wc[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

}  // namespace

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total, hipStream_t st)
{
    std::vector<DevBuf> keep;
    exclusive_scan_impl<uint32_t>(in, out, n, total, st, keep);
    if (!keep.empty()) ABHSF_HIP(hipStreamSynchronize(st));
}

void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total, hipStream_t st)
{
    std::vector<DevBuf> keep;
    exclusive_scan_impl<uint64_t>(in, out, n, total, st, keep);
    if (!keep.empty()) ABHSF_HIP(hipStreamSynchronize(st));
}

void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, hipStream_t st)
{
    if (n < 2 || key_bits <= 0) return;
    ABHSF_REQUIRE(n < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "radix sort limited to 2^32 elements");
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    ABHSF_REQUIRE(ntiles * kRadix < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "radix histogram too large");
    DevBuf k2(n * sizeof(uint64_t)), v2(n * sizeof(uint32_t)), hist(ntiles * kRadix * sizeof(uint32_t));
    uint64_t *ka = keys, *kb = k2.get<uint64_t>();
    uint32_t *va = vals, *vb = v2.get<uint32_t>();
    std::vector<DevBuf> keep;
    for (int shift = 0; shift < key_bits; shift += kRadixBits) {
        k_radix_hist<<<(unsigned)ntiles, kThreads, 0, st>>>(ka, n, shift, hist.get<uint32_t>(), (uint32_t)ntiles);
        ABHSF_HIP(hipGetLastError());
        exclusive_scan_impl<uint32_t>(hist.get<uint32_t>(), hist.get<uint32_t>(), ntiles * kRadix, nullptr, st, keep);
        k_radix_scatter<<<(unsigned)ntiles, kThreads, 0, st>>>(ka, va, kb, vb, n, shift, hist.get<uint32_t>(),
                                                                (uint32_t)ntiles);
        ABHSF_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        ABHSF_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        ABHSF_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
    ABHSF_HIP(hipStreamSynchronize(st));
}

}  // namespace abhsf
