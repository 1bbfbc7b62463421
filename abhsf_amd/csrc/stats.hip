// stats.hip -- block statistics on the GPU: the histogram "entries per
// nonzero block -> number of blocks" for r x s blocks (r = 2^kr, s = 2^ks).
//
// Restates block_statistics/block_stats.cpp:69-100 (sort by block key, count
// run lengths) as data-parallel passes: per-row sample offsets (scan), block
// keys, stable LSD radix sort, head flags, run lengths, histogram by integer
// atomics (order-independent, so deterministic).  With strip_stride > 1 only
// every strip_stride-th strip of r rows is counted (a sample for the tuner's
// features); strip_stride = 1 is the reference's full statistic.
#include "prims.hpp"

namespace abhsf {
namespace {

constexpr int kT = 256;

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kT - 1) / kT); }

int bits_for(uint64_t maxval)
{
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) b++;
    return b;
}

__global__ void k_sample_counts(const int64_t *__restrict__ rp, int64_t m, int kr, int64_t stride,
                                uint32_t *__restrict__ cnt)
{
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r >= m) return;
    const bool take = ((r >> kr) % stride) == 0;
    cnt[r] = take ? (uint32_t)(rp[r + 1] - rp[r]) : 0u;
}

__global__ void k_sample_keys(const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int64_t m,
                              uint64_t nnz, int kr, int ks, int jbits, int64_t stride,
                              const uint32_t *__restrict__ off, uint64_t *__restrict__ key,
                              uint32_t *__restrict__ pay)
{
    const uint64_t e = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= nnz) return;
    int64_t lo = 0, hi = m - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((uint64_t)rp[mid] <= e) lo = mid; else hi = mid - 1;
    }
    if (((lo >> kr) % stride) != 0) return;
    const uint64_t o = off[lo] + (e - (uint64_t)rp[lo]);
    key[o] = ((uint64_t)(lo >> kr) << jbits) | (uint64_t)((uint32_t)col[e] >> ks);
    pay[o] = 0;
}

// Run heads measure their run (the block's entry count z) and the wave
// aggregates: one atomic per distinct bin per wave (most blocks of sparse
// matrices share a handful of small z, so per-lane atomics would contend).
__global__ void k_runs_hist(const uint64_t *__restrict__ key, uint64_t n, unsigned long long *__restrict__ hist,
                            uint64_t cap, unsigned long long *__restrict__ nblocks)
{
    const uint64_t e = (uint64_t)blockIdx.x * kT + threadIdx.x;
    const int lane = (int)(threadIdx.x & 63);
    uint64_t z = 0;  // 0: not a run head
    if (e < n && (e == 0 || key[e] != key[e - 1])) {
        // run end by galloping then bisection (keys are sorted): O(log z)
        const uint64_t k0 = key[e];
        uint64_t lo = 1, hi = 1;  // key[e + lo - 1] == k0; answer z in [lo, hi]
        while (e + hi < n && key[e + hi] == k0) {
            lo = hi + 1;
            hi = 2 * hi;
        }
        if (e + hi > n) hi = n - e;
        while (lo < hi) {  // smallest z in [lo, hi] with e + z == n or key[e + z] != k0
            const uint64_t mid = (lo + hi) >> 1;
            if (key[e + mid] == k0) lo = mid + 1; else hi = mid;
        }
        z = lo;
        if (z >= cap) z = cap - 1;
    }
    const uint64_t heads = __ballot(z != 0);
    if (lane == 0 && heads) atomicAdd(nblocks, (unsigned long long)__popcll(heads));
    bool pending = z != 0;
    for (uint64_t live = heads; live; live = __ballot(pending)) {
        const int leader = __ffsll((unsigned long long)live) - 1;
        const uint64_t zl = __shfl(z, leader, 64);
        const uint64_t same = __ballot(pending && z == zl);
        if (lane == leader) atomicAdd(&hist[zl], (unsigned long long)__popcll(same));
        if (z == zl) pending = false;
    }
}

}  // namespace
}  // namespace abhsf

using namespace abhsf;

extern "C" abhsf_status abhsf_block_hist_impl(const int64_t *rowptr, const int32_t *col, int64_t m, int64_t n,
                                              int64_t nnz, int32_t on_device, int32_t kr, int32_t ks,
                                              int64_t strip_stride, uint64_t *hist, int64_t hist_cap,
                                              uint64_t *nblocks, uint64_t *sampled_nnz, void *stream,
                                              std::string *err)
{
    try {
        ABHSF_REQUIRE(rowptr && hist && hist_cap >= 2 && nblocks && m >= 1 && n >= 1 && nnz >= 0,
                      ABHSF_ERR_INVALID, "invalid argument");
        ABHSF_REQUIRE(kr >= 0 && kr <= 16 && ks >= 0 && ks <= 16 && strip_stride >= 1, ABHSF_ERR_INVALID,
                      "block size or stride out of range");
        ABHSF_REQUIRE((uint64_t)nnz < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "nnz must be below 2^32");
        hipStream_t st = static_cast<hipStream_t>(stream);
        DevBuf drp, dc;
        const int64_t *rp = rowptr;
        const int32_t *cl = col;
        if (!on_device) {
            drp.alloc(((uint64_t)m + 1) * 8);
            dc.alloc((uint64_t)nnz * 4);
            ABHSF_HIP(hipMemcpyAsync(drp.get(), rowptr, drp.bytes(), hipMemcpyHostToDevice, st));
            if (nnz) ABHSF_HIP(hipMemcpyAsync(dc.get(), col, dc.bytes(), hipMemcpyHostToDevice, st));
            rp = drp.get<int64_t>();
            cl = dc.get<int32_t>();
        }
        DevBuf cnt((uint64_t)m * 4), tot(4);
        k_sample_counts<<<grid_for((uint64_t)m), kT, 0, st>>>(rp, m, kr, strip_stride, cnt.get<uint32_t>());
        ABHSF_HIP(hipGetLastError());
        exclusive_scan_u32(cnt.get<uint32_t>(), cnt.get<uint32_t>(), (uint64_t)m, tot.get<uint32_t>(), st);
        uint32_t ns = 0;
        ABHSF_HIP(hipMemcpyAsync(&ns, tot.get(), 4, hipMemcpyDeviceToHost, st));
        ABHSF_HIP(hipStreamSynchronize(st));
        if (sampled_nnz) *sampled_nnz = ns;
        DevBuf key((uint64_t)ns * 8), pay((uint64_t)ns * 4), dh((uint64_t)hist_cap * 8), dnb(8);
        ABHSF_HIP(hipMemsetAsync(dh.get(), 0, dh.bytes(), st));
        ABHSF_HIP(hipMemsetAsync(dnb.get(), 0, 8, st));
        const int jbits = bits_for((uint64_t)((n - 1) >> ks));
        const int ibits = bits_for((uint64_t)((m - 1) >> kr));
        if (ns) {
            k_sample_keys<<<grid_for((uint64_t)nnz), kT, 0, st>>>(rp, cl, m, (uint64_t)nnz, kr, ks, jbits,
                                                                  strip_stride, cnt.get<uint32_t>(),
                                                                  key.get<uint64_t>(), pay.get<uint32_t>());
            ABHSF_HIP(hipGetLastError());
            radix_sort_pairs(key.get<uint64_t>(), pay.get<uint32_t>(), ns, jbits + ibits, st);
            k_runs_hist<<<grid_for(ns), kT, 0, st>>>(key.get<uint64_t>(), ns, dh.get<unsigned long long>(),
                                                      (uint64_t)hist_cap, dnb.get<unsigned long long>());
            ABHSF_HIP(hipGetLastError());
        }
        ABHSF_HIP(hipMemcpyAsync(hist, dh.get(), dh.bytes(), hipMemcpyDeviceToHost, st));
        ABHSF_HIP(hipMemcpyAsync(nblocks, dnb.get(), 8, hipMemcpyDeviceToHost, st));
        ABHSF_HIP(hipStreamSynchronize(st));
        return ABHSF_OK;
    } catch (const Error &e) {
        *err = e.what();
        return e.code;
    }
}
