// delta.hip -- delta-coded coordinate storage and its SpMV (the delta codecs
// of spmv/unified/spmv.cpp:628-1033, re-laid out for wavefront decode).
//
// The reference decodes its varint stream with a serial running sum, one
// thread per contiguous range (:923-944): the width of element k is only
// known after element k-1 has been read.  Here a wavefront decodes 64
// elements per step: each lane reads its element's 2-bit tag from the tag
// words, ballots per width give the element's index in its width stream
// (so all loads are issued at once), and an inclusive wave scan of the
// deltas plus the chunk's base key gives every key.  Rows are
// non-decreasing along the elements, so products are summed per row with a
// segmented scan and one lane per row segment adds the sum to y atomically
// (a row may continue into the next step or chunk).  Both scans are DPP row
// shifts plus the row broadcasts (no LDS traffic).
#include "delta.hpp"

#include <algorithm>
#include <chrono>
#include <vector>

namespace abhsf {
namespace {

constexpr int kWaveD = 64;
constexpr int kStepsD = kDeltaChunk / kWaveD;

__device__ inline uint64_t shfl_up_u64(uint64_t v, int d)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d, kWaveD);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d, kWaveD);
    return ((uint64_t)hi << 32) | lo;
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ inline uint64_t dpp_u64(uint64_t v)
{
    return ((uint64_t)dpp_u32<CTRL, ROW_MASK>((uint32_t)(v >> 32)) << 32) | dpp_u32<CTRL, ROW_MASK>((uint32_t)v);
}

// Inclusive 64-bit prefix sum over the wave (DPP row shifts, then the row
// broadcasts carry row totals up), as wave_incl_scan_dpp for 32 bits.
__device__ inline uint64_t wave_incl_scan_u64(uint64_t v)
{
    v += dpp_u64<DPP_ROW_SHR1>(v);
    v += dpp_u64<DPP_ROW_SHR2>(v);
    v += dpp_u64<DPP_ROW_SHR4>(v);
    v += dpp_u64<DPP_ROW_SHR8>(v);
    v += dpp_u64<DPP_ROW_BCAST15, 0xa>(v);
    v += dpp_u64<DPP_ROW_BCAST31, 0xc>(v);
    return v;
}

// One step of a segmented inclusive sum: lanes not yet closed by a segment
// head add the shifted partial sum, and inherit its head flag.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline void seg_step(double &v, uint32_t &f)
{
    const double vd = dpp_f64<CTRL, ROW_MASK>(v);
    const uint32_t fd = dpp_u32<CTRL, ROW_MASK>(f);
    if (!f) v += vd;
    f |= fd;
}

__global__ __launch_bounds__(256) void k_delta_spmv(const uint32_t *__restrict__ tags,
                                                    const uint8_t *__restrict__ w8,
                                                    const uint16_t *__restrict__ w16,
                                                    const uint32_t *__restrict__ w32,
                                                    const uint64_t *__restrict__ w64,
                                                    const uint64_t *__restrict__ base,
                                                    const uint32_t *__restrict__ start,
                                                    const double *__restrict__ val, int64_t nnz, int K,
                                                    int64_t nchunk, const double *__restrict__ x, double *y,
                                                    double alpha)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * 4 + wave;
    if (c >= nchunk) return;
    uint64_t carry = base[c];
    uint32_t o0 = start[c * 4 + 0], o1 = start[c * 4 + 1], o2 = start[c * 4 + 2], o3 = start[c * 4 + 3];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t kmask = (1ull << K) - 1ull;
#pragma unroll 1
    for (int step = 0; step < kStepsD; step++) {
        const int64_t e0 = c * kDeltaChunk + (int64_t)step * kWaveD;
        if (e0 >= nnz) break;  // wave-uniform
        const int64_t e = e0 + lane;
        const bool valid = e < nnz;
        const uint32_t tag = valid ? (tags[e >> 4] >> (2 * (e & 15))) & 3u : 4u;
        const uint64_t m0 = __ballot(tag == 0), m1 = __ballot(tag == 1), m2 = __ballot(tag == 2),
                       m3 = __ballot(tag == 3);
        uint64_t word = 0;
        switch (tag) {
        case 0: word = w8[o0 + (uint32_t)__popcll(m0 & lt)]; break;
        case 1: word = w16[o1 + (uint32_t)__popcll(m1 & lt)]; break;
        case 2: word = w32[o2 + (uint32_t)__popcll(m2 & lt)]; break;
        case 3: word = w64[o3 + (uint32_t)__popcll(m3 & lt)]; break;
        default: break;
        }
        o0 += (uint32_t)__popcll(m0);
        o1 += (uint32_t)__popcll(m1);
        o2 += (uint32_t)__popcll(m2);
        o3 += (uint32_t)__popcll(m3);
        // keys: inclusive scan of the deltas (invalid lanes add 0)
        uint64_t key = wave_incl_scan_u64(word >> 2) + carry;
        carry = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(key >> 32), kWaveD - 1) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, kWaveD - 1);
        const uint64_t row = valid ? key >> K : ~0ull;
        double p = valid ? alpha * val[e] * x[key & kmask] : 0.0;
        // segmented inclusive sum over runs of equal rows (head = first lane
        // of a run); the last lane of each run adds the run's sum to y
        const uint64_t prev = shfl_up_u64(row, 1);
        const uint32_t head = (lane == 0 || prev != row) ? 1u : 0u;
        const uint32_t next_head = (uint32_t)__shfl_down((int)head, 1, kWaveD);
        uint32_t f = head;
        seg_step<DPP_ROW_SHR1>(p, f);
        seg_step<DPP_ROW_SHR2>(p, f);
        seg_step<DPP_ROW_SHR4>(p, f);
        seg_step<DPP_ROW_SHR8>(p, f);
        seg_step<DPP_ROW_BCAST15, 0xa>(p, f);
        seg_step<DPP_ROW_BCAST31, 0xc>(p, f);
        if (valid && (lane == kWaveD - 1 || next_head)) atomicAdd(&y[row], p);
    }
}

// Inclusive max over the wave of values >= 0 (out-of-row DPP sources read 0).
__device__ inline uint32_t wave_incl_max_u32(uint32_t v)
{
    v = max(v, dpp_u32<DPP_ROW_SHR1>(v));
    v = max(v, dpp_u32<DPP_ROW_SHR2>(v));
    v = max(v, dpp_u32<DPP_ROW_SHR4>(v));
    v = max(v, dpp_u32<DPP_ROW_SHR8>(v));
    v = max(v, dpp_u32<DPP_ROW_BCAST15, 0xa>(v));
    v = max(v, dpp_u32<DPP_ROW_BCAST31, 0xc>(v));
    return v;
}

// u32 delta codecs: a wavefront decodes 64 elements per step.  Deltas are
// one coalesced u32 load; escapes (0 at any element but a chunk's first)
// take their key from the exception list at the chunk's offset plus the
// escapes before them (ballot + popcount).  Keys: the inclusive 64-bit scan
// S of the deltas (escapes add 0) plus, per lane, the base of its segment --
// the chunk carry before the first escape, (exception key - S at the escape)
// after it, found by a max-scan of escape lane indices and one shuffle.
// Row-major keys: products are summed over runs of equal rows (segmented DPP
// scan) and the last lane of a run adds its sum to y atomically (one run per
// row and step).  8 x 8 cache-block keys alternate rows inside each cache
// block, so runs are one element long; their products go to a per-wave LDS
// bin of kBandsD 8-row bands (rows further on take a global add), a window
// that moves up to each step's first band: the bands it leaves are complete
// and their touched rows are added to y then (R-MAT scale 22: one global add
// per element took 2.84 ms, a flush of the bin after every step 0.52 ms).
constexpr int kBandsD = 8;
static_assert(kBandsD * 8 <= 64 && (kBandsD & (kBandsD - 1)) == 0, "one bin slot per lane, power of two bands");

template <int CODEC>
__global__ __launch_bounds__(256) void k_delta32_spmv(const uint32_t *__restrict__ da,
                                                      const uint64_t *__restrict__ base,
                                                      const uint32_t *__restrict__ xstart,
                                                      const uint64_t *__restrict__ exc,
                                                      const double *__restrict__ val, int64_t nnz, int64_t nchunk,
                                                      const double *__restrict__ x, double *y, double alpha)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    __shared__ double s_bin[4][kBandsD * 8];
    double *bin = s_bin[wave];
    if constexpr (CODEC == kDeltaU32Block) {
        if (lane < kBandsD * 8) bin[lane] = 0.0;
        wave_sync();
    }
    const int64_t c = (int64_t)blockIdx.x * 4 + wave;
    if (c >= nchunk) return;
    uint64_t carry = base[c];
    uint32_t xo = xstart[c];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // cache-block keys: bin slot (band % kBandsD) * 8 + row % 8 holds the rows
    // of one band of the window [wb, wb + kBandsD); flush_bins(lim) adds the
    // bins of the window's bands below lim to y and clears them
    uint64_t wb = 0;
    bool have_wb = false;
    auto flush_bins = [&](uint64_t lim) {
        if (lane < kBandsD * 8) {
            const uint64_t sb = (uint64_t)(lane >> 3);
            const uint64_t b = wb + ((sb - (wb & (kBandsD - 1))) & (kBandsD - 1));
            const double v = bin[lane];
            if (b < lim && v != 0.0) {
                atomicAdd(&y[b * 8 + (uint64_t)(lane & 7)], v);
                bin[lane] = 0.0;
            }
        }
        wave_sync();
    };
#pragma unroll 1
    for (int step = 0; step < kStepsD; step++) {
        const int64_t e0 = c * kDeltaChunk + (int64_t)step * kWaveD;
        if (e0 >= nnz) break;  // wave-uniform
        const int64_t e = e0 + lane;
        const bool valid = e < nnz;
        const uint32_t d = valid ? da[e] : 0u;
        const bool esc = valid && d == 0u && !(step == 0 && lane == 0);
        const uint64_t mesc = __ballot(esc);
        uint64_t ekey = 0;
        if (esc) ekey = exc[xo + (uint32_t)__popcll(mesc & lt)];
        xo += (uint32_t)__popcll(mesc);
        const uint64_t S = wave_incl_scan_u64(esc ? 0ull : (uint64_t)d);
        uint64_t key = carry + S;
        if (mesc) {  // wave-uniform: some lane escapes in this step
            const uint32_t last = wave_incl_max_u32(esc ? (uint32_t)lane + 1u : 0u);
            const uint64_t b = ekey - S;  // meaningful at escape lanes
            const uint32_t src = last ? last - 1u : 0u;
            const uint32_t blo = (uint32_t)__shfl((int)(uint32_t)b, (int)src, kWaveD);
            const uint32_t bhi = (uint32_t)__shfl((int)(uint32_t)(b >> 32), (int)src, kWaveD);
            if (last) key = (((uint64_t)bhi << 32) | blo) + S;
        }
        carry = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(key >> 32), kWaveD - 1) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, kWaveD - 1);
        uint64_t row, col;
        if constexpr (CODEC == kDeltaU32Row) {
            row = key >> 31;
            col = key & ((1ull << 31) - 1ull);
        } else {
            row = ((key >> 31) << 3) | ((key >> 3) & 7ull);
            col = (((key >> 6) & ((1ull << 25) - 1ull)) << 3) | (key & 7ull);
        }
        if (!valid) row = ~0ull;
        double p = valid ? alpha * val[e] * x[col] : 0.0;
        if constexpr (CODEC == kDeltaU32Block) {
            // lane 0 holds the step's first (smallest) row: e0 < nnz.  Bands
            // never decrease, so the bands before it are complete: their bins
            // go to y and the window moves up to it
            const uint64_t band0 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(row >> 32), 0)
                                    << 32 | (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row, 0)) >> 3;
            if (!have_wb) {
                wb = band0;
                have_wb = true;
            } else if (band0 > wb) {
                flush_bins(band0);
                wb = band0;
            }
            if (valid) {
                if ((row >> 3) - wb < (uint64_t)kBandsD) atomicAdd(&bin[((row >> 3) & (kBandsD - 1)) * 8 + (row & 7ull)], p);
                else atomicAdd(&y[row], p);
            }
            wave_sync();
            continue;
        }
        const uint64_t prev = shfl_up_u64(row, 1);
        const uint32_t head = (lane == 0 || prev != row) ? 1u : 0u;
        const uint32_t next_head = (uint32_t)__shfl_down((int)head, 1, kWaveD);
        uint32_t f = head;
        seg_step<DPP_ROW_SHR1>(p, f);
        seg_step<DPP_ROW_SHR2>(p, f);
        seg_step<DPP_ROW_SHR4>(p, f);
        seg_step<DPP_ROW_SHR8>(p, f);
        seg_step<DPP_ROW_BCAST15, 0xa>(p, f);
        seg_step<DPP_ROW_BCAST31, 0xc>(p, f);
        if (valid && (lane == kWaveD - 1 || next_head)) atomicAdd(&y[row], p);
    }
    if constexpr (CODEC == kDeltaU32Block) {
        if (have_wb) flush_bins(~0ull);  // the window's last bands
    }
}

__global__ void k_delta_scale(double *y, int64_t m, double beta)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < m) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
}

}  // namespace

uint64_t DeltaStorage::stored_bytes() const
{
    uint64_t b = tags.bytes() + base.bytes() + start.bytes() + val.bytes() + da.bytes() + xstart.bytes() +
                 exc.bytes();
    for (int t = 0; t < 4; t++) b += words[t].bytes();
    return b;
}

namespace {

// Host encoder of the u32 codecs (layout documented in delta.hpp); the CSR
// has been validated.
void delta32_encode(DeltaStorage &D, const int64_t *rowptr, const int32_t *col, const double *val, hipStream_t st)
{
    const int64_t nnz = D.nnz, m = D.m;
    std::vector<uint64_t> key((size_t)nnz);
    std::vector<uint32_t> idx((size_t)nnz);
    for (int64_t r = 0; r < m; r++)
        for (int64_t q = rowptr[r]; q < rowptr[r + 1]; q++) {
            const uint64_t rr = (uint64_t)r, cc = (uint64_t)col[q];
            key[(size_t)q] = D.codec == kDeltaU32Row ? (rr << 31) + cc
                                                     : ((rr >> 3) << 31) + ((cc >> 3) << 6) + ((rr & 7) << 3) + (cc & 7);
            idx[(size_t)q] = (uint32_t)q;
        }
    if (D.codec == kDeltaU32Block) {
        // (row/8, col/8, row%8, col%8) order = ascending key; elements only
        // move within their band of 8 rows, whose range is the same in CSR
        // order, so each band is sorted on its own
        for (int64_t r0 = 0; r0 < m; r0 += 8) {
            const int64_t q0 = rowptr[r0], q1 = rowptr[std::min<int64_t>(r0 + 8, m)];
            std::sort(idx.begin() + q0, idx.begin() + q1, [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
        }
    }
    std::vector<uint32_t> da((size_t)nnz), xstart((size_t)D.nchunk + 1);
    std::vector<uint64_t> base((size_t)D.nchunk), exc;
    std::vector<double> v((size_t)nnz);
    uint64_t prev = 0;
    for (int64_t e = 0; e < nnz; e++) {
        const uint32_t q = idx[(size_t)e];
        const uint64_t k = key[q], d = k - prev;
        const bool fits = d < (1ull << 32);
        if (e % kDeltaChunk == 0) {
            base[(size_t)(e / kDeltaChunk)] = fits ? prev : k;
            da[(size_t)e] = fits ? (uint32_t)d : 0u;
            xstart[(size_t)(e / kDeltaChunk)] = (uint32_t)exc.size();
        } else if (fits) {
            da[(size_t)e] = (uint32_t)d;
        } else {
            da[(size_t)e] = 0u;
            exc.push_back(k);
        }
        v[(size_t)e] = val ? val[q] : 1.0;
        prev = k;
    }
    xstart[(size_t)D.nchunk] = (uint32_t)exc.size();
    D.count[2] = (uint64_t)nnz - exc.size();
    D.count[3] = exc.size();
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
        b.alloc(bytes);
        if (bytes) ABHSF_HIP(hipMemcpyAsync(b.get(), src, bytes, hipMemcpyHostToDevice, st));
    };
    up(D.da, da.data(), da.size() * 4);
    up(D.base, base.data(), base.size() * 8);
    up(D.xstart, xstart.data(), xstart.size() * 4);
    up(D.exc, exc.data(), exc.size() * 8);
    up(D.val, v.data(), v.size() * 8);
    ABHSF_HIP(hipStreamSynchronize(st));
}

}  // namespace

void delta_build(DeltaStorage &D, int64_t m, int64_t n, const int64_t *rowptr, const int32_t *col,
                 const double *val, hipStream_t st, int codec)
{
    const auto t0 = std::chrono::steady_clock::now();
    D = DeltaStorage();
    ABHSF_REQUIRE(codec == kDeltaVarint || codec == kDeltaU32Row || codec == kDeltaU32Block, ABHSF_ERR_INVALID,
                  "unknown delta codec");
    ABHSF_REQUIRE(codec != kDeltaU32Block || n <= (1ll << 28), ABHSF_ERR_UNSUPPORTED,
                  "the 8 x 8 cache-block key holds column blocks in 25 bits: n must be at most 2^28");
    D.codec = codec;
    ABHSF_REQUIRE(m >= 1 && n >= 1 && m <= INT32_MAX && n <= INT32_MAX, ABHSF_ERR_INVALID,
                  "matrix dimensions must be in [1, 2^31-1]");
    ABHSF_REQUIRE(rowptr && rowptr[0] == 0, ABHSF_ERR_INVALID, "row pointers must start at 0");
    const int64_t nnz = rowptr[m];
    ABHSF_REQUIRE(nnz >= 0 && (uint64_t)nnz < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "nnz must be in [0, 2^32)");
    ABHSF_REQUIRE(nnz == 0 || col, ABHSF_ERR_INVALID, "null column array");
    D.m = m;
    D.n = n;
    D.nnz = nnz;
    D.K = std::max<int>(1, (int)ceil_log2_u64((uint64_t)n));
    D.nchunk = (nnz + kDeltaChunk - 1) / kDeltaChunk;
    std::vector<uint32_t> tags((size_t)(nnz + 15) / 16, 0u);
    std::vector<uint8_t> w8;
    std::vector<uint16_t> w16;
    std::vector<uint32_t> w32;
    std::vector<uint64_t> w64;
    std::vector<uint64_t> base((size_t)D.nchunk);
    std::vector<uint32_t> start((size_t)D.nchunk * 4);
    uint64_t prev = 0;
    int64_t e = 0;
    for (int64_t r = 0; r < m; r++) {
        ABHSF_REQUIRE(rowptr[r + 1] >= rowptr[r] && rowptr[r + 1] <= nnz, ABHSF_ERR_INVALID,
                      "row pointers not monotone");
        for (int64_t q = rowptr[r]; q < rowptr[r + 1]; q++, e++) {
            const int32_t cc = col[q];
            ABHSF_REQUIRE(cc >= 0 && (int64_t)cc < n, ABHSF_ERR_INVALID, "column out of range");
            ABHSF_REQUIRE(q == rowptr[r] || cc > col[q - 1], ABHSF_ERR_INVALID,
                          "columns not strictly increasing within a row");
            if (codec != kDeltaVarint) continue;  // validated here, encoded by delta32_encode
            const uint64_t key = ((uint64_t)r << D.K) + (uint64_t)cc;
            if (e % kDeltaChunk == 0) {
                const int64_t c = e / kDeltaChunk;
                base[(size_t)c] = prev;
                start[(size_t)c * 4 + 0] = (uint32_t)w8.size();
                start[(size_t)c * 4 + 1] = (uint32_t)w16.size();
                start[(size_t)c * 4 + 2] = (uint32_t)w32.size();
                start[(size_t)c * 4 + 3] = (uint32_t)w64.size();
            }
            const uint64_t delta = key - prev;
            uint32_t tag;
            if (delta < (1ull << 6)) { tag = 0; w8.push_back((uint8_t)(delta << 2)); }
            else if (delta < (1ull << 14)) { tag = 1; w16.push_back((uint16_t)((delta << 2) | 1u)); }
            else if (delta < (1ull << 30)) { tag = 2; w32.push_back((uint32_t)((delta << 2) | 2u)); }
            else { tag = 3; w64.push_back((delta << 2) | 3u); }
            tags[(size_t)(e >> 4)] |= tag << (2 * (e & 15));
            prev = key;
        }
    }
    if (codec != kDeltaVarint) {
        delta32_encode(D, rowptr, col, val, st);
        D.built = true;
        D.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return;
    }
    D.count[0] = w8.size();
    D.count[1] = w16.size();
    D.count[2] = w32.size();
    D.count[3] = w64.size();
    auto up = [&](DevBuf &d, const void *src, size_t bytes) {
        d.alloc(bytes);
        if (bytes) ABHSF_HIP(hipMemcpyAsync(d.get(), src, bytes, hipMemcpyHostToDevice, st));
    };
    up(D.tags, tags.data(), tags.size() * 4);
    up(D.words[0], w8.data(), w8.size());
    up(D.words[1], w16.data(), w16.size() * 2);
    up(D.words[2], w32.data(), w32.size() * 4);
    up(D.words[3], w64.data(), w64.size() * 8);
    up(D.base, base.data(), base.size() * 8);
    up(D.start, start.data(), start.size() * 4);
    D.val.alloc((size_t)nnz * 8);
    if (nnz) {
        if (val) {
            ABHSF_HIP(hipMemcpyAsync(D.val.get(), val, (size_t)nnz * 8, hipMemcpyHostToDevice, st));
        } else {
            std::vector<double> ones((size_t)nnz, 1.0);
            ABHSF_HIP(hipMemcpyAsync(D.val.get(), ones.data(), (size_t)nnz * 8, hipMemcpyHostToDevice, st));
            ABHSF_HIP(hipStreamSynchronize(st));
        }
    }
    ABHSF_HIP(hipStreamSynchronize(st));
    D.built = true;
    D.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void delta_spmv(const DeltaStorage &D, double alpha, const double *x, double beta, double *y, hipStream_t st)
{
    ABHSF_REQUIRE(D.built, ABHSF_ERR_STATE, "matrix not built");
    if (beta != 1.0) {
        hipLaunchKernelGGL(k_delta_scale, dim3((unsigned)((D.m + 255) / 256)), dim3(256), 0, st, y, D.m, beta);
        ABHSF_HIP(hipGetLastError());
    }
    if (D.nchunk == 0 || alpha == 0.0) return;
    const int64_t nblk = (D.nchunk + 3) / 4;
    ABHSF_REQUIRE(nblk < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "too many chunks");
    if (D.codec != kDeltaVarint) {
        const uint32_t *da = D.da.get<uint32_t>(), *xs = D.xstart.get<uint32_t>();
        const uint64_t *bs = D.base.get<uint64_t>(), *ex = D.exc.get<uint64_t>();
        if (D.codec == kDeltaU32Row)
            hipLaunchKernelGGL(k_delta32_spmv<kDeltaU32Row>, dim3((unsigned)nblk), dim3(256), 0, st, da, bs, xs, ex,
                               D.val.get<double>(), D.nnz, D.nchunk, x, y, alpha);
        else
            hipLaunchKernelGGL(k_delta32_spmv<kDeltaU32Block>, dim3((unsigned)nblk), dim3(256), 0, st, da, bs, xs,
                               ex, D.val.get<double>(), D.nnz, D.nchunk, x, y, alpha);
        ABHSF_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_delta_spmv, dim3((unsigned)nblk), dim3(256), 0, st, D.tags.get<uint32_t>(),
                       D.words[0].get<uint8_t>(), D.words[1].get<uint16_t>(), D.words[2].get<uint32_t>(),
                       D.words[3].get<uint64_t>(), D.base.get<uint64_t>(), D.start.get<uint32_t>(),
                       D.val.get<double>(), D.nnz, D.K, D.nchunk, x, y, alpha);
    ABHSF_HIP(hipGetLastError());
}

}  // namespace abhsf
