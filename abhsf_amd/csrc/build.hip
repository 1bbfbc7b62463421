// build.hip -- CSR -> ABHSF conversion on the GPU.
//
// The reference computes the storage footprint by sorting entries by block key
// and walking maximal runs of equal (I,J) (preprocessing/test.cpp:237-368),
// calling a per-block cost processor (test.cpp:36-108).  Here every step is a
// data-parallel pass over HBM-resident arrays:
//   1. validate CSR, expand rows, key = (I << bits(N)) | J, payload = entry id
//   2. stable LSD radix sort by key -> entries in (I, J, lrow, lcol) order
//      (CSR order is already (row, col), so stability gives the in-block order
//      of spmv/orig/spmv.cpp:85-110)
//   3. head flags + exclusive scan -> block id per entry, block starts
//   4. per block: z, scheme by the cost model, footprint bits
//   5. per scheme: exclusive scans of (is scheme X) and (values of X) -> each
//      block's ordinal and value offset inside its scheme group
//   6. strip pointers, block metadata, and the entry emit (COO index, CSR row
//      pointers + columns, bitmap bits by atomicOr, dense positions)
// All arithmetic is integer, so the result is bit-exact and deterministic.
#include "storage.hpp"

#include <algorithm>
#include <chrono>

namespace abhsf {
namespace {

constexpr int kT = 256;

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kT - 1) / kT); }

__device__ inline void st_uint(void *p, int bytes, uint64_t i, uint32_t v)
{
    if (bytes == 1) static_cast<uint8_t *>(p)[i] = (uint8_t)v;
    else if (bytes == 2) static_cast<uint16_t *>(p)[i] = (uint16_t)v;
    else static_cast<uint32_t *>(p)[i] = v;
}

__device__ inline void st_val(void *p, int bytes, uint64_t i, double v)
{
    if (bytes == 8) static_cast<double *>(p)[i] = v;
    else static_cast<float *>(p)[i] = (float)v;
}

template <typename T>
struct Four {
    T p[4];
};

__global__ void k_check_rowptr(const int64_t *__restrict__ rp, int64_t m, int64_t nnz, uint32_t *err)
{
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (r > m) return;
    if (r == 0 && rp[0] != 0) atomicOr(err, 1u);
    if (r == m && rp[m] != nnz) atomicOr(err, 2u);
    if (r < m && rp[r + 1] < rp[r]) atomicOr(err, 4u);
}

// Triangle flags of a CSR matrix (columns sorted within rows): bit 0 some
// entry below the diagonal, bit 1 some entry above, bit 2 some diagonal entry.
__global__ void k_triangle(const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int64_t m,
                           uint32_t *flags)
{
    const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
    uint32_t f = 0;
    if (r < m && rp[r + 1] > rp[r]) {
        int64_t lo = rp[r], hi = rp[r + 1] - 1;
        if (col[lo] < r) f |= 1u;
        if (col[hi] > r) f |= 2u;
        while (lo < hi) {  // first column >= r
            const int64_t mid = (lo + hi) >> 1;
            if (col[mid] < r) lo = mid + 1; else hi = mid;
        }
        if (col[lo] == r) f |= 4u;
    }
    // one atomic per wave and flag value
    const uint64_t any = __ballot(f != 0);
    if (any) {
        uint32_t all = f;
        for (int o = 32; o >= 1; o >>= 1) all |= (uint32_t)__shfl_xor((int)all, o, 64);
        if ((threadIdx.x & 63) == 0) atomicOr(flags, all);
    }
}

__global__ void k_rows_keys(const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int64_t m, int64_t n,
                            uint64_t nnz, int k, int jbits, uint32_t *__restrict__ rows,
                            uint64_t *__restrict__ key, uint32_t *__restrict__ perm, uint32_t *err)
{
    const uint64_t e = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= nnz) return;
    int64_t lo = 0, hi = m - 1;  // last row r with rp[r] <= e
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((uint64_t)rp[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const int32_t c = col[e];
    if (c < 0 || (int64_t)c >= n) atomicOr(err, 8u);
    if ((int64_t)e > rp[lo] && c <= col[e - 1]) atomicOr(err, 16u);
    rows[e] = (uint32_t)lo;
    key[e] = ((uint64_t)((uint64_t)lo >> k) << jbits) | (uint64_t)((uint32_t)c >> k);
    perm[e] = (uint32_t)e;
}

__global__ void k_heads(const uint64_t *__restrict__ key, uint64_t nnz, uint32_t *__restrict__ flag)
{
    const uint64_t e = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= nnz) return;
    flag[e] = (e == 0 || key[e] != key[e - 1]) ? 1u : 0u;
}

__global__ void k_block_starts(const uint64_t *__restrict__ key, const uint32_t *__restrict__ flag,
                               const uint32_t *__restrict__ excl, uint64_t nnz, uint32_t nblk,
                               uint32_t *__restrict__ bstart, uint64_t *__restrict__ bkey)
{
    const uint64_t e = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (e == 0) bstart[nblk] = (uint32_t)nnz;
    if (e >= nnz || !flag[e]) return;
    const uint32_t b = excl[e];
    bstart[b] = (uint32_t)e;
    bkey[b] = key[e];
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// z, scheme, per-scheme indicator/value counts; per-workgroup partial sums of
// the footprints and value counts (partial[q * gridDim.x + workgroup]).
__global__ __launch_bounds__(kT) void k_block_scheme(const uint32_t *__restrict__ bstart, uint32_t nblk, int s,
                                                     int policy, uint64_t cost_b, int is_binary, uint64_t hdr_bits,
                                                     uint8_t *__restrict__ sch, Four<uint32_t *> cnt,
                                                     Four<uint32_t *> ecnt, unsigned long long *__restrict__ partial)
{
    __shared__ unsigned long long red[kT / kWave][6];
    const uint32_t b = blockIdx.x * kT + threadIdx.x;
    unsigned long long v[6] = {0, 0, 0, 0, 0, 0};
    if (b < nblk) {
        const uint64_t z = bstart[b + 1] - bstart[b];
        uint64_t bits;
        const int x = choose_scheme(policy, (uint64_t)s, z, cost_b, is_binary, nullptr);
        sch[b] = (uint8_t)x;
        const uint32_t evx = (x == DENSE) ? (uint32_t)(s * s) : (uint32_t)z;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            cnt.p[q][b] = (q == x) ? 1u : 0u;
            ecnt.p[q][b] = (q == x) ? evx : 0u;
            v[2 + q] = (q == x) ? evx : 0u;
        }
        choose_scheme(ABHSF_POLICY_PREPROC, (uint64_t)s, z, cost_b, is_binary, &bits);
        v[0] = hdr_bits + bits;  // preprocessing/test.cpp:76
        choose_scheme(ABHSF_POLICY_MSMF, (uint64_t)s, z, cost_b, is_binary, &bits);
        v[1] = bits + 2;         // block_statistics/mmf.cpp:93-96
    }
#pragma unroll
    for (int q = 0; q < 6; q++) v[q] = wave_sum_u64(v[q]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 6; q++) red[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 6) {
        unsigned long long a = 0;
        for (int w = 0; w < kT / kWave; w++) a += red[w][threadIdx.x];
        partial[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
    }
}

__global__ void k_strip_ptrs(const uint64_t *__restrict__ bkey, uint32_t nblk, int jbits, int64_t M,
                             Four<const uint32_t *> ord, Four<const uint32_t *> eoff, Four<uint32_t> nbx,
                             Four<uint32_t> nex, Four<uint32_t *> sptr, Four<uint32_t *> eptr)
{
    const uint32_t b = blockIdx.x * kT + threadIdx.x;
    if (b >= nblk) return;
    const int64_t I = (int64_t)(bkey[b] >> jbits);
    const int64_t Ip = b ? (int64_t)(bkey[b - 1] >> jbits) : -1;
    for (int64_t i = Ip + 1; i <= I; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            sptr.p[q][i] = ord.p[q][b];
            eptr.p[q][i] = eoff.p[q][b];
        }
    if (b == nblk - 1)
        for (int64_t i = I + 1; i <= M; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                sptr.p[q][i] = nbx.p[q];
                eptr.p[q][i] = nex.p[q];
            }
}

__global__ void k_block_meta(const uint64_t *__restrict__ bkey, const uint32_t *__restrict__ bstart,
                             const uint8_t *__restrict__ sch, uint32_t nblk, uint64_t jmask, int bcol_bytes,
                             Four<const uint32_t *> ord, Four<void *> bcol, Four<uint16_t *> z1)
{
    const uint32_t b = blockIdx.x * kT + threadIdx.x;
    if (b >= nblk) return;
    const int x = sch[b];
    const uint32_t o = ord.p[x][b];
    st_uint(bcol.p[x], bcol_bytes, o, (uint32_t)(bkey[b] & jmask));
    z1.p[x][o] = (uint16_t)(bstart[b + 1] - bstart[b] - 1);
}

struct EmitOut {
    void *val[4];
    void *coo_idx;
    void *csr_rowptr;
    uint8_t *csr_col;
    unsigned long long *bm_words;
    int coo_idx_bytes, csr_rp_bytes, val_bytes, W;
};

__global__ void k_emit(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ excl,
                       const uint32_t *__restrict__ bstart, const uint8_t *__restrict__ sch,
                       const uint32_t *__restrict__ perm, const uint32_t *__restrict__ rows,
                       const int32_t *__restrict__ col, const double *__restrict__ val, uint64_t nnz, int s, int k,
                       Four<const uint32_t *> ord, Four<const uint32_t *> eoff, EmitOut out)
{
    const uint64_t e = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= nnz) return;
    const uint32_t b = excl[e] + flag[e] - 1u;
    const uint32_t q = (uint32_t)e - bstart[b];
    const uint32_t z = bstart[b + 1] - bstart[b];
    const uint32_t el = perm[e];
    const uint32_t smask = (uint32_t)s - 1u;
    const uint32_t lr = rows[el] & smask, lc = (uint32_t)col[el] & smask;
    const double v = val ? val[el] : 1.0;
    const int x = sch[b];
    const uint32_t o = ord.p[x][b];
    const uint64_t eo = eoff.p[x][b];
    switch (x) {
    case COO:
        st_uint(out.coo_idx, out.coo_idx_bytes, eo + q, (lr << k) | lc);
        st_val(out.val[COO], out.val_bytes, eo + q, v);
        break;
    case CSR: {
        out.csr_col[eo + q] = (uint8_t)lc;
        st_val(out.val[CSR], out.val_bytes, eo + q, v);
        const uint64_t rb = (uint64_t)o * (uint64_t)(s + 1);
        const int prev = (q == 0) ? -1 : (int)(rows[perm[e - 1]] & smask);
        for (int r = prev + 1; r <= (int)lr; r++) st_uint(out.csr_rowptr, out.csr_rp_bytes, rb + r, q);
        if (q == z - 1)
            for (int r = (int)lr + 1; r <= s; r++) st_uint(out.csr_rowptr, out.csr_rp_bytes, rb + r, z);
        break;
    }
    case BITMAP: {
        const uint32_t p = lr * (uint32_t)s + lc;
        atomicOr(&out.bm_words[(uint64_t)o * out.W + (p >> 6)], 1ull << (p & 63));
        st_val(out.val[BITMAP], out.val_bytes, eo + q, v);
        break;
    }
    default:
        st_val(out.val[DENSE], out.val_bytes, eo + (uint64_t)lr * s + lc, v);
    }
}

// Tiled COO order, step 1: sort key of every COO block b = (group of its
// strip, block column, strip within the group).  Groups are consecutive strip
// ranges numbered in strip order (a tile, or one strip), so sorting by the
// key reorders blocks only inside tiles.  The strip of b is the last I with
// sptr[I] <= b (empty strips share their start with the next strip).
__global__ void k_tile_keys(const uint32_t *__restrict__ sptr, int64_t M, uint64_t nb, const void *__restrict__ bcol,
                            int bcol_bytes, const uint32_t *__restrict__ grp, const uint32_t *__restrict__ tpos,
                            int bits_bc, int bits_t, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    const uint64_t b = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    int64_t lo = 0, hi = M - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((uint64_t)sptr[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const uint64_t J = bcol_bytes == 2 ? static_cast<const uint16_t *>(bcol)[b] : static_cast<const uint32_t *>(bcol)[b];
    keys[b] = ((uint64_t)grp[lo] << (bits_bc + bits_t)) | (J << bits_t) | (uint64_t)tpos[lo];
    vals[b] = (uint32_t)b;
}

// Step 2: new block i takes old block perm[i]: block column copied, z1 gets
// the strip within the tile in its bits [zb, 16); znew[i] = entry count.
__global__ void k_tile_gather(uint64_t nb, const uint32_t *__restrict__ perm, const uint64_t *__restrict__ keys,
                              int bits_t, int zb, const void *__restrict__ bc_old, void *__restrict__ bc_new,
                              int bcol_bytes, const uint16_t *__restrict__ z1_old, uint16_t *__restrict__ z1_new,
                              uint32_t *__restrict__ znew)
{
    const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= nb) return;
    const uint32_t b = perm[i];
    const uint32_t t = (uint32_t)(keys[i] & ((1ull << bits_t) - 1ull));
    if (bcol_bytes == 2) static_cast<uint16_t *>(bc_new)[i] = static_cast<const uint16_t *>(bc_old)[b];
    else static_cast<uint32_t *>(bc_new)[i] = static_cast<const uint32_t *>(bc_old)[b];
    const uint32_t z1 = z1_old[b];
    z1_new[i] = (uint16_t)(z1 | (t << zb));
    znew[i] = z1 + 1u;
}

__global__ void k_z_count(uint64_t nb, const uint16_t *__restrict__ z1, uint32_t *__restrict__ z)
{
    const uint64_t b = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (b < nb) z[b] = (uint32_t)z1[b] + 1u;
}

// Step 3: the entries (values and in-block indices) of new block i move from
// the old block's offset to the new one.
__global__ void k_tile_move(uint64_t nb, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ off_old,
                            const uint32_t *__restrict__ off_new, const uint32_t *__restrict__ znew,
                            const void *__restrict__ v_old, void *__restrict__ v_new, int val_bytes,
                            const void *__restrict__ i_old, void *__restrict__ i_new, int idx_bytes)
{
    const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= nb) return;
    const uint64_t src = off_old[perm[i]], dst = off_new[i];
    const uint32_t z = znew[i];
    for (uint32_t j = 0; j < z; j++) {
        if (val_bytes == 8) static_cast<double *>(v_new)[dst + j] = static_cast<const double *>(v_old)[src + j];
        else static_cast<float *>(v_new)[dst + j] = static_cast<const float *>(v_old)[src + j];
        if (idx_bytes == 1) static_cast<uint8_t *>(i_new)[dst + j] = static_cast<const uint8_t *>(i_old)[src + j];
        else static_cast<uint16_t *>(i_new)[dst + j] = static_cast<const uint16_t *>(i_old)[src + j];
    }
}

// First block column of every unit (units[u] = {scheme | strip << 2, b0, b1, e0}).
__global__ void k_unit_cols(const uint32_t *__restrict__ units, int64_t n_units, Four<const void *> bcol,
                            int bcol_bytes, uint32_t *__restrict__ out)
{
    const int64_t u = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (u >= n_units) return;
    const uint32_t x = units[4 * u] & 3u, b0 = units[4 * u + 1];
    out[u] = bcol_bytes == 2 ? static_cast<const uint16_t *>(bcol.p[x])[b0] : static_cast<const uint32_t *>(bcol.p[x])[b0];
}

template <typename T>
T read_dev(const T *p, hipStream_t st)
{
    T h;
    ABHSF_HIP(hipMemcpyAsync(&h, p, sizeof(T), hipMemcpyDeviceToHost, st));
    ABHSF_HIP(hipStreamSynchronize(st));
    return h;
}

int bits_for(uint64_t maxval)
{
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) b++;
    return b;
}

}  // namespace

void Storage::set_shape(int64_t m_, int64_t n_, int s_)
{
    m = m_;
    n = n_;
    s = s_;
    k = (int)ceil_log2_u64((uint64_t)s_);
    M = (m + s - 1) / s;
    N = (n + s - 1) / s;
    W = s * s / 64 > 0 ? s * s / 64 : 1;
    bcol_bytes = N <= 65536 ? 2 : 4;
    coo_idx_bytes = s <= 16 ? 1 : 2;
    csr_rp_bytes = s <= 8 ? 1 : 2;
}

DevMat Storage::view() const
{
    DevMat A;
    A.m = m; A.n = n; A.M = M; A.N = N;
    A.s = s; A.k = k; A.W = W;
    A.zb = coo_zb;
    for (int x = 0; x < 4; x++) {
        A.sptr[x] = sptr[x].get<uint32_t>();
        A.eptr[x] = eptr[x].get<uint32_t>();
        A.bcol[x] = bcol[x].get();
        A.z1[x] = z1[x].get<uint16_t>();
        A.val[x] = val[x].get();
    }
    A.coo_idx = coo_idx.get();
    A.csr_rowptr = csr_rowptr.get();
    A.csr_col = csr_col.get<uint8_t>();
    A.bm_words = bm_words.get<uint64_t>();
    return A;
}

SpmvLaunch Storage::launch() const
{
    SpmvLaunch L;
    L.bcol_bytes = bcol_bytes;
    L.val_bytes = val_bytes;
    L.coo_single = nb[COO] > 0 && 4 * ne[COO] < 5 * nb[COO];  // mean entries per COO block below 1.25
    L.has_dense = nb[DENSE] > 0;
    L.bal.light = bal_light.get<uint32_t>();
    L.bal.pack = bal_pack.get<uint32_t>();
    L.bal.units = bal_units.get<uint32_t>();
    L.bal.heavy = bal_heavy.get<uint32_t>();
    L.bal.fix = n_fix ? bal_heavy.get<uint32_t>() + 3 * n_heavy : nullptr;
    L.bal.part = bal_part.get<double>();
    L.bal.order = bal_order.get<uint32_t>();
    L.bal.n_light = n_light;
    L.bal.n_units = n_units;
    L.bal.n_heavy = n_heavy;
    L.bal.n_slots = n_slots;
    L.bal.n_fix = n_fix;
    L.bal.tiled = tiled;
    L.bal.tile_zs1 = tile_zs1;
    L.bal.coo_rng = bal_coo_rng.get<uint32_t>();
    return L;
}

// Reorder the COO group of S by the key (rank of the block's strip, block
// column, position of the strip in its tile): grp[I] ranks strip I's group
// (a tile or a single strip; ranks follow the groups' first strips), tpos[I]
// is its position in the tile (0 outside tiles).  z1 gains that position in
// its bits [coo_zb, 16).
static void tile_coo_order(Storage &S, const std::vector<uint32_t> &grp, const std::vector<uint32_t> &tpos,
                           uint64_t nrank, int64_t tcap, hipStream_t st)
{
    const int64_t M = S.M;
    const uint64_t nb = S.nb[COO];
    const uint64_t g = nrank - 1;
    const int bits_bc = bits_for((uint64_t)std::max<int64_t>(S.N - 1, 1));
    const int bits_t = bits_for((uint64_t)std::max<int64_t>(tcap - 1, 1));
    const int key_bits = bits_for(g) + bits_bc + bits_t;
    ABHSF_REQUIRE(key_bits <= 64, ABHSF_ERR_UNSUPPORTED, "tile sort key exceeds 64 bits");
    DevBuf dgrp((size_t)M * 4), dtpos((size_t)M * 4), keys(nb * 8), perm(nb * 4);
    ABHSF_HIP(hipMemcpyAsync(dgrp.get(), grp.data(), dgrp.bytes(), hipMemcpyHostToDevice, st));
    ABHSF_HIP(hipMemcpyAsync(dtpos.get(), tpos.data(), dtpos.bytes(), hipMemcpyHostToDevice, st));
    k_tile_keys<<<grid_for(nb), kT, 0, st>>>(S.sptr[COO].get<uint32_t>(), M, nb, S.bcol[COO].get(), S.bcol_bytes,
                                             dgrp.get<uint32_t>(), dtpos.get<uint32_t>(), bits_bc, bits_t,
                                             keys.get<uint64_t>(), perm.get<uint32_t>());
    ABHSF_HIP(hipGetLastError());
    radix_sort_pairs(keys.get<uint64_t>(), perm.get<uint32_t>(), nb, key_bits, st);
    DevBuf bc_new(S.bcol[COO].bytes()), z1_new(S.z1[COO].bytes()), znew(nb * 4), off_old(nb * 4), off_new(nb * 4);
    k_z_count<<<grid_for(nb), kT, 0, st>>>(nb, S.z1[COO].get<uint16_t>(), off_old.get<uint32_t>());
    ABHSF_HIP(hipGetLastError());
    exclusive_scan_u32(off_old.get<uint32_t>(), off_old.get<uint32_t>(), nb, nullptr, st);
    k_tile_gather<<<grid_for(nb), kT, 0, st>>>(nb, perm.get<uint32_t>(), keys.get<uint64_t>(), bits_t, S.coo_zb,
                                               S.bcol[COO].get(), bc_new.get(), S.bcol_bytes,
                                               S.z1[COO].get<uint16_t>(), z1_new.get<uint16_t>(),
                                               znew.get<uint32_t>());
    ABHSF_HIP(hipGetLastError());
    exclusive_scan_u32(znew.get<uint32_t>(), off_new.get<uint32_t>(), nb, nullptr, st);
    DevBuf v_new(S.val[COO].bytes()), i_new(S.coo_idx.bytes());
    k_tile_move<<<grid_for(nb), kT, 0, st>>>(nb, perm.get<uint32_t>(), off_old.get<uint32_t>(),
                                             off_new.get<uint32_t>(), znew.get<uint32_t>(), S.val[COO].get(),
                                             v_new.get(), S.val_bytes, S.coo_idx.get(), i_new.get(),
                                             S.coo_idx_bytes);
    ABHSF_HIP(hipGetLastError());
    ABHSF_HIP(hipStreamSynchronize(st));
    S.bcol[COO] = std::move(bc_new);
    S.z1[COO] = std::move(z1_new);
    S.val[COO] = std::move(v_new);
    S.coo_idx = std::move(i_new);
}

void Storage::build_balance(int64_t unit, hipStream_t st, int panels, int tile_rows)
{
    ABHSF_REQUIRE(!tiled, ABHSF_ERR_STATE, "the COO blocks are already in tile order");
    bal_light.release();
    bal_pack.release();
    bal_units.release();
    bal_heavy.release();
    bal_part.release();
    bal_order.release();
    bal_coo_rng.release();
    n_light = n_units = n_heavy = n_slots = n_fix = 0;
    tile_zs1 = false;
    if (unit <= 0 || M == 0) return;
    ABHSF_REQUIRE(panels == 1 || panels == kXcdCount, ABHSF_ERR_INVALID, "column panels must match the XCD count");
    ABHSF_REQUIRE(tile_rows <= 0 || panels == 1, ABHSF_ERR_INVALID, "tiles and column panels do not combine");
    std::vector<uint32_t> ep[4], sp[4];
    for (int x = 0; x < 4; x++) {
        ep[x].resize((size_t)M + 1);
        sp[x].resize((size_t)M + 1);
        ABHSF_HIP(hipMemcpyAsync(ep[x].data(), eptr[x].get(), ((size_t)M + 1) * 4, hipMemcpyDeviceToHost, st));
        ABHSF_HIP(hipMemcpyAsync(sp[x].data(), sptr[x].get(), ((size_t)M + 1) * 4, hipMemcpyDeviceToHost, st));
    }
    ABHSF_HIP(hipStreamSynchronize(st));
    const int64_t H = 2 * unit;
    const int64_t T = unit / 2;     // light strips are packed into items of up to T entries
    constexpr int64_t kTiny = 128;  // ... if each holds at most kTiny entries
    auto strip_work = [&](int64_t I) {
        int64_t w = 0;
        for (int x = 0; x < 4; x++) w += (int64_t)ep[x][I + 1] - ep[x][I];
        return w;
    };
    if (panels == 1 && tile_rows <= 0) {  // nothing to split or pack: one item per strip, no tables
        bool any = false;
        for (int64_t I = 0; I < M && !any; I++) {
            const int64_t w = strip_work(I);
            any = w > H || (I + 1 < M && w <= kTiny && strip_work(I + 1) <= kTiny);
        }
        if (!any) return;
    }
    std::vector<uint16_t> zz[4];
    for (int x = 0; x < 3; x++) {
        zz[x].resize(nb[x]);
        if (nb[x]) ABHSF_HIP(hipMemcpyAsync(zz[x].data(), z1[x].get(), nb[x] * 2, hipMemcpyDeviceToHost, st));
    }
    // block columns (only needed for panels)
    std::vector<uint32_t> bc[4];
    std::vector<uint16_t> bc16;
    if (panels > 1) {
        for (int x = 0; x < 4; x++) {
            bc[x].resize(nb[x]);
            if (!nb[x]) continue;
            if (bcol_bytes == 4) {
                ABHSF_HIP(hipMemcpyAsync(bc[x].data(), bcol[x].get(), nb[x] * 4, hipMemcpyDeviceToHost, st));
            } else {
                bc16.resize(nb[x]);
                ABHSF_HIP(hipMemcpyAsync(bc16.data(), bcol[x].get(), nb[x] * 2, hipMemcpyDeviceToHost, st));
                ABHSF_HIP(hipStreamSynchronize(st));
                for (uint64_t b = 0; b < nb[x]; b++) bc[x][b] = bc16[b];
            }
        }
    }
    ABHSF_HIP(hipStreamSynchronize(st));
    const int64_t ss = (int64_t)s * s;
    auto work = [&](int x, uint32_t b) -> int64_t { return (x == DENSE) ? ss : (int64_t)zz[x][b] + 1; };
    // panel of every block column: cut where the running work passes p/panels
    std::vector<uint8_t> panel_of;
    if (panels > 1) {
        std::vector<uint64_t> cw((size_t)N, 0);
        uint64_t total = 0;
        for (int x = 0; x < 4; x++)
            for (uint64_t b = 0; b < nb[x]; b++) {
                cw[bc[x][b]] += (uint64_t)work(x, (uint32_t)b);
                total += (uint64_t)work(x, (uint32_t)b);
            }
        panel_of.resize((size_t)N);
        uint64_t acc = 0;
        int p = 0;
        for (int64_t J = 0; J < N; J++) {
            panel_of[J] = (uint8_t)p;
            acc += cw[J];
            while (p < panels - 1 && acc * (uint64_t)panels >= total * (uint64_t)(p + 1)) p++;
        }
    }
    std::vector<uint32_t> light, units, heavy;
    std::vector<uint64_t> unit_work;  // entries (dense: positions) of every unit
    std::vector<uint8_t> light_panel, unit_panel;
    std::vector<int64_t> light_work;
    light.reserve((size_t)M);
    for (int64_t I = 0; I < M; I++) {
        int64_t w = 0;
        int p0 = -1;
        bool multi = false;
        for (int x = 0; x < 4; x++) {
            w += (int64_t)ep[x][I + 1] - ep[x][I];
            if (panels > 1)
                for (uint32_t b = sp[x][I]; b < sp[x][I + 1] && !multi; b++) {
                    const int p = panel_of[bc[x][b]];
                    if (p0 < 0) p0 = p; else multi = p != p0;
                }
        }
        if (w <= H && !multi) {
            light.push_back((uint32_t)I);
            light_panel.push_back((uint8_t)(p0 >= 0 ? p0 : (int)(I % panels)));
            light_work.push_back(w);
            continue;
        }
        const uint32_t first = (uint32_t)(units.size() / 4);
        for (int x = 0; x < 4; x++) {
            const uint32_t b0 = sp[x][I], b1 = sp[x][I + 1];
            uint64_t e = ep[x][I], ue = e;
            uint32_t ub = b0;
            int up = 0;
            int64_t acc = 0;
            for (uint32_t b = b0; b < b1; b++) {
                const int64_t wz = work(x, b);
                const int p = panels > 1 ? panel_of[bc[x][b]] : 0;
                if (b > ub && (p != up || acc + wz > unit)) {
                    units.insert(units.end(), {(uint32_t)x | ((uint32_t)I << 2), ub, b, (uint32_t)ue});
                    unit_panel.push_back((uint8_t)up);
                    unit_work.push_back((uint64_t)acc);
                    ub = b;
                    ue = e;
                    acc = 0;
                }
                if (b == ub) up = p;
                acc += wz;
                e += (uint64_t)wz;
            }
            if (ub < b1) {
                units.insert(units.end(), {(uint32_t)x | ((uint32_t)I << 2), ub, b1, (uint32_t)ue});
                unit_panel.push_back((uint8_t)up);
                unit_work.push_back((uint64_t)acc);
            }
        }
        heavy.insert(heavy.end(), {(uint32_t)I, first, (uint32_t)(units.size() / 4 - first)});
    }
    // pack runs of consecutive tiny light strips (<= kTiny entries each, one
    // panel) into items of <= T entries and at most kPackMax strips; larger
    // strips stay items of their own (packing those serialises work that
    // separate wavefronts overlap: block_mix s = 16 lost 6 %)
    constexpr size_t kPackMax = 64;
    if (panels > 1) {  // group the light strips by panel, strip order kept inside a panel
        std::vector<size_t> idx(light.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return light_panel[a] < light_panel[b]; });
        std::vector<uint32_t> l2(light.size());
        std::vector<uint8_t> p2(light.size());
        std::vector<int64_t> w2(light.size());
        for (size_t i = 0; i < idx.size(); i++) {
            l2[i] = light[idx[i]];
            p2[i] = light_panel[idx[i]];
            w2[i] = light_work[idx[i]];
        }
        light.swap(l2);
        light_panel.swap(p2);
        light_work.swap(w2);
    }
    std::vector<uint32_t> pack;
    std::vector<uint8_t> item_panel;
    // tiles: the strip within a tile goes to the COO z1 bits above the
    // largest z - 1, and the sort key (group, block column, strip) must fit
    // 64 bits; otherwise every tile holds one strip
    int zb = 16;
    int64_t tcap = 1;
    if (tile_rows > 0) {
        uint32_t zmax = 0;
        for (uint64_t b = 0; b < nb[COO]; b++) zmax = std::max<uint32_t>(zmax, zz[COO][b]);
        zb = bits_for(zmax);
        tcap = std::min<int64_t>(tile_rows / s, zb < 16 ? (int64_t)1 << (16 - zb) : 1);
        if (bits_for((uint64_t)M) + bits_for((uint64_t)N) + bits_for((uint64_t)std::max<int64_t>(tcap - 1, 1)) > 64)
            tcap = 1;
    }
    std::vector<uint32_t> grank, gtpos, coo_rng;
    uint64_t nrank = 0;
    if (tile_rows > 0) {
        // items: tiles of light strips whose entries are all COO (or none), up
        // to tcap strips and kTileWork entries, gathered in strip order past
        // strips with other groups (each of those stays an item of its own)
        std::vector<std::vector<uint32_t>> items;
        int64_t open = -1, open_work = 0;
        for (size_t i = 0; i < light.size(); i++) {
            const uint32_t I = light[i];
            const int64_t cw = (int64_t)ep[COO][I + 1] - ep[COO][I];
            const bool pure = tcap > 1 && light_work[i] == cw;
            if (pure && open >= 0 && (int64_t)items[(size_t)open].size() < tcap && open_work + cw <= kTileWork) {
                items[(size_t)open].push_back(I);
                open_work += cw;
                continue;
            }
            items.push_back({I});
            if (pure) {
                open = (int64_t)items.size() - 1;
                open_work = cw;
            }
        }
        // blocks of 3 or more entries among the tiles' COO blocks (z - 1 >= 2):
        // at one in 64 or more, most 64-block sub-chunks run phase B anyway
        // and phase A takes one entry per block (R-MAT scale 22, s = 256:
        // 9 % of them, 0.527 -> 0.51 ms; uniform scatter: none, and one entry
        // per block would send the 0.15 % of blocks with two to phase B)
        uint64_t tile_blocks = 0, tile_ge3 = 0;
        for (const auto &it : items) {
            if (it.size() < 2) continue;
            for (uint32_t I : it)
                for (uint64_t b = sp[COO][I]; b < sp[COO][I + 1]; b++) {
                    tile_blocks++;
                    tile_ge3 += zz[COO][b] >= 2u;
                }
        }
        tile_zs1 = tile_blocks > 0 && 64 * tile_ge3 >= tile_blocks;
        // tiles first: they start together and sweep x in step (their COO
        // streams run in block-column order), the single strips fill the
        // remaining slots and follow
        std::stable_partition(items.begin(), items.end(), [](const std::vector<uint32_t> &it) { return it.size() > 1; });
        std::vector<uint32_t> l2;
        l2.reserve(light.size());
        for (auto &it : items) {
            pack.push_back((uint32_t)l2.size());
            l2.insert(l2.end(), it.begin(), it.end());
        }
        pack.push_back((uint32_t)l2.size());
        light.swap(l2);
        // new COO block order: groups (an item, or a heavy strip) ranked by
        // their first strip; inside a group by (block column, strip)
        std::vector<int64_t> item_of((size_t)M, -1), rank_of_item(items.size(), -1), strip_rank((size_t)M);
        grank.assign((size_t)M, 0u);
        gtpos.assign((size_t)M, 0u);
        for (size_t q = 0; q + 1 < pack.size(); q++)
            for (uint32_t p = pack[q]; p < pack[q + 1]; p++) {
                item_of[light[p]] = (int64_t)q;
                gtpos[light[p]] = p - pack[q];
            }
        std::vector<uint64_t> gb, ge;  // blocks and values of every rank
        for (int64_t I = 0; I < M; I++) {
            const int64_t it = item_of[(size_t)I];
            int64_t r = it >= 0 ? rank_of_item[(size_t)it] : -1;
            if (r < 0) {
                r = (int64_t)gb.size();
                gb.push_back(0);
                ge.push_back(0);
                if (it >= 0) rank_of_item[(size_t)it] = r;
            }
            strip_rank[(size_t)I] = r;
            grank[(size_t)I] = (uint32_t)r;
            gb[(size_t)r] += sp[COO][I + 1] - sp[COO][I];
            ge[(size_t)r] += ep[COO][I + 1] - ep[COO][I];
        }
        nrank = gb.size();
        std::vector<uint64_t> bs(nrank + 1, 0), es(nrank + 1, 0);
        for (size_t r = 0; r < nrank; r++) {
            bs[r + 1] = bs[r] + gb[r];
            es[r + 1] = es[r] + ge[r];
        }
        coo_rng.resize(items.size() * 3);
        for (size_t q = 0; q < items.size(); q++) {
            const size_t r = (size_t)rank_of_item[q];
            coo_rng[3 * q] = (uint32_t)bs[r];
            coo_rng[3 * q + 1] = (uint32_t)bs[r + 1];
            coo_rng[3 * q + 2] = (uint32_t)es[r];
        }
        // the COO units of a heavy strip move with the strip
        for (size_t u = 0; u < units.size() / 4; u++) {
            if ((units[4 * u] & 3u) != (uint32_t)COO) continue;
            const uint32_t I = units[4 * u] >> 2;
            const size_t r = (size_t)strip_rank[I];
            units[4 * u + 1] = (uint32_t)(units[4 * u + 1] - sp[COO][I] + bs[r]);
            units[4 * u + 2] = (uint32_t)(units[4 * u + 2] - sp[COO][I] + bs[r]);
            units[4 * u + 3] = (uint32_t)(units[4 * u + 3] - ep[COO][I] + es[r]);
        }
    } else {
        int64_t iw = 0;
        bool tiny_item = false;  // the open item packs tiny strips
        for (size_t i = 0; i < light.size(); i++) {
            const bool tiny = light_work[i] <= kTiny;
            const bool fits = !pack.empty() && tiny && tiny_item && light_panel[i] == item_panel.back() &&
                              iw + light_work[i] <= T && i - pack.back() < kPackMax;
            if (!fits) {
                pack.push_back((uint32_t)i);
                item_panel.push_back(light_panel[i]);
                iw = 0;
                tiny_item = tiny;
            }
            iw += light_work[i];
        }
        pack.push_back((uint32_t)light.size());
    }
    n_light = (int64_t)pack.size() - 1;
    n_units = (int64_t)(units.size() / 4);
    n_heavy = (int64_t)(heavy.size() / 3);
    ABHSF_REQUIRE(n_light + n_units < (int64_t)kNoItem, ABHSF_ERR_UNSUPPORTED, "too many work items");
    // fix-up workgroups: row slices of every heavy strip, narrower for the
    // strips with more units (stored after the heavy strips)
    std::vector<uint32_t> fix;
    for (int64_t h = 0; h < n_heavy; h++) {
        const int rs = fix_rows(s, heavy[(size_t)h * 3 + 2], val_bytes == 4);
        for (int r0 = 0; r0 < s; r0 += rs) fix.insert(fix.end(), {(uint32_t)h, (uint32_t)r0 | ((uint32_t)rs << 16)});
    }
    n_fix = (int64_t)(fix.size() / 2);
    ABHSF_REQUIRE(n_fix < (1ll << 31), ABHSF_ERR_UNSUPPORTED, "too many heavy strips");
    heavy.insert(heavy.end(), fix.begin(), fix.end());
    std::vector<uint32_t> order;
    if (panels > 1) {
        std::vector<std::vector<uint32_t>> lists((size_t)panels);
        for (int64_t i = 0; i < n_light; i++) lists[item_panel[i]].push_back((uint32_t)i);
        for (int64_t u = 0; u < n_units; u++) lists[unit_panel[u]].push_back((uint32_t)(n_light + u));
        size_t G = 0;
        const size_t w = (size_t)kStripWaves;
        for (auto &l : lists) G = std::max(G, (l.size() + w - 1) / w);
        order.assign(G * (size_t)panels * w, kNoItem);
        for (int p = 0; p < panels; p++)
            for (size_t j = 0; j < lists[p].size(); j++)
                order[((j / w) * (size_t)panels + (size_t)p) * w + j % w] = lists[p][j];
        n_slots = (int64_t)order.size();
    }
    bal_light.alloc(light.size() * 4);
    bal_pack.alloc(pack.size() * 4);
    ABHSF_HIP(hipMemcpyAsync(bal_pack.get(), pack.data(), pack.size() * 4, hipMemcpyHostToDevice, st));
    bal_units.alloc(units.size() * 4);
    bal_heavy.alloc(heavy.size() * 4);
    bal_order.alloc(order.size() * 4);
    bal_part.alloc((size_t)n_units * (size_t)s * sizeof(double) * (val_bytes == 4 ? 2 : 1));
    if (!light.empty())
        ABHSF_HIP(hipMemcpyAsync(bal_light.get(), light.data(), light.size() * 4, hipMemcpyHostToDevice, st));
    if (!units.empty())
        ABHSF_HIP(hipMemcpyAsync(bal_units.get(), units.data(), units.size() * 4, hipMemcpyHostToDevice, st));
    if (!heavy.empty())
        ABHSF_HIP(hipMemcpyAsync(bal_heavy.get(), heavy.data(), heavy.size() * 4, hipMemcpyHostToDevice, st));
    if (!order.empty())
        ABHSF_HIP(hipMemcpyAsync(bal_order.get(), order.data(), order.size() * 4, hipMemcpyHostToDevice, st));
    ABHSF_HIP(hipStreamSynchronize(st));
    if (tile_rows > 0) {
        bool multi = false;
        for (size_t i = 0; i + 1 < pack.size() && !multi; i++) multi = pack[i + 1] - pack[i] > 1;
        coo_zb = zb;
        if (multi && nb[COO] > 0) tile_coo_order(*this, grank, gtpos, nrank, tcap, st);
        bal_coo_rng.alloc(coo_rng.size() * 4);
        if (!coo_rng.empty())
            ABHSF_HIP(hipMemcpyAsync(bal_coo_rng.get(), coo_rng.data(), coo_rng.size() * 4, hipMemcpyHostToDevice,
                                     st));
        // launch order, when COO units of heavy strips hold a quarter of the
        // entries or more: the tiles, then the units by their first block
        // column (units running together cover nearby columns, like the
        // tiles' common sweep), then the single-strip items (R-MAT scale 22,
        // s = 256: 0.654 -> 0.571 ms; where the units are mostly dense /
        // bitmap blocks the order only loses the XCD grouping of the default
        // numbering: block_mix s = 128 1.034 -> 1.103 ms, hence the threshold)
        uint64_t coo_unit_work = 0;
        for (int64_t u = 0; u < n_units; u++)
            if ((units[(size_t)u * 4] & 3u) == (uint32_t)COO) coo_unit_work += unit_work[(size_t)u];
        if (n_units > 0 && 4 * coo_unit_work >= (uint64_t)nnz) {
            DevBuf dcol((size_t)n_units * 4);
            Four<const void *> bcp{{bcol[0].get(), bcol[1].get(), bcol[2].get(), bcol[3].get()}};
            k_unit_cols<<<grid_for((uint64_t)n_units), kT, 0, st>>>(bal_units.get<uint32_t>(), n_units, bcp,
                                                                     bcol_bytes, dcol.get<uint32_t>());
            ABHSF_HIP(hipGetLastError());
            std::vector<uint32_t> ucol((size_t)n_units);
            ABHSF_HIP(hipMemcpyAsync(ucol.data(), dcol.get(), ucol.size() * 4, hipMemcpyDeviceToHost, st));
            ABHSF_HIP(hipStreamSynchronize(st));
            std::vector<uint32_t> uo((size_t)n_units);
            for (size_t u = 0; u < uo.size(); u++) uo[u] = (uint32_t)u;
            std::stable_sort(uo.begin(), uo.end(), [&](uint32_t a, uint32_t b) { return ucol[a] < ucol[b]; });
            int64_t ntiles = 0;
            while (ntiles < n_light && pack[(size_t)ntiles + 1] - pack[(size_t)ntiles] > 1) ntiles++;
            std::vector<uint32_t> ord;
            ord.reserve((size_t)(n_light + n_units + 3));
            for (int64_t i = 0; i < ntiles; i++) ord.push_back((uint32_t)i);
            for (uint32_t u : uo) ord.push_back((uint32_t)(n_light + u));
            for (int64_t i = ntiles; i < n_light; i++) ord.push_back((uint32_t)i);
            while (ord.size() % kStripWaves) ord.push_back(kNoItem);
            bal_order.alloc(ord.size() * 4);
            ABHSF_HIP(hipMemcpyAsync(bal_order.get(), ord.data(), ord.size() * 4, hipMemcpyHostToDevice, st));
            n_slots = (int64_t)ord.size();
        }
        ABHSF_HIP(hipStreamSynchronize(st));
        tiled = true;
    }
}

uint64_t Storage::value_bytes() const
{
    uint64_t b = 0;
    for (int x = 0; x < 4; x++) b += val[x].bytes();
    return b;
}

uint64_t Storage::stored_bytes() const
{
    uint64_t b = value_bytes() + coo_idx.bytes() + csr_rowptr.bytes() + csr_col.bytes() + bm_words.bytes() +
                 bal_light.bytes() + bal_pack.bytes() + bal_units.bytes() + bal_heavy.bytes() +
                 bal_order.bytes() + bal_coo_rng.bytes();
    for (int x = 0; x < 4; x++) b += sptr[x].bytes() + eptr[x].bytes() + bcol[x].bytes() + z1[x].bytes();
    return b;
}

uint32_t csr_triangle_flags(const int64_t *rowptr, const int32_t *col, int64_t m, hipStream_t st)
{
    DevBuf f(sizeof(uint32_t));
    ABHSF_HIP(hipMemsetAsync(f.get(), 0, 4, st));
    k_triangle<<<grid_for((uint64_t)m), kT, 0, st>>>(rowptr, col, m, f.get<uint32_t>());
    ABHSF_HIP(hipGetLastError());
    return read_dev(f.get<uint32_t>(), st);
}

void build_from_device_csr(Storage &S, const int64_t *rowptr, const int32_t *col, const double *val, hipStream_t st)
{
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t nnz = (uint64_t)S.nnz;
    const int64_t m = S.m, M = S.M;
    ABHSF_REQUIRE(nnz < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "nnz must be below 2^32 per matrix");
    const int jbits = bits_for((uint64_t)(S.N > 0 ? S.N - 1 : 0));
    const int ibits = bits_for((uint64_t)(M > 0 ? M - 1 : 0));
    ABHSF_REQUIRE(jbits + ibits <= 64, ABHSF_ERR_UNSUPPORTED, "block key does not fit 64 bits");

    DevBuf err(sizeof(uint32_t));
    ABHSF_HIP(hipMemsetAsync(err.get(), 0, sizeof(uint32_t), st));
    k_check_rowptr<<<grid_for((uint64_t)m + 1), kT, 0, st>>>(rowptr, m, (int64_t)nnz, err.get<uint32_t>());
    ABHSF_HIP(hipGetLastError());

    DevBuf rows(nnz * 4), key(nnz * 8), perm(nnz * 4), flag(nnz * 4), excl(nnz * 4);
    if (nnz) {
        k_rows_keys<<<grid_for(nnz), kT, 0, st>>>(rowptr, col, m, S.n, nnz, S.k, jbits, rows.get<uint32_t>(),
                                                  key.get<uint64_t>(), perm.get<uint32_t>(), err.get<uint32_t>());
        ABHSF_HIP(hipGetLastError());
    }
    const uint32_t ecode = read_dev(err.get<uint32_t>(), st);
    ABHSF_REQUIRE(ecode == 0, ABHSF_ERR_INVALID,
                  std::string("malformed CSR input (") + ((ecode & 7) ? "row pointers" : "") +
                      ((ecode & 8) ? " column out of range" : "") +
                      ((ecode & 16) ? " columns not strictly increasing within a row" : "") + ")");

    radix_sort_pairs(key.get<uint64_t>(), perm.get<uint32_t>(), nnz, jbits + ibits, st);

    uint32_t nblk = 0;
    DevBuf d_nblk(sizeof(uint32_t));
    if (nnz) {
        k_heads<<<grid_for(nnz), kT, 0, st>>>(key.get<uint64_t>(), nnz, flag.get<uint32_t>());
        ABHSF_HIP(hipGetLastError());
        exclusive_scan_u32(flag.get<uint32_t>(), excl.get<uint32_t>(), nnz, d_nblk.get<uint32_t>(), st);
        nblk = read_dev(d_nblk.get<uint32_t>(), st);
    }

    DevBuf bstart(((uint64_t)nblk + 1) * 4), bkey((uint64_t)nblk * 8), sch((uint64_t)nblk);
    DevBuf cnt[4], ecnt[4];
    for (int x = 0; x < 4; x++) {
        cnt[x].alloc((uint64_t)nblk * 4);
        ecnt[x].alloc((uint64_t)nblk * 4);
    }
    DevBuf totals(6 * sizeof(unsigned long long));
    ABHSF_HIP(hipMemsetAsync(totals.get(), 0, totals.bytes(), st));
    unsigned long long htot[6] = {0, 0, 0, 0, 0, 0};
    const uint64_t nwg = grid_for(nblk);
    DevBuf partial(6 * nwg * sizeof(unsigned long long));
    if (nnz) {
        k_block_starts<<<grid_for(nnz), kT, 0, st>>>(key.get<uint64_t>(), flag.get<uint32_t>(),
                                                     excl.get<uint32_t>(), nnz, nblk, bstart.get<uint32_t>(),
                                                     bkey.get<uint64_t>());
        ABHSF_HIP(hipGetLastError());
        Four<uint32_t *> c4, e4;
        for (int x = 0; x < 4; x++) { c4.p[x] = cnt[x].get<uint32_t>(); e4.p[x] = ecnt[x].get<uint32_t>(); }
        const uint64_t hdr = ceil_log2_u64((uint64_t)M) + ceil_log2_u64((uint64_t)S.N) + 2;  // test.cpp:49
        k_block_scheme<<<(unsigned)nwg, kT, 0, st>>>(bstart.get<uint32_t>(), nblk, S.s, S.policy, S.cost_b,
                                                     S.is_binary, hdr, sch.get<uint8_t>(), c4, e4,
                                                     partial.get<unsigned long long>());
        ABHSF_HIP(hipGetLastError());
        for (int q = 0; q < 6; q++) {
            uint64_t *pq = partial.get<uint64_t>() + (uint64_t)q * nwg;
            exclusive_scan_u64(pq, pq, nwg, totals.get<uint64_t>() + q, st);
        }
        ABHSF_HIP(hipMemcpyAsync(htot, totals.get(), sizeof(htot), hipMemcpyDeviceToHost, st));
        ABHSF_HIP(hipStreamSynchronize(st));
    }
    for (int x = 0; x < 4; x++)
        ABHSF_REQUIRE(htot[2 + x] < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "stored values of one scheme exceed 2^32");
    S.smf_preproc = htot[0];
    // block_statistics/mmf.cpp:68: top-level block CSR of the adaptive format
    S.smf_msmf = htot[1] + (uint64_t)M * ceil_log2_u64((uint64_t)S.N + 1) + (uint64_t)nblk * ceil_log2_u64((uint64_t)S.N);

    // per-scheme ordinals and value offsets (in place)
    DevBuf tot4(8 * sizeof(uint32_t));
    ABHSF_HIP(hipMemsetAsync(tot4.get(), 0, tot4.bytes(), st));
    for (int x = 0; x < 4; x++) {
        exclusive_scan_u32(cnt[x].get<uint32_t>(), cnt[x].get<uint32_t>(), nblk, tot4.get<uint32_t>() + x, st);
        exclusive_scan_u32(ecnt[x].get<uint32_t>(), ecnt[x].get<uint32_t>(), nblk, tot4.get<uint32_t>() + 4 + x, st);
    }
    uint32_t h4[8];
    ABHSF_HIP(hipMemcpyAsync(h4, tot4.get(), sizeof(h4), hipMemcpyDeviceToHost, st));
    ABHSF_HIP(hipStreamSynchronize(st));
    for (int x = 0; x < 4; x++) {
        S.nb[x] = h4[x];
        S.ne[x] = h4[4 + x];
    }

    // final storage
    const int vb = S.val_bytes;
    for (int x = 0; x < 4; x++) {
        S.sptr[x].alloc(((uint64_t)M + 1) * 4);
        S.eptr[x].alloc(((uint64_t)M + 1) * 4);
        S.bcol[x].alloc(S.nb[x] * (uint64_t)S.bcol_bytes);
        S.z1[x].alloc(S.nb[x] * 2);
        S.val[x].alloc(S.ne[x] * (uint64_t)vb);
    }
    S.coo_idx.alloc(S.ne[COO] * (uint64_t)S.coo_idx_bytes);
    S.csr_rowptr.alloc(S.nb[CSR] * (uint64_t)(S.s + 1) * (uint64_t)S.csr_rp_bytes);
    S.csr_col.alloc(S.ne[CSR]);
    S.bm_words.alloc(S.nb[BITMAP] * (uint64_t)S.W * 8);
    if (S.bm_words.bytes()) ABHSF_HIP(hipMemsetAsync(S.bm_words.get(), 0, S.bm_words.bytes(), st));
    if (S.val[DENSE].bytes()) ABHSF_HIP(hipMemsetAsync(S.val[DENSE].get(), 0, S.val[DENSE].bytes(), st));

    Four<const uint32_t *> ord, eoff;
    Four<uint32_t> nbx, nex;
    Four<uint32_t *> sp, ep;
    Four<void *> bc;
    Four<uint16_t *> zz;
    for (int x = 0; x < 4; x++) {
        ord.p[x] = cnt[x].get<uint32_t>();
        eoff.p[x] = ecnt[x].get<uint32_t>();
        nbx.p[x] = (uint32_t)S.nb[x];
        nex.p[x] = (uint32_t)S.ne[x];
        sp.p[x] = S.sptr[x].get<uint32_t>();
        ep.p[x] = S.eptr[x].get<uint32_t>();
        bc.p[x] = S.bcol[x].get();
        zz.p[x] = S.z1[x].get<uint16_t>();
    }
    if (nblk == 0) {
        for (int x = 0; x < 4; x++) {
            ABHSF_HIP(hipMemsetAsync(S.sptr[x].get(), 0, S.sptr[x].bytes(), st));
            ABHSF_HIP(hipMemsetAsync(S.eptr[x].get(), 0, S.eptr[x].bytes(), st));
        }
    } else {
        k_strip_ptrs<<<grid_for(nblk), kT, 0, st>>>(bkey.get<uint64_t>(), nblk, jbits, M, ord, eoff, nbx, nex, sp, ep);
        ABHSF_HIP(hipGetLastError());
        const uint64_t jmask = (jbits >= 64) ? ~0ull : ((1ull << jbits) - 1);
        k_block_meta<<<grid_for(nblk), kT, 0, st>>>(bkey.get<uint64_t>(), bstart.get<uint32_t>(), sch.get<uint8_t>(),
                                                    nblk, jmask, S.bcol_bytes, ord, bc, zz);
        ABHSF_HIP(hipGetLastError());
        EmitOut out;
        for (int x = 0; x < 4; x++) out.val[x] = S.val[x].get();
        out.coo_idx = S.coo_idx.get();
        out.csr_rowptr = S.csr_rowptr.get();
        out.csr_col = S.csr_col.get<uint8_t>();
        out.bm_words = S.bm_words.get<unsigned long long>();
        out.coo_idx_bytes = S.coo_idx_bytes;
        out.csr_rp_bytes = S.csr_rp_bytes;
        out.val_bytes = vb;
        out.W = S.W;
        k_emit<<<grid_for(nnz), kT, 0, st>>>(flag.get<uint32_t>(), excl.get<uint32_t>(), bstart.get<uint32_t>(),
                                             sch.get<uint8_t>(), perm.get<uint32_t>(), rows.get<uint32_t>(), col, val,
                                             nnz, S.s, S.k, ord, eoff, out);
        ABHSF_HIP(hipGetLastError());
    }
    ABHSF_HIP(hipStreamSynchronize(st));
    S.built = true;
    S.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace abhsf
