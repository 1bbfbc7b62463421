// blob.hip -- recreate ABHSF storage from exported arrays (a serialised blob).
//
// Every array is checked on the host before it reaches the device: sizes
// against the shape, strip pointers monotone and consistent with the block
// sizes, block columns and in-block positions inside the matrix, CSR row
// pointers and bitmap popcounts consistent with z.  A blob that passes cannot
// make a kernel read outside its arrays or x / y.
#include "storage.hpp"

#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace abhsf {
namespace {

template <typename T>
std::vector<T> as_vec(const void *p, size_t bytes)
{
    std::vector<T> v(bytes / sizeof(T));
    if (bytes) std::memcpy(v.data(), p, bytes);
    return v;
}

uint64_t get_u(const void *p, int bytes, uint64_t i)
{
    if (bytes == 1) return static_cast<const uint8_t *>(p)[i];
    if (bytes == 2) return static_cast<const uint16_t *>(p)[i];
    return static_cast<const uint32_t *>(p)[i];
}

}  // namespace

void import_storage(Storage &S, int32_t count, const char *const *names, const void *const *data,
                    const size_t *bytes, hipStream_t st)
{
    static const char *pre[4] = {"coo_", "csr_", "bm_", "dn_"};
    std::map<std::string, std::pair<const void *, size_t>> in;
    for (int32_t i = 0; i < count; i++) {
        ABHSF_REQUIRE(names && names[i] && (bytes[i] == 0 || data[i]), ABHSF_ERR_INVALID, "null array in blob");
        in[names[i]] = {data[i], bytes[i]};
    }
    auto get = [&](const std::string &nm) -> std::pair<const void *, size_t> {
        auto it = in.find(nm);
        ABHSF_REQUIRE(it != in.end(), ABHSF_ERR_INVALID, "blob lacks array " + nm);
        return it->second;
    };
    const int64_t M = S.M, N = S.N, s = S.s;
    const uint64_t ss = (uint64_t)s * (uint64_t)s;
    const int vb = S.val_bytes;
    std::vector<uint32_t> sp[4], ep[4];
    for (int x = 0; x < 4; x++) {
        auto a = get(std::string(pre[x]) + "sptr"), b = get(std::string(pre[x]) + "eptr");
        ABHSF_REQUIRE(a.second == ((size_t)M + 1) * 4 && b.second == ((size_t)M + 1) * 4, ABHSF_ERR_INVALID,
                      "strip pointer arrays do not match the shape");
        sp[x] = as_vec<uint32_t>(a.first, a.second);
        ep[x] = as_vec<uint32_t>(b.first, b.second);
        ABHSF_REQUIRE(sp[x][0] == 0 && ep[x][0] == 0, ABHSF_ERR_INVALID, "strip pointers must start at 0");
        for (int64_t I = 0; I < M; I++)
            ABHSF_REQUIRE(sp[x][I] <= sp[x][I + 1] && ep[x][I] <= ep[x][I + 1], ABHSF_ERR_INVALID,
                          "strip pointers not monotone");
        S.nb[x] = sp[x][M];
        S.ne[x] = ep[x][M];
        auto bc = get(std::string(pre[x]) + "bcol"), zz = get(std::string(pre[x]) + "z1"),
             vv = get(std::string(pre[x]) + "val");
        ABHSF_REQUIRE(bc.second == S.nb[x] * (uint64_t)S.bcol_bytes && zz.second == S.nb[x] * 2 &&
                          vv.second == S.ne[x] * (uint64_t)vb,
                      ABHSF_ERR_INVALID, std::string("array sizes of scheme ") + pre[x] + " inconsistent");
    }
    const void *bcol[4], *zp[4];
    for (int x = 0; x < 4; x++) {
        bcol[x] = get(std::string(pre[x]) + "bcol").first;
        zp[x] = get(std::string(pre[x]) + "z1").first;
    }
    auto coo = get("coo_idx"), crp = get("csr_rowptr"), ccol = get("csr_col"), bmw = get("bm_words");
    ABHSF_REQUIRE(coo.second == S.ne[COO] * (uint64_t)S.coo_idx_bytes &&
                      crp.second == S.nb[CSR] * (uint64_t)(s + 1) * (uint64_t)S.csr_rp_bytes &&
                      ccol.second == S.ne[CSR] && bmw.second == S.nb[BITMAP] * (uint64_t)S.W * 8 &&
                      S.ne[DENSE] == S.nb[DENSE] * ss,
                  ABHSF_ERR_INVALID, "in-block array sizes inconsistent");
    // contents: every entry inside the matrix, offsets consistent with z
    uint64_t nnz = 0, smf_pre = 0, smf_msmf = 0;
    const uint64_t hdr = ceil_log2_u64((uint64_t)M) + ceil_log2_u64((uint64_t)N) + 2;
    for (int64_t I = 0; I < M; I++) {
        const int64_t rows = std::min<int64_t>(s, S.m - I * s);
        for (int x = 0; x < 4; x++) {
            uint64_t e = ep[x][I];
            for (uint32_t b = sp[x][I]; b < sp[x][I + 1]; b++) {
                const uint64_t J = get_u(bcol[x], S.bcol_bytes, b);
                const uint64_t z = (uint64_t)static_cast<const uint16_t *>(zp[x])[b] + 1;
                ABHSF_REQUIRE((int64_t)J < N && z <= ss, ABHSF_ERR_INVALID, "block column or size out of range");
                const int64_t cols = std::min<int64_t>(s, S.n - (int64_t)J * s);
                ABHSF_REQUIRE(choose_scheme(S.policy, (uint64_t)s, z, S.cost_b, S.is_binary, nullptr) == x,
                              ABHSF_ERR_INVALID, "block stored in a scheme the cost model does not choose");
                uint64_t bits;
                choose_scheme(ABHSF_POLICY_PREPROC, (uint64_t)s, z, S.cost_b, S.is_binary, &bits);
                smf_pre += hdr + bits;
                choose_scheme(ABHSF_POLICY_MSMF, (uint64_t)s, z, S.cost_b, S.is_binary, &bits);
                smf_msmf += bits + 2;
                nnz += z;
                if (x == COO) {
                    for (uint64_t q = 0; q < z; q++) {
                        const uint64_t id = get_u(coo.first, S.coo_idx_bytes, e + q);
                        ABHSF_REQUIRE((int64_t)(id >> S.k) < rows && (int64_t)(id & (uint64_t)(s - 1)) < cols &&
                                          (id >> S.k) < (uint64_t)s,
                                      ABHSF_ERR_INVALID, "COO entry outside the matrix");
                    }
                } else if (x == CSR) {
                    uint64_t prev = 0;
                    for (int64_t r = 0; r <= s; r++) {
                        const uint64_t v = get_u(crp.first, S.csr_rp_bytes, (uint64_t)b * (s + 1) + r);
                        const bool ok = (r == 0) ? v == 0 : (v >= prev && (r - 1 < rows || v == prev));
                        ABHSF_REQUIRE(ok && v <= z, ABHSF_ERR_INVALID, "CSR row pointers inconsistent");
                        prev = v;
                    }
                    ABHSF_REQUIRE(prev == z, ABHSF_ERR_INVALID, "CSR row pointers do not end at z");
                    for (uint64_t q = 0; q < z; q++)
                        ABHSF_REQUIRE((int64_t)static_cast<const uint8_t *>(ccol.first)[e + q] < cols,
                                      ABHSF_ERR_INVALID, "CSR column outside the matrix");
                } else if (x == BITMAP) {
                    uint64_t pc = 0;
                    for (int64_t w = 0; w < S.W; w++) {
                        const uint64_t word = static_cast<const uint64_t *>(bmw.first)[(uint64_t)b * S.W + w];
                        for (int i = 0; i < 64; i++)
                            if ((word >> i) & 1) {
                                const uint64_t p = (uint64_t)w * 64 + i;
                                ABHSF_REQUIRE(p < ss && (int64_t)(p >> S.k) < rows &&
                                                  (int64_t)(p & (uint64_t)(s - 1)) < cols,
                                              ABHSF_ERR_INVALID, "bitmap bit outside the matrix");
                                pc++;
                            }
                    }
                    ABHSF_REQUIRE(pc == z, ABHSF_ERR_INVALID, "bitmap popcount differs from z");
                }
                e += (x == DENSE) ? ss : z;
            }
            ABHSF_REQUIRE(e == ep[x][I + 1], ABHSF_ERR_INVALID, "value offsets inconsistent with block sizes");
        }
    }
    ABHSF_REQUIRE((int64_t)nnz == S.nnz, ABHSF_ERR_INVALID, "blob holds a different number of entries");
    S.smf_preproc = smf_pre;
    uint64_t nblk = 0;
    for (int x = 0; x < 4; x++) nblk += S.nb[x];
    S.smf_msmf = smf_msmf + (uint64_t)M * ceil_log2_u64((uint64_t)N + 1) + nblk * ceil_log2_u64((uint64_t)N);
    // upload
    auto up = [&](DevBuf &d, const std::string &nm) {
        auto a = get(nm);
        d.alloc(a.second);
        if (a.second) ABHSF_HIP(hipMemcpyAsync(d.get(), a.first, a.second, hipMemcpyHostToDevice, st));
    };
    for (int x = 0; x < 4; x++) {
        up(S.sptr[x], std::string(pre[x]) + "sptr");
        up(S.eptr[x], std::string(pre[x]) + "eptr");
        up(S.bcol[x], std::string(pre[x]) + "bcol");
        up(S.z1[x], std::string(pre[x]) + "z1");
        up(S.val[x], std::string(pre[x]) + "val");
    }
    up(S.coo_idx, "coo_idx");
    up(S.csr_rowptr, "csr_rowptr");
    up(S.csr_col, "csr_col");
    up(S.bm_words, "bm_words");
    ABHSF_HIP(hipStreamSynchronize(st));
    S.built = true;
}

}  // namespace abhsf
