// capi.hip -- the C ABI of libabhsf.so (declared in include/abhsf_c.h).
#include "delta.hpp"
#include "storage.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

extern "C" abhsf_status abhsf_block_hist_impl(const int64_t *rowptr, const int32_t *col, int64_t m, int64_t n,
                                              int64_t nnz, int32_t on_device, int32_t kr, int32_t ks,
                                              int64_t strip_stride, uint64_t *hist, int64_t hist_cap,
                                              uint64_t *nblocks, uint64_t *sampled_nnz, void *stream,
                                              std::string *err);

extern "C" abhsf_status abhsf_mlp_forward_impl(const float *theta, int32_t F, const float *X, int64_t B, float *H1,
                                               float *H2, float *out, void *stream, std::string *err);
extern "C" abhsf_status abhsf_mlp_backward_impl(const float *theta, int32_t F, const float *X, int64_t B,
                                                const float *H1, const float *H2, const float *out, const float *y,
                                                const float *w, float wsum, float *D1, float *D2, float *D3,
                                                float *grad, void *stream, std::string *err);
extern "C" abhsf_status abhsf_mlp_train_impl(float *theta, int32_t F, const float *X, const float *y, const float *w,
                                             int64_t B, int64_t steps, float lr, float *loss_host, void *stream,
                                             std::string *err);
extern "C" abhsf_status abhsf_mlp_predict_impl(const float *thetas, int32_t nm, int32_t F, const float *X,
                                               int64_t B, double *out, void *stream, std::string *err);
extern "C" abhsf_status abhsf_mlp_train_noalloc_impl(float *theta, int32_t F, const float *X, const float *y,
                                                     const float *w, int64_t B, int64_t steps, float lr,
                                                     float *loss_dev, void *stream, std::string *err);

struct abhsf_handle {
    abhsf_options opts;
    hipStream_t stream;
    abhsf::Storage S;
    abhsf::DevBuf col_x, col_y;  // tiled SpMM: one right-hand side at a time, made contiguous
    int32_t symmetry = ABHSF_GENERAL;  // applies to the next build
    int32_t built_symmetry = ABHSF_GENERAL;
};

struct abhsf_delta {
    hipStream_t stream = nullptr;
    int32_t codec = ABHSF_DELTA_VARINT;
    abhsf::DeltaStorage D;
};

struct abhsf_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<abhsf::BatchEntry> host;
    std::vector<const abhsf_handle *> mats;  // members (their storage must outlive the batch)
    abhsf::DevBuf table, heavy, item_mat;
    int64_t total_items = 0, items_f64 = 0, nheavy = 0;
    bool vectors_set = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double galpha = 0.0, gbeta = 0.0;
};

namespace {

thread_local std::string g_last_error;

abhsf_status fail(abhsf_status code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

template <typename F>
abhsf_status guarded(F &&f)
{
    try {
        f();
        g_last_error.clear();
        return ABHSF_OK;
    } catch (const abhsf::Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(ABHSF_ERR_ALLOC, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(ABHSF_ERR_INVALID, e.what());
    }
}

bool is_pow2_block(int s) { return s >= 2 && s <= 256 && (s & (s - 1)) == 0; }

void reset_storage(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz)
{
    h->S.~Storage();
    new (&h->S) abhsf::Storage();
    abhsf::Storage &S = h->S;
    S.set_shape(m, n, h->opts.block_size);
    S.nnz = nnz;
    S.policy = h->opts.policy;
    S.is_binary = h->opts.is_binary;
    S.val_bytes = h->opts.precision == ABHSF_F32 ? 4 : 8;
    // value bits of the cost model: test.cpp:581-593 (REAL -> 64, BINARY -> 0),
    // mmf.cpp:34-47 (single precision -> 32)
    S.cost_b = h->opts.is_binary ? 0 : (uint64_t)S.val_bytes * 8;
}

// Work decomposition for the single-RHS kernel: units of about 2048 entries
// (at least 8 rows' worth), heavy strips above twice that.
#ifndef ABHSF_UNIT_ENTRIES
#define ABHSF_UNIT_ENTRIES 2048
#endif
int64_t balance_unit(const abhsf::Storage &S) { return std::max<int64_t>(ABHSF_UNIT_ENTRIES, 8 * (int64_t)S.s); }

int64_t max_strip_work(const abhsf::Storage &S, hipStream_t st)
{
    if (S.M == 0) return 0;
    std::vector<uint32_t> ep((size_t)S.M + 1);
    std::vector<int64_t> w((size_t)S.M, 0);
    for (int x = 0; x < 4; x++) {
        ABHSF_HIP(hipMemcpyAsync(ep.data(), S.eptr[x].get(), ep.size() * 4, hipMemcpyDeviceToHost, st));
        ABHSF_HIP(hipStreamSynchronize(st));
        for (int64_t I = 0; I < S.M; I++) w[(size_t)I] += (int64_t)ep[(size_t)I + 1] - ep[(size_t)I];
    }
    return *std::max_element(w.begin(), w.end());
}

void finish_build(abhsf_handle *h)
{
    h->S.max_strip_work = max_strip_work(h->S, h->stream);
    const int tile_rows =
        h->opts.variant != ABHSF_VARIANT_TILED ? 0 : (h->S.val_bytes == 4 ? abhsf::kTileRowsComp : abhsf::kTileRows);
    if (h->opts.variant != ABHSF_VARIANT_STRIP)
        h->S.build_balance(balance_unit(h->S), h->stream,
                           h->opts.variant == ABHSF_VARIANT_PANEL ? abhsf::kXcdCount : 1, tile_rows);
}

void check_symmetric_shape(const abhsf_handle *h, int64_t m, int64_t n)
{
    if (h->symmetry == ABHSF_GENERAL) return;
    ABHSF_REQUIRE(h->opts.variant != ABHSF_VARIANT_TILED, ABHSF_ERR_UNSUPPORTED,
                  "symmetric storage needs strip-ordered COO blocks (not the tiled variant)");
    ABHSF_REQUIRE(m == n, ABHSF_ERR_INVALID, "symmetric storage needs a square matrix");
    ABHSF_REQUIRE(h->opts.precision == ABHSF_F64, ABHSF_ERR_UNSUPPORTED, "symmetric storage needs FP64 values");
}

// Symmetric storage: the stored entries must lie in one triangle
// (matrix_market_reader.h:191-197 rejects both), skew-symmetric has no
// diagonal; the CSR arrays are valid (the build checked them).
void check_symmetry(abhsf_handle *h, const int64_t *d_rowptr, const int32_t *d_col)
{
    h->built_symmetry = ABHSF_GENERAL;
    if (h->symmetry == ABHSF_GENERAL) return;
    const uint32_t f = abhsf::csr_triangle_flags(d_rowptr, d_col, h->S.m, h->stream);
    if ((f & 3u) == 3u) {
        h->S.built = false;
        throw abhsf::Error(ABHSF_ERR_INVALID, "Elements from both L and U parts stored for not-unsymmetric matrix.");
    }
    if (h->symmetry == ABHSF_SKEW_SYMMETRIC && (f & 4u)) {
        h->S.built = false;
        throw abhsf::Error(ABHSF_ERR_INVALID, "skew-symmetric matrix with a diagonal entry");
    }
    h->built_symmetry = h->symmetry;
}

void check_shape(int64_t m, int64_t n, int64_t nnz)
{
    ABHSF_REQUIRE(m >= 1 && n >= 1 && m <= INT32_MAX && n <= INT32_MAX, ABHSF_ERR_INVALID,
                  "matrix dimensions must be in [1, 2^31-1]");
    ABHSF_REQUIRE(nnz >= 0, ABHSF_ERR_INVALID, "nnz must be non-negative");
    ABHSF_REQUIRE((uint64_t)nnz < (1ull << 32), ABHSF_ERR_UNSUPPORTED, "nnz must be below 2^32 per matrix");
}

}  // namespace

extern "C" {

const char *abhsf_last_error(void) { return g_last_error.c_str(); }

const char *abhsf_version(void) { return "abhsf-mi355x 0.1 (gfx950)"; }

abhsf_status abhsf_create(const abhsf_options *opts, abhsf_handle **out)
{
    return guarded([&] {
        ABHSF_REQUIRE(opts && out, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(is_pow2_block(opts->block_size), ABHSF_ERR_INVALID, "block_size must be a power of two in [2, 256]");
        ABHSF_REQUIRE(opts->policy == ABHSF_POLICY_PREPROC || opts->policy == ABHSF_POLICY_MSMF, ABHSF_ERR_INVALID,
                      "unknown policy");
        ABHSF_REQUIRE(opts->precision == ABHSF_F64 || opts->precision == ABHSF_F32, ABHSF_ERR_INVALID,
                      "unknown precision");
        ABHSF_REQUIRE(opts->variant >= ABHSF_VARIANT_AUTO && opts->variant <= ABHSF_VARIANT_TILED, ABHSF_ERR_INVALID,
                      "unknown variant");
        *out = nullptr;
        abhsf_handle *h = new abhsf_handle();
        h->opts = *opts;
        h->stream = static_cast<hipStream_t>(opts->stream);
        *out = h;
    });
}

abhsf_status abhsf_destroy(abhsf_handle *h)
{
    return guarded([&] {
        if (!h) return;
        (void)hipStreamSynchronize(h->stream);
        delete h;
    });
}

abhsf_status abhsf_build_csr(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz, const int64_t *rowptr,
                             const int32_t *col, const double *val, int32_t on_device)
{
    return guarded([&] {
        ABHSF_REQUIRE(h && rowptr, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(nnz == 0 || col, ABHSF_ERR_INVALID, "null column array");
        check_shape(m, n, nnz);
        check_symmetric_shape(h, m, n);
        reset_storage(h, m, n, nnz);
        if (on_device) {
            abhsf::build_from_device_csr(h->S, rowptr, col, val, h->stream);
            check_symmetry(h, rowptr, col);
            finish_build(h);
            return;
        }
        ABHSF_REQUIRE(rowptr[0] == 0 && rowptr[m] == nnz, ABHSF_ERR_INVALID, "row pointers do not span nnz");
        abhsf::DevBuf drp(((uint64_t)m + 1) * 8), dc((uint64_t)nnz * 4), dv(val ? (uint64_t)nnz * 8 : 0);
        ABHSF_HIP(hipMemcpyAsync(drp.get(), rowptr, drp.bytes(), hipMemcpyHostToDevice, h->stream));
        if (nnz) ABHSF_HIP(hipMemcpyAsync(dc.get(), col, dc.bytes(), hipMemcpyHostToDevice, h->stream));
        if (val && nnz) ABHSF_HIP(hipMemcpyAsync(dv.get(), val, dv.bytes(), hipMemcpyHostToDevice, h->stream));
        abhsf::build_from_device_csr(h->S, drp.get<int64_t>(), dc.get<int32_t>(), val ? dv.get<double>() : nullptr,
                                     h->stream);
        check_symmetry(h, drp.get<int64_t>(), dc.get<int32_t>());
        finish_build(h);
    });
}

abhsf_status abhsf_build_coo(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz, const int32_t *rows,
                             const int32_t *cols, const double *vals)
{
    return guarded([&] {
        ABHSF_REQUIRE(h, ABHSF_ERR_INVALID, "null handle");
        ABHSF_REQUIRE(nnz == 0 || (rows && cols), ABHSF_ERR_INVALID, "null triplet array");
        check_shape(m, n, nnz);
        check_symmetric_shape(h, m, n);
        for (int64_t i = 0; i < nnz; i++)
            ABHSF_REQUIRE(rows[i] >= 0 && rows[i] < m && cols[i] >= 0 && cols[i] < n, ABHSF_ERR_INVALID,
                          "triplet index out of range");
        // private copy sorted row-major (the reference sorts the caller's
        // vector in place, spmv/unified/spmv.cpp:77)
        std::vector<uint64_t> order((size_t)nnz);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
            return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
        });
        std::vector<int64_t> rp((size_t)m + 1, 0);
        std::vector<int32_t> c((size_t)nnz);
        std::vector<double> v(vals ? (size_t)nnz : 0);
        for (int64_t i = 0; i < nnz; i++) {
            const uint64_t o = order[(size_t)i];
            if (i > 0) {
                const uint64_t p = order[(size_t)i - 1];
                ABHSF_REQUIRE(!(rows[o] == rows[p] && cols[o] == cols[p]), ABHSF_ERR_INVALID, "duplicate triplet");
            }
            rp[(size_t)rows[o] + 1]++;
            c[(size_t)i] = cols[o];
            if (vals) v[(size_t)i] = vals[o];
        }
        for (int64_t r = 0; r < m; r++) rp[(size_t)r + 1] += rp[(size_t)r];
        reset_storage(h, m, n, nnz);
        abhsf::DevBuf drp(((uint64_t)m + 1) * 8), dc((uint64_t)nnz * 4), dv(vals ? (uint64_t)nnz * 8 : 0);
        ABHSF_HIP(hipMemcpyAsync(drp.get(), rp.data(), drp.bytes(), hipMemcpyHostToDevice, h->stream));
        if (nnz) ABHSF_HIP(hipMemcpyAsync(dc.get(), c.data(), dc.bytes(), hipMemcpyHostToDevice, h->stream));
        if (vals && nnz) ABHSF_HIP(hipMemcpyAsync(dv.get(), v.data(), dv.bytes(), hipMemcpyHostToDevice, h->stream));
        abhsf::build_from_device_csr(h->S, drp.get<int64_t>(), dc.get<int32_t>(), vals ? dv.get<double>() : nullptr,
                                     h->stream);
        check_symmetry(h, drp.get<int64_t>(), dc.get<int32_t>());
        finish_build(h);
    });
}

abhsf_status abhsf_import(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz, int32_t count,
                          const char *const *names, const void *const *data, const size_t *bytes)
{
    return guarded([&] {
        ABHSF_REQUIRE(h && names && data && bytes && count > 0, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(h->symmetry == ABHSF_GENERAL, ABHSF_ERR_UNSUPPORTED,
                      "a serialised blob holds general storage (build symmetric storage from CSR)");
        check_shape(m, n, nnz);
        reset_storage(h, m, n, nnz);
        abhsf::import_storage(h->S, count, names, data, bytes, h->stream);
        finish_build(h);
    });
}

abhsf_status abhsf_get_stats(const abhsf_handle *h, abhsf_stats *out)
{
    return guarded([&] {
        ABHSF_REQUIRE(h && out, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(h->S.built, ABHSF_ERR_STATE, "matrix not built");
        const abhsf::Storage &S = h->S;
        std::memset(out, 0, sizeof(*out));
        out->m = S.m; out->n = S.n; out->nnz = S.nnz;
        out->s = S.s; out->k = S.k; out->M = S.M; out->N = S.N;
        for (int x = 0; x < 4; x++) { out->blocks[x] = S.nb[x]; out->values[x] = S.ne[x]; }
        out->smf_bits_preproc = S.smf_preproc;
        out->smf_bits_msmf = S.smf_msmf;
        out->stored_bytes = S.stored_bytes();
        out->index_bytes = S.stored_bytes() - S.value_bytes();
        out->build_ms = S.build_ms;
        out->variant = h->opts.variant == ABHSF_VARIANT_AUTO ? ABHSF_VARIANT_BALANCED : h->opts.variant;
        out->tile_zs1 = S.tile_zs1 ? 1 : 0;
        out->work_items = h->opts.variant == ABHSF_VARIANT_STRIP ? S.M : S.n_light + S.n_units;
        out->max_strip_work = S.max_strip_work;
    });
}

abhsf_status abhsf_spmm(abhsf_handle *h, int32_t nrhs, double alpha, const double *X, int64_t ldx, double beta,
                        double *Y, int64_t ldy)
{
    return guarded([&] {
        ABHSF_REQUIRE(h && X && Y, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(h->S.built, ABHSF_ERR_STATE, "matrix not built");
        ABHSF_REQUIRE(nrhs >= 1 && nrhs <= 64, ABHSF_ERR_UNSUPPORTED, "nrhs must be in [1, 64]");
        ABHSF_REQUIRE(ldx >= nrhs && ldy >= nrhs, ABHSF_ERR_INVALID, "leading dimension smaller than nrhs");
        // The SpMM kernel gives each strip to one wave: where one strip holds
        // more than 1/4096 of the entries (power-law rows: R-MAT scale 22's
        // first strip holds 1.2e6 of 6.5e7), that wave outlasts the rest of
        // the launch (k = 8: 37.6 ms) and each right-hand side goes through
        // the SpMV kernels instead, whose units split heavy strips (5.8 ms).
        const bool serial_strip = h->S.max_strip_work > std::max<int64_t>(4096, h->S.nnz / 4096);
        if ((h->S.tiled || serial_strip) && !(nrhs == 1 && ldx == 1 && ldy == 1)) {
            // tiles reorder the COO blocks, which the SpMM kernels walk strip by
            // strip: each right-hand side goes through the (tiled) SpMV kernels,
            // gathered into and scattered from contiguous columns the handle
            // keeps (allocated at the first such call)
            // (blocks of up to col_block() columns move through a transposed tile)
            const int cbw = abhsf::col_block();
            const int64_t n = h->S.n, m = h->S.m;
            if (h->col_x.bytes() < (size_t)n * 8 * cbw) h->col_x.alloc((size_t)n * 8 * cbw);
            if (h->col_y.bytes() < (size_t)m * 8 * cbw) h->col_y.alloc((size_t)m * 8 * cbw);
            double *cx = h->col_x.get<double>(), *cy = h->col_y.get<double>();
            for (int32_t j0 = 0; j0 < nrhs; j0 += cbw) {
                const int cols = std::min<int32_t>(cbw, nrhs - j0);
                abhsf::gather_cols(X + j0, ldx, n, cols, cx, h->stream);
                if (beta != 0.0) abhsf::gather_cols(Y + j0, ldy, m, cols, cy, h->stream);
                for (int c = 0; c < cols; c++) {
                    double *xc = cx + (int64_t)c * n, *yc = cy + (int64_t)c * m;
                    abhsf::spmm(h->S.launch(), h->S.view(), 1, alpha, xc, 1, beta, yc, 1, h->stream);
                    if (h->built_symmetry != ABHSF_GENERAL && alpha != 0.0)  // the mirrored triangle
                        abhsf::spmv_transposed(h->S.launch(), h->S.view(), 1,
                                               h->built_symmetry == ABHSF_SKEW_SYMMETRIC ? -alpha : alpha, xc, 1,
                                               yc, 1, h->stream);
                }
                abhsf::scatter_cols(cy, m, cols, Y + j0, ldy, h->stream);
            }
            return;
        }
        abhsf::spmm(h->S.launch(), h->S.view(), nrhs, alpha, X, ldx, beta, Y, ldy, h->stream);
        if (h->built_symmetry != ABHSF_GENERAL && alpha != 0.0)  // the mirrored triangle
            abhsf::spmv_transposed(h->S.launch(), h->S.view(), nrhs,
                                   h->built_symmetry == ABHSF_SKEW_SYMMETRIC ? -alpha : alpha, X, ldx, Y, ldy,
                                   h->stream);
    });
}

abhsf_status abhsf_set_symmetry(abhsf_handle *h, int32_t symmetry)
{
    return guarded([&] {
        ABHSF_REQUIRE(h, ABHSF_ERR_INVALID, "null handle");
        ABHSF_REQUIRE(symmetry == ABHSF_GENERAL || symmetry == ABHSF_SYMMETRIC || symmetry == ABHSF_SKEW_SYMMETRIC,
                      ABHSF_ERR_INVALID, "unknown symmetry");
        h->symmetry = symmetry;
    });
}

abhsf_status abhsf_spmv(abhsf_handle *h, double alpha, const double *x, double beta, double *y)
{
    return abhsf_spmm(h, 1, alpha, x, 1, beta, y, 1);
}

abhsf_status abhsf_export(const abhsf_handle *h, const char *name, void *dst, size_t cap, size_t *bytes)
{
    return guarded([&] {
        ABHSF_REQUIRE(h && name, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(h->S.built, ABHSF_ERR_STATE, "matrix not built");
        ABHSF_REQUIRE(!h->S.tiled, ABHSF_ERR_UNSUPPORTED,
                      "a tiled matrix cannot be exported (its COO blocks are in tile order)");
        const abhsf::Storage &S = h->S;
        static const char *pre[4] = {"coo_", "csr_", "bm_", "dn_"};
        const abhsf::DevBuf *buf = nullptr;
        const std::string nm(name);
        for (int x = 0; x < 4 && !buf; x++) {
            if (nm == std::string(pre[x]) + "sptr") buf = &S.sptr[x];
            else if (nm == std::string(pre[x]) + "eptr") buf = &S.eptr[x];
            else if (nm == std::string(pre[x]) + "bcol") buf = &S.bcol[x];
            else if (nm == std::string(pre[x]) + "z1") buf = &S.z1[x];
            else if (nm == std::string(pre[x]) + "val") buf = &S.val[x];
        }
        if (!buf) {
            if (nm == "coo_idx") buf = &S.coo_idx;
            else if (nm == "csr_rowptr") buf = &S.csr_rowptr;
            else if (nm == "csr_col") buf = &S.csr_col;
            else if (nm == "bm_words") buf = &S.bm_words;
        }
        ABHSF_REQUIRE(buf, ABHSF_ERR_INVALID, "unknown array name: " + nm);
        if (bytes) *bytes = buf->bytes();
        if (!dst) return;
        ABHSF_REQUIRE(cap >= buf->bytes(), ABHSF_ERR_INVALID, "destination too small");
        if (buf->bytes()) {
            ABHSF_HIP(hipMemcpyAsync(dst, buf->get(), buf->bytes(), hipMemcpyDeviceToHost, h->stream));
            ABHSF_HIP(hipStreamSynchronize(h->stream));
        }
    });
}

abhsf_status abhsf_block_hist(const int64_t *rowptr, const int32_t *col, int64_t m, int64_t n, int64_t nnz,
                              int32_t on_device, int32_t kr, int32_t ks, int64_t strip_stride, uint64_t *hist,
                              int64_t hist_cap, uint64_t *nblocks, uint64_t *sampled_nnz, void *stream)
{
    std::string err;
    const abhsf_status st = abhsf_block_hist_impl(rowptr, col, m, n, nnz, on_device, kr, ks, strip_stride, hist,
                                                  hist_cap, nblocks, sampled_nnz, stream, &err);
    if (st != ABHSF_OK) return fail(st, err);
    g_last_error.clear();
    return ABHSF_OK;
}

abhsf_status abhsf_mlp_forward(const float *theta, int32_t nfeat, const float *X, int64_t nsamples, float *H1,
                               float *H2, float *out, void *stream)
{
    std::string err;
    const abhsf_status st = abhsf_mlp_forward_impl(theta, nfeat, X, nsamples, H1, H2, out, stream, &err);
    if (st != ABHSF_OK) return fail(st, err);
    g_last_error.clear();
    return ABHSF_OK;
}

abhsf_status abhsf_mlp_predict(const float *thetas, int32_t nmembers, int32_t nfeat, const float *X,
                               int64_t nsamples, double *out, void *stream)
{
    std::string err;
    const abhsf_status st = abhsf_mlp_predict_impl(thetas, nmembers, nfeat, X, nsamples, out, stream, &err);
    if (st != ABHSF_OK) return fail(st, err);
    g_last_error.clear();
    return ABHSF_OK;
}

abhsf_status abhsf_mlp_backward(const float *theta, int32_t nfeat, const float *X, int64_t nsamples,
                                const float *H1, const float *H2, const float *out, const float *y,
                                const float *w, float wsum, float *D1, float *D2, float *D3, float *grad,
                                void *stream)
{
    std::string err;
    const abhsf_status st = abhsf_mlp_backward_impl(theta, nfeat, X, nsamples, H1, H2, out, y, w, wsum, D1, D2, D3,
                                                    grad, stream, &err);
    if (st != ABHSF_OK) return fail(st, err);
    g_last_error.clear();
    return ABHSF_OK;
}

abhsf_status abhsf_mlp_train(float *theta, int32_t nfeat, const float *X, const float *y, const float *w,
                             int64_t nsamples, int64_t steps, float lr, float *loss_host, void *stream)
{
    std::string err;
    const abhsf_status st = abhsf_mlp_train_impl(theta, nfeat, X, y, w, nsamples, steps, lr, loss_host, stream, &err);
    if (st != ABHSF_OK) return fail(st, err);
    g_last_error.clear();
    return ABHSF_OK;
}

abhsf_status abhsf_mlp_train_noalloc(float *theta, int32_t nfeat, const float *X, const float *y, const float *w,
                                     int64_t nsamples, int64_t steps, float lr, float *loss, void *stream)
{
    std::string err;
    const abhsf_status st =
        abhsf_mlp_train_noalloc_impl(theta, nfeat, X, y, w, nsamples, steps, lr, loss, stream, &err);
    if (st != ABHSF_OK) return fail(st, err);
    g_last_error.clear();
    return ABHSF_OK;
}

namespace {
// Symmetric members: the mirrored triangle's pass (y_i += +-alpha * T_off^T
// x_i, global atomics) after the batched launch, one launch per such member.
void batch_mirrored(abhsf_batch *b, double alpha)
{
    if (alpha == 0.0) return;
    for (size_t i = 0; i < b->mats.size(); i++) {
        const abhsf_handle *h = b->mats[i];
        if (h->built_symmetry == ABHSF_GENERAL) continue;
        abhsf::spmv_transposed(h->S.launch(), h->S.view(), 1,
                               h->built_symmetry == ABHSF_SKEW_SYMMETRIC ? -alpha : alpha, b->host[i].x, 1,
                               b->host[i].y, 1, b->stream);
    }
}
}  // namespace

abhsf_status abhsf_batch_create(abhsf_handle *const *mats, int32_t count, void *stream, abhsf_batch **out)
{
    return guarded([&] {
        ABHSF_REQUIRE(mats && out && count >= 1, ABHSF_ERR_INVALID, "invalid argument");
        *out = nullptr;
        auto b = std::make_unique<abhsf_batch>();
        if (stream) {
            b->stream = static_cast<hipStream_t>(stream);
        } else {
            ABHSF_HIP(hipStreamCreate(&b->stream));
            b->own_stream = true;
        }
        std::vector<uint32_t> hv;
        for (int32_t i = 0; i < count; i++) {
            const abhsf_handle *h = mats[i];
            ABHSF_REQUIRE(h && h->S.built, ABHSF_ERR_STATE, "batch member not built");
            abhsf::BatchEntry e{};
            e.A = h->S.view();
            e.B = h->S.launch().bal;
            if (!e.B.light && e.B.n_units == 0) e.B.n_light = h->S.M;
            e.code = abhsf::batch_code(h->S.bcol_bytes, h->S.s);
            e.comp = h->S.val_bytes == 4 ? 1 : 0;
            for (int64_t q = 0; q < e.B.n_heavy; q++) hv.insert(hv.end(), {(uint32_t)i, (uint32_t)q});
            b->host.push_back(e);
            b->mats.push_back(h);
        }
        // number the items member by member, FP64 members first (one launch),
        // then FP32 members (a second launch); within each, longest first:
        // members in decreasing order of entries per work item (one kernel
        // serves all index-type combinations; the grid-stride loop starts the
        // items in this order, so the long ones do not form the launch's tail)
        std::vector<size_t> ord(b->host.size());
        std::iota(ord.begin(), ord.end(), (size_t)0);
        auto per_item = [&](size_t i) {
            const auto &e = b->host[i];
            return (double)mats[i]->S.nnz / (double)std::max<int64_t>(1, e.B.n_light + e.B.n_units);
        };
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t c) {
            if (b->host[a].comp != b->host[c].comp) return b->host[a].comp < b->host[c].comp;
            return per_item(a) > per_item(c);
        });
        int64_t item = 0;
        for (size_t i : ord) {
            if (!b->host[i].comp) b->items_f64 = item + b->host[i].B.n_light + b->host[i].B.n_units;
            b->host[i].item0 = item;
            item += b->host[i].B.n_light + b->host[i].B.n_units;
        }
        b->total_items = item;
        b->nheavy = (int64_t)hv.size() / 2;
        ABHSF_REQUIRE(item < (int64_t)UINT32_MAX, ABHSF_ERR_UNSUPPORTED, "too many work items in one batch");
        std::vector<uint32_t> im((size_t)item);
        for (size_t i = 0; i < b->host.size(); i++) {
            const auto &e = b->host[i];
            std::fill(im.begin() + e.item0, im.begin() + e.item0 + e.B.n_light + e.B.n_units, (uint32_t)i);
        }
        b->item_mat.alloc(im.size() * 4);
        if (!im.empty())
            ABHSF_HIP(hipMemcpyAsync(b->item_mat.get(), im.data(), im.size() * 4, hipMemcpyHostToDevice, b->stream));
        b->table.alloc(b->host.size() * sizeof(abhsf::BatchEntry));
        b->heavy.alloc(hv.size() * 4);
        if (!hv.empty())
            ABHSF_HIP(hipMemcpyAsync(b->heavy.get(), hv.data(), hv.size() * 4, hipMemcpyHostToDevice, b->stream));
        ABHSF_HIP(hipStreamSynchronize(b->stream));
        *out = b.release();
    });
}

abhsf_status abhsf_batch_set_vectors(abhsf_batch *b, const double *const *xs, double *const *ys)
{
    return guarded([&] {
        ABHSF_REQUIRE(b && xs && ys, ABHSF_ERR_INVALID, "null argument");
        for (size_t i = 0; i < b->host.size(); i++) {
            ABHSF_REQUIRE(xs[i] && ys[i], ABHSF_ERR_INVALID, "null vector in batch");
            b->host[i].x = xs[i];
            b->host[i].y = ys[i];
        }
        ABHSF_HIP(hipMemcpyAsync(b->table.get(), b->host.data(), b->table.bytes(), hipMemcpyHostToDevice, b->stream));
        ABHSF_HIP(hipStreamSynchronize(b->stream));
        b->vectors_set = true;
    });
}

abhsf_status abhsf_batch_spmv(abhsf_batch *b, double alpha, double beta, int32_t use_graph)
{
    return guarded([&] {
        ABHSF_REQUIRE(b, ABHSF_ERR_INVALID, "null batch");
        ABHSF_REQUIRE(b->vectors_set, ABHSF_ERR_STATE, "abhsf_batch_set_vectors not called");
        auto launch = [&] {
            abhsf::batch_spmv(b->table.get<abhsf::BatchEntry>(), b->item_mat.get<uint32_t>(), b->items_f64,
                              b->total_items, b->heavy.get<uint32_t>(), b->nheavy, alpha, beta, b->stream);
            batch_mirrored(b, alpha);
        };
        if (!use_graph) {
            launch();
            return;
        }
        if (!b->exec || b->galpha != alpha || b->gbeta != beta) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
            ABHSF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            try {
                launch();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            ABHSF_HIP(hipStreamEndCapture(b->stream, &b->graph));
            ABHSF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->galpha = alpha;
            b->gbeta = beta;
        }
        ABHSF_HIP(hipGraphLaunch(b->exec, b->stream));
    });
}

abhsf_status abhsf_batch_synchronize(abhsf_batch *b)
{
    return guarded([&] {
        ABHSF_REQUIRE(b, ABHSF_ERR_INVALID, "null batch");
        ABHSF_HIP(hipStreamSynchronize(b->stream));
    });
}

abhsf_status abhsf_batch_profile(abhsf_batch *b, double alpha, double beta, uint64_t *cycles)
{
    return guarded([&] {
        ABHSF_REQUIRE(b && cycles, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(b->vectors_set, ABHSF_ERR_STATE, "abhsf_batch_set_vectors not called");
        const size_t nm = b->host.size();
        abhsf::DevBuf cyc(nm * sizeof(unsigned long long));
        ABHSF_HIP(hipMemsetAsync(cyc.get(), 0, cyc.bytes(), b->stream));
        abhsf::batch_spmv(b->table.get<abhsf::BatchEntry>(), b->item_mat.get<uint32_t>(), b->items_f64,
                          b->total_items, b->heavy.get<uint32_t>(), b->nheavy, alpha, beta, b->stream,
                          cyc.get<unsigned long long>());
        batch_mirrored(b, alpha);  // results complete; its time is not in the members' cycles
        ABHSF_HIP(hipMemcpyAsync(cycles, cyc.get(), cyc.bytes(), hipMemcpyDeviceToHost, b->stream));
        ABHSF_HIP(hipStreamSynchronize(b->stream));
    });
}

abhsf_status abhsf_batch_destroy(abhsf_batch *b)
{
    return guarded([&] {
        if (!b) return;
        (void)hipStreamSynchronize(b->stream);
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->own_stream) (void)hipStreamDestroy(b->stream);
        delete b;
    });
}

abhsf_status abhsf_delta_create(void *stream, abhsf_delta **out)
{
    return guarded([&] {
        ABHSF_REQUIRE(out, ABHSF_ERR_INVALID, "null argument");
        *out = new abhsf_delta();
        (*out)->stream = static_cast<hipStream_t>(stream);
    });
}

abhsf_status abhsf_delta_create_codec(void *stream, int32_t codec, abhsf_delta **out)
{
    return guarded([&] {
        ABHSF_REQUIRE(out, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(codec == ABHSF_DELTA_VARINT || codec == ABHSF_DELTA_U32_ROW || codec == ABHSF_DELTA_U32_BLOCK,
                      ABHSF_ERR_INVALID, "unknown delta codec");
        *out = new abhsf_delta();
        (*out)->stream = static_cast<hipStream_t>(stream);
        (*out)->codec = codec;
    });
}

abhsf_status abhsf_delta_build_csr(abhsf_delta *d, int64_t m, int64_t n, int64_t nnz, const int64_t *rowptr,
                                   const int32_t *col, const double *val, int32_t on_device)
{
    return guarded([&] {
        ABHSF_REQUIRE(d && rowptr, ABHSF_ERR_INVALID, "null argument");
        check_shape(m, n, nnz);
        if (!on_device) {
            ABHSF_REQUIRE(rowptr[m] == nnz, ABHSF_ERR_INVALID, "row pointers do not span nnz");
            abhsf::delta_build(d->D, m, n, rowptr, col, val, d->stream, d->codec);
            return;
        }
        // the encoder runs on the host: bring the device CSR over
        std::vector<int64_t> rp((size_t)m + 1);
        std::vector<int32_t> c((size_t)nnz);
        std::vector<double> v(val ? (size_t)nnz : 0);
        ABHSF_HIP(hipMemcpyAsync(rp.data(), rowptr, rp.size() * 8, hipMemcpyDeviceToHost, d->stream));
        if (nnz) ABHSF_HIP(hipMemcpyAsync(c.data(), col, c.size() * 4, hipMemcpyDeviceToHost, d->stream));
        if (val && nnz) ABHSF_HIP(hipMemcpyAsync(v.data(), val, v.size() * 8, hipMemcpyDeviceToHost, d->stream));
        ABHSF_HIP(hipStreamSynchronize(d->stream));
        ABHSF_REQUIRE(rp[(size_t)m] == nnz, ABHSF_ERR_INVALID, "row pointers do not span nnz");
        abhsf::delta_build(d->D, m, n, rp.data(), c.data(), val ? v.data() : nullptr, d->stream, d->codec);
    });
}

abhsf_status abhsf_delta_spmv(abhsf_delta *d, double alpha, const double *x, double beta, double *y)
{
    return guarded([&] {
        ABHSF_REQUIRE(d && x && y, ABHSF_ERR_INVALID, "null argument");
        abhsf::delta_spmv(d->D, alpha, x, beta, y, d->stream);
    });
}

abhsf_status abhsf_delta_get_stats(const abhsf_delta *d, abhsf_delta_stats *out)
{
    return guarded([&] {
        ABHSF_REQUIRE(d && out, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(d->D.built, ABHSF_ERR_STATE, "matrix not built");
        std::memset(out, 0, sizeof(*out));
        out->m = d->D.m;
        out->n = d->D.n;
        out->nnz = d->D.nnz;
        out->key_bits = d->D.codec == ABHSF_DELTA_VARINT ? d->D.K : 31;
        out->codec = d->D.codec;
        out->chunk = abhsf::kDeltaChunk;
        for (int t = 0; t < 4; t++) out->words[t] = d->D.count[t];
        out->stored_bytes = d->D.stored_bytes();
        out->build_ms = d->D.build_ms;
    });
}

abhsf_status abhsf_delta_export(const abhsf_delta *d, const char *name, void *dst, size_t cap, size_t *bytes)
{
    return guarded([&] {
        ABHSF_REQUIRE(d && name && bytes, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(d->D.built, ABHSF_ERR_STATE, "matrix not built");
        const std::string nm(name);
        const abhsf::DevBuf *buf = nm == "tags" ? &d->D.tags : nm == "w8" ? &d->D.words[0]
                                   : nm == "w16" ? &d->D.words[1] : nm == "w32" ? &d->D.words[2]
                                   : nm == "w64" ? &d->D.words[3] : nm == "base" ? &d->D.base
                                   : nm == "start" ? &d->D.start : nm == "val" ? &d->D.val
                                   : nm == "da" ? &d->D.da : nm == "xstart" ? &d->D.xstart
                                   : nm == "exc" ? &d->D.exc : nullptr;
        ABHSF_REQUIRE(buf, ABHSF_ERR_INVALID, "unknown array " + nm);
        *bytes = buf->bytes();
        if (!dst) return;
        ABHSF_REQUIRE(cap >= buf->bytes(), ABHSF_ERR_INVALID, "destination too small");
        if (buf->bytes())
            ABHSF_HIP(hipMemcpyAsync(dst, buf->get(), buf->bytes(), hipMemcpyDeviceToHost, d->stream));
        ABHSF_HIP(hipStreamSynchronize(d->stream));
    });
}

abhsf_status abhsf_delta_synchronize(abhsf_delta *d)
{
    return guarded([&] {
        ABHSF_REQUIRE(d, ABHSF_ERR_INVALID, "null argument");
        ABHSF_HIP(hipStreamSynchronize(d->stream));
    });
}

abhsf_status abhsf_delta_destroy(abhsf_delta *d)
{
    return guarded([&] {
        if (!d) return;
        (void)hipStreamSynchronize(d->stream);
        delete d;
    });
}

abhsf_status abhsf_synchronize(abhsf_handle *h)
{
    return guarded([&] {
        ABHSF_REQUIRE(h, ABHSF_ERR_INVALID, "null handle");
        ABHSF_HIP(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
