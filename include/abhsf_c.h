/*
 * abhsf_c.h -- C ABI of the MI355X ABHSF SpMV library (libabhsf.so).
 *
 * Adaptive-blocking hierarchical storage format: the matrix is tiled into
 * s x s blocks (s = 2^k, 2 <= s <= 256); every nonzero block is stored in the
 * scheme (COO, CSR, bitmap or dense) that the cost model makes smallest, and the
 * nonzero blocks of each scheme are indexed by a block-level CSR over block
 * rows ("strips").
 *
 * Conventions
 *   - Every call returns abhsf_status; 0 is success.  abhsf_last_error()
 *     gives the message of the last failure on the calling thread.
 *   - The handle owns all device memory it allocates.  One handle must not
 *     be used from two threads at once.
 *   - x / y / X / Y are DEVICE pointers (hipMalloc'd).  Compute calls are
 *     enqueued on the handle's stream and return without synchronising;
 *     abhsf_synchronize() waits.
 *   - The caller's input arrays are never modified (the reference's
 *     csr_matrix::from_elements sorts the caller's vector in place,
 *     spmv/unified/spmv.cpp:77; this library sorts a private copy).
 *
 * Storage layout in HBM (one set of arrays per scheme X in {coo,csr,bm,dn}):
 *   X_sptr[M+1]  u32   block offsets per strip (block-level CSR row pointers)
 *   X_eptr[M+1]  u32   value offsets per strip
 *   X_bcol[nb]   u16 if N <= 65536 else u32   block column J
 *   X_z1[nb]     u16   entries in the block minus one
 *   X_val[ne]    f64 or f32 (ABHSF_F32)
 *   coo_idx[ne]  (lrow << k) | lcol, u8 if s <= 16 else u16
 *   csr_rowptr[nb*(s+1)] u8 if s <= 8 else u16; csr_col[ne] u8 = lcol
 *   bm_words[nb*max(1,s*s/64)] u64, bit lrow*s+lcol LSB-first; values in
 *                set-bit order
 *   dn_val[nb*s*s] row-major, explicit zeros
 * Blocks of a strip are in block-column order inside each scheme group;
 * entries of a block are in (lrow, lcol) order.
 */
#ifndef ABHSF_C_H
#define ABHSF_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t abhsf_status;
enum {
    ABHSF_OK = 0,
    ABHSF_ERR_INVALID = 1,     /* bad argument or malformed matrix */
    ABHSF_ERR_HIP = 2,         /* HIP runtime failure */
    ABHSF_ERR_ALLOC = 3,       /* device or host allocation failed */
    ABHSF_ERR_STATE = 4,       /* call out of order (e.g. spmv before build) */
    ABHSF_ERR_UNSUPPORTED = 5  /* size or option outside what is implemented */
};

/* Scheme-choice cost models.
 * PREPROC: preprocessing/test.cpp:55-77 (abhsf_structure_memory_footprint_processor)
 * MSMF:    include/abhsf/msmf.h:9-48 (block_min_msmf) */
enum { ABHSF_POLICY_PREPROC = 0, ABHSF_POLICY_MSMF = 1 };

/* Value storage.  F64 is the reference-equivalent path (real_type = double,
 * spmv/unified/spmv.cpp:59).  F32 stores values in single precision and
 * accumulates in double with compensated (TwoSum) sums. */
enum { ABHSF_F64 = 0, ABHSF_F32 = 1 };

/* SpMV kernel variants (the tuner chooses among them). */
enum {
    ABHSF_VARIANT_AUTO = 0,
    ABHSF_VARIANT_STRIP = 1,    /* one wavefront per strip */
    ABHSF_VARIANT_BALANCED = 2, /* heavy strips split over several wavefronts */
    ABHSF_VARIANT_PANEL = 3,    /* balanced, plus 8 column panels of equal work, each
                                   processed on one XCD so that its slice of x stays in
                                   that XCD's L2; strips spanning panels are split there */
    ABHSF_VARIANT_TILED = 4     /* balanced, plus runs of consecutive light strips (up to
                                   512 rows, 256 with F32 values) processed as one tile
                                   whose COO blocks are stored in block-column order, so
                                   that all wavefronts sweep x together.  SpMV (single
                                   and batched) only: SpMM, symmetric storage and export
                                   refuse a tiled handle (ABHSF_ERR_UNSUPPORTED) */
};

/* Symmetry of the stored matrix (matrix_symmetry_t, include/abhsf/utils/
 * matrix_properties.h; Hermitian equals symmetric for real values).  With
 * SYMMETRIC / SKEW_SYMMETRIC only one triangle is stored (the entries given
 * to the build, which must all lie on one side of the diagonal; no diagonal
 * for skew) and SpMV/SpMM apply the whole matrix: A = T + T_off^T or
 * T - T^T, so a stored triangle of nnz_stored entries does the work of
 * nnz_all (spmv/unified/spmv.cpp:1160-1166).  FP64 values only. */
enum { ABHSF_GENERAL = 0, ABHSF_SYMMETRIC = 1, ABHSF_SKEW_SYMMETRIC = 2 };

typedef struct abhsf_options {
    int32_t block_size;   /* s: power of two in [2, 256] */
    int32_t policy;       /* ABHSF_POLICY_* */
    int32_t precision;    /* ABHSF_F64 / ABHSF_F32 */
    int32_t variant;      /* ABHSF_VARIANT_* */
    int32_t is_binary;    /* pattern matrix: dense scheme excluded (test.cpp:70) */
    int32_t reserved;
    void *stream;         /* hipStream_t; NULL = null stream */
} abhsf_options;

typedef struct abhsf_stats {
    int64_t m, n, nnz;
    int32_t s, k;
    int64_t M, N;                  /* block rows / block columns */
    uint64_t blocks[4];            /* nonzero blocks per scheme: COO, CSR, bitmap, dense */
    uint64_t values[4];            /* stored values per scheme (dense includes zeros) */
    uint64_t smf_bits_preproc;     /* structure bits, preprocessing/test.cpp:76-85 */
    uint64_t smf_bits_msmf;        /* structure bits, block_statistics/mmf.cpp:68-101 (adaptive, no values) */
    uint64_t stored_bytes;         /* bytes of all storage arrays as stored in HBM */
    uint64_t index_bytes;          /* stored_bytes minus value bytes */
    double build_ms;               /* device time of the last build */
    int32_t variant;               /* ABHSF_VARIANT_* the storage was built for (AUTO resolved) */
    int32_t tile_zs1;              /* tiled: 1 if tile COO decode takes one entry per block in its
                                      first phase (blocks of 3+ entries are 1 in 64 or more), else 0 */
    int64_t work_items;            /* single-RHS work items: light items + heavy-strip units */
    int64_t max_strip_work;        /* entries of the largest strip (dense blocks: positions); above
                                      max(4096, nnz / 4096) abhsf_spmm runs each right-hand side
                                      through the SpMV kernels */
} abhsf_stats;

typedef struct abhsf_handle abhsf_handle;

/* Create an empty matrix handle.  Replaces the construction of the
 * reference's matrix classes (spmv/unified/spmv.cpp:71, 1205). */
abhsf_status abhsf_create(const abhsf_options *opts, abhsf_handle **out);

/* Build the ABHSF storage from a CSR matrix with 0-based indices, columns
 * sorted ascending and unique within each row.  on_device = 1: the three
 * arrays are device pointers; 0: host pointers (copied to the device).
 * The conversion runs on the GPU (key, sort, head-flag scan, per-block cost
 * and choice, per-scheme compaction, emit).
 * Replaces csr_matrix::from_elements (spmv/unified/spmv.cpp:74-104) and the
 * block iteration + footprint processor of preprocessing/test.cpp:36-108,237-368. */
abhsf_status abhsf_build_csr(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz,
                             const int64_t *rowptr, const int32_t *col, const double *val,
                             int32_t on_device);

/* Build from host COO triplets (any order, no duplicates); a private copy is
 * sorted row-major first.  Mirrors element_t = (row, col, val)
 * (spmv/unified/spmv.cpp:68-69) as three arrays. */
abhsf_status abhsf_build_coo(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz,
                             const int32_t *rows, const int32_t *cols, const double *vals);

/* Storage statistics (blocks per scheme, footprints under both cost models). */
abhsf_status abhsf_get_stats(const abhsf_handle *h, abhsf_stats *out);

/* Symmetry (ABHSF_GENERAL / _SYMMETRIC / _SKEW_SYMMETRIC) for the next build of
 * h; the reference carries it in matrix_properties::symmetry, passed to
 * from_elements (spmv/unified/spmv.cpp:74).  Building symmetric storage from
 * entries in both triangles fails with the reference reader's message
 * (matrix_market_reader.h:191-197). */
abhsf_status abhsf_set_symmetry(abhsf_handle *h, int32_t symmetry);

/* y = alpha * A * x + beta * y (device pointers; beta == 0 does not read y).
 * With alpha = beta = 1 this is the reference's accumulating
 * csr_matrix::spmv(x, y) (spmv/unified/spmv.cpp:140-159) and
 * spmv_cusparse_hyb (spmv/unified/spmv.cpp:199-207, which also syncs). */
abhsf_status abhsf_spmv(abhsf_handle *h, double alpha, const double *x, double beta, double *y);

/* Y = alpha * A * X + beta * Y for nrhs right-hand sides, X (n x nrhs) and
 * Y (m x nrhs) row-major with leading dimensions ldx, ldy >= nrhs.
 * nrhs in [1, 64].  No reference counterpart (batched multi-RHS extension).
 * Any alignment of X is accepted; X at a 16-byte boundary with an even ldx
 * reads whole column tiles in 16-byte loads (faster).
 * A tiled handle runs each right-hand side through its SpMV kernels (the
 * column gathered into, and the result scattered from, contiguous vectors
 * the handle allocates at the first such call). */
abhsf_status abhsf_spmm(abhsf_handle *h, int32_t nrhs, double alpha, const double *X, int64_t ldx,
                        double beta, double *Y, int64_t ldy);

/* Copy one storage array to host memory.  name is one of the array names in
 * the layout above ("coo_sptr", "coo_eptr", "coo_bcol", "coo_z1", "coo_val",
 * "coo_idx", "csr_...", "csr_rowptr", "csr_col", "bm_...", "bm_words",
 * "dn_..."), *bytes receives the array size; dst may be NULL to query it.
 * Used to check the GPU conversion byte-for-byte. */
abhsf_status abhsf_export(const abhsf_handle *h, const char *name, void *dst, size_t cap,
                          size_t *bytes);

/* Recreate the storage from arrays obtained with abhsf_export (a serialised
 * ABHSF blob: count (name, host pointer, bytes) triples covering every array of
 * the layout above), skipping the conversion.  The handle's options must match
 * the ones the arrays were built with.  Every array is validated on the host
 * (sizes, offsets, entries inside the m x n matrix, scheme = cost-model
 * choice) before upload; an inconsistent blob returns ABHSF_ERR_INVALID.
 * The reference keeps converted matrices only as text tables or the binary
 * COO .mm file (utils/mtx2mm/mtx2mm.cpp:101-136). */
abhsf_status abhsf_import(abhsf_handle *h, int64_t m, int64_t n, int64_t nnz, int32_t count,
                          const char *const *names, const void *const *data, const size_t *bytes);

/* Block statistics on the GPU: hist[z] = number of nonzero r x s blocks
 * (r = 2^kr, s = 2^ks, 0 <= kr, ks <= 16) holding z entries (z >= hist_cap - 1
 * lands in hist[hist_cap - 1]); *nblocks = number of nonzero blocks.  Only
 * every strip_stride-th strip of r rows is counted (1 = all rows, the
 * reference's statistic); *sampled_nnz receives the entries counted.  CSR as
 * for abhsf_build_csr.  Synchronous on `stream` (hipStream_t, may be NULL).
 * Replaces process_block_size_powers (block_statistics/block_stats.cpp:69-100),
 * which sorts the whole matrix once per block size. */
abhsf_status abhsf_block_hist(const int64_t *rowptr, const int32_t *col, int64_t m, int64_t n, int64_t nnz,
                              int32_t on_device, int32_t kr, int32_t ks, int64_t strip_stride, uint64_t *hist,
                              int64_t hist_cap, uint64_t *nblocks, uint64_t *sampled_nnz, void *stream);

/* ---- Batched launch of many matrices ---------------------------------------
 * No reference counterpart.  One grid-stride launch processes the work items
 * of every FP64 matrix (each with its own block size), a second one those of
 * the FP32-valued members (compensated accumulation), followed by the
 * heavy-strip fix-up of all members and, for members with symmetric or
 * skew-symmetric storage, the mirrored triangle's pass (one launch each);
 * with use_graph the sequence is captured once into a hipGraph and replayed.
 * The member handles must stay alive and unchanged while the batch exists. */
typedef struct abhsf_batch abhsf_batch;
abhsf_status abhsf_batch_create(abhsf_handle *const *mats, int32_t count, void *stream, abhsf_batch **out);
/* xs[i] / ys[i]: device vectors of matrix i (host arrays of device pointers). */
abhsf_status abhsf_batch_set_vectors(abhsf_batch *b, const double *const *xs, double *const *ys);
/* y_i = alpha * A_i * x_i + beta * y_i for every member, enqueued on the batch stream. */
abhsf_status abhsf_batch_spmv(abhsf_batch *b, double alpha, double beta, int32_t use_graph);
abhsf_status abhsf_batch_synchronize(abhsf_batch *b);
/* One eager batched SpMV that also measures, per member matrix, the
 * shader-clock cycles its work items occupied (sum over its wavefronts;
 * fix-up workgroups count once per wavefront).  cycles: host uint64[count].
 * Inside a full launch this is the matrix's share of the GPU, the label the
 * throughput tuner learns.  No counterpart in the reference (its tuner works
 * on storage footprints, block_statistics/mmf.h:77-102). */
abhsf_status abhsf_batch_profile(abhsf_batch *b, double alpha, double beta, uint64_t *cycles);
abhsf_status abhsf_batch_destroy(abhsf_batch *b);

/* ---- Tuner cost model: 2 x 64 hidden ReLU MLP, FP32, on the GPU ----------
 * No reference counterpart (the reference searches block sizes exhaustively
 * over storage bits, block_statistics/mmf.h:77-102); these are the learned
 * tuner's training and inference kernels.  All arrays are device pointers;
 * theta holds W1[64][nfeat], b1[64], W2[64][64], b2[64], W3[64], b3.
 * X is nsamples x nfeat row-major (1 <= nfeat <= 64).  H1, H2, D1, D2 are
 * nsamples x 64 work arrays; out, y, w, D3 have nsamples entries (w may be
 * NULL for unit weights, wsum = sum of w).  The loss is
 * sum_b w_b (out_b - y_b)^2 / wsum. */
int64_t abhsf_mlp_num_params(int32_t nfeat);
abhsf_status abhsf_mlp_forward(const float *theta, int32_t nfeat, const float *X, int64_t nsamples, float *H1,
                               float *H2, float *out, void *stream);
/* Ensemble inference in one launch: out (FP64, nsamples) = mean over the
 * nmembers parameter vectors thetas[m * num_params ..] of the forward
 * outputs, summed in member order. */
abhsf_status abhsf_mlp_predict(const float *thetas, int32_t nmembers, int32_t nfeat, const float *X,
                               int64_t nsamples, double *out, void *stream);
/* grad (num_params) = dLoss/dtheta, summed over samples in sample order. */
abhsf_status abhsf_mlp_backward(const float *theta, int32_t nfeat, const float *X, int64_t nsamples,
                                const float *H1, const float *H2, const float *out, const float *y,
                                const float *w, float wsum, float *D1, float *D2, float *D3, float *grad,
                                void *stream);
/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) updating
 * theta in place; loss_host (host, `steps` floats, may be NULL) receives the
 * loss before each step.  Synchronous on stream. */
abhsf_status abhsf_mlp_train(float *theta, int32_t nfeat, const float *X, const float *y, const float *w,
                             int64_t nsamples, int64_t steps, float lr, float *loss_host, void *stream);
/* abhsf_mlp_train without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving or tuning loops, harnesses that forbid hipMalloc).
 * The work buffer (activations, deltas, gradient, Adam moments) is a
 * file-scope device array of at most 4.4 MB per GPU, kept until the process
 * ends; limits: nsamples <= 4096 and steps <= 16384 (beyond them:
 * ABHSF_ERR_UNSUPPORTED, nothing enqueued).  Same kernels in the same order
 * as abhsf_mlp_train, so theta and the losses are bit-identical to its
 * results.  Differences: asynchronous on stream (no host synchronisation,
 * no memset); loss is a DEVICE array of `steps` floats, or NULL.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race. */
abhsf_status abhsf_mlp_train_noalloc(float *theta, int32_t nfeat, const float *X, const float *y, const float *w,
                                     int64_t nsamples, int64_t steps, float lr, float *loss, void *stream);

/* ---- Delta-coded coordinate storage (spmv/unified/spmv.cpp:628-1033) ----
 * The reference's delta_matrix family stores the deltas of linearised keys
 * key = (row << K) + col of the row-major elements as tagged varint words
 * ((delta << 2) | tag, 1/2/4/8 bytes for delta < 2^6 / 2^14 / 2^30 / more,
 * :859-896) and decodes them with a serial running sum (:923-944).  Here the
 * same words are kept in one aligned stream per width, the 2-bit tags apart,
 * and chunks of 512 elements carry their base key and stream offsets, so a
 * wavefront decodes 64 elements per step (ballots, scan).  Replaces
 * delta_matrix::from_elements / spmv / release (:635-710, 712-789, 790-800). */
typedef struct abhsf_delta abhsf_delta;

enum {
    ABHSF_DELTA_VARINT = 0,      /* tagged varint words of key = (row << K) + col (:859-944) */
    ABHSF_DELTA_U32_ROW = 1,     /* u32 deltas of key = (row << 31) + col, row-major order
                                    (spmv/orig/delta.cpp:124-151) */
    ABHSF_DELTA_U32_BLOCK = 2    /* u32 deltas of the 8 x 8 cache-block key ((row/8) << 31) +
                                    ((col/8) << 6) + ((row%8) << 3) + col%8, ordered by
                                    (row/8, col/8, row%8, col%8) (:636-689); n <= 2^28 */
};

typedef struct abhsf_delta_stats {
    int64_t m, n, nnz;
    int32_t key_bits;          /* varint: K of key = (row << K) + col; u32 codecs: 31 */
    int32_t chunk;             /* elements per chunk */
    uint64_t words[4];         /* varint: elements coded in 1 / 2 / 4 / 8 bytes; u32 codecs:
                                  words[2] elements with a u32 delta, words[3] exceptions */
    uint64_t stored_bytes;     /* all arrays as stored in HBM */
    double build_ms;           /* host encode + upload */
    int32_t codec;             /* ABHSF_DELTA_* */
    int32_t reserved;
} abhsf_delta_stats;

/* A varint handle (ABHSF_DELTA_VARINT). */
abhsf_status abhsf_delta_create(void *stream, abhsf_delta **out);
/* A handle for one of the ABHSF_DELTA_* codecs.  The u32 codecs keep the
 * reference's fixed-width words (replacing delta_matrix of
 * spmv/orig/delta.cpp:118-160 and the cache-blocking delta_matrix of
 * spmv/unified/spmv.cpp:628-710); where the reference asserts that a delta
 * fits 32 bits, a chunk's first element moves its key into the chunk base
 * (delta 0) and any other element stores 0 -- never a real delta there --
 * with its key in an exception list. */
abhsf_status abhsf_delta_create_codec(void *stream, int32_t codec, abhsf_delta **out);
/* CSR input as for abhsf_build_csr; the encoder runs on the host. */
abhsf_status abhsf_delta_build_csr(abhsf_delta *d, int64_t m, int64_t n, int64_t nnz, const int64_t *rowptr,
                                   const int32_t *col, const double *val, int32_t on_device);
/* y = alpha*A*x + beta*y on the handle's stream; beta = 1 is the reference's
 * accumulating delta_matrix::spmv(x, y). */
abhsf_status abhsf_delta_spmv(abhsf_delta *d, double alpha, const double *x, double beta, double *y);
abhsf_status abhsf_delta_get_stats(const abhsf_delta *d, abhsf_delta_stats *out);
/* Arrays: varint "tags" u32, "w8" u8, "w16" u16, "w32" u32, "w64" u64,
 * "start" u32[chunks][4]; u32 codecs "da" u32[nnz], "xstart" u32[chunks + 1],
 * "exc" u64; all "base" u64[chunks], "val" f64 (element order).  dst NULL:
 * size only. */
abhsf_status abhsf_delta_export(const abhsf_delta *d, const char *name, void *dst, size_t cap, size_t *bytes);
abhsf_status abhsf_delta_synchronize(abhsf_delta *d);
abhsf_status abhsf_delta_destroy(abhsf_delta *d);

/* Wait for all work enqueued on the handle's stream. */
abhsf_status abhsf_synchronize(abhsf_handle *h);

/* Release device memory and the handle.  Replaces csr_matrix::release
 * (spmv/unified/spmv.cpp:210-226). */
abhsf_status abhsf_destroy(abhsf_handle *h);

/* Message of the last failing call on this thread ("" if none). */
const char *abhsf_last_error(void);

/* Library version string. */
const char *abhsf_version(void);

#ifdef __cplusplus
}
#endif
#endif
