/*
 * abhsf_oracle.h -- CPU restatement of the reference's ABHSF algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C ABI or the host-side package) may link, load or call this code.  It is
 * used by tests/ as the parity checker, by __graft_entry__.smoke() as the
 * checker of one small GPU invocation, and by bench.py as the timed CPU
 * baseline ("port").
 *
 * Every function cites the reference file:line it restates.  What pins each
 * function (golden vectors, reference binaries, closed forms) is listed in
 * the test files under tests/ that exercise it.
 */
#ifndef ABHSF_ORACLE_H
#define ABHSF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_COO = 0, ORC_CSR = 1, ORC_BITMAP = 2, ORC_DENSE = 3 };
enum { ORC_POLICY_PREPROC = 0, ORC_POLICY_MSMF = 1 };

/* ceil(log2 x), log2(1) = 0.  include/abhsf/utils/math.h:17-31 and
 * preprocessing/test.cpp:88-98 (identical for x >= 1). */
uint64_t orc_ceil_log2(uint64_t x);

/* Per-block scheme cost and choice for an s x s block holding z entries.
 * PREPROC: preprocessing/test.cpp:55-77 (COO 2kz+log2 s^2, CSR kz+(s+1)log2 s^2,
 *          bitmap s^2, dense (s^2-z)b; strict '<' in the order COO,CSR,bitmap,dense).
 * MSMF:    include/abhsf/msmf.h:9-48 with r = s (COO log2(s^2)+z(2 log2 s),
 *          CSR s log2(s^2)+z log2 s, bitmap s^2, dense (s^2-z)b; std::min keeps
 *          the first of equal operands, which is the same order).
 * Writes the chosen scheme's bits (without header or tag) to *bits. */
int orc_choose(int policy, uint64_t s, uint64_t z, uint64_t b, int is_binary, uint64_t *bits);

/* Structure footprint in bits of processor type 2 (preprocessing/test.cpp:36-85):
 * sum over nonzero blocks of (log2 M + log2 N + 2 + min scheme bits).
 * rows/cols are 0-based coordinates of the nnz stored entries (any order). */
uint64_t orc_smf_preproc(uint64_t m, uint64_t n, uint64_t nnz, const uint32_t *rows,
                         const uint32_t *cols, uint64_t s, uint64_t b, int is_binary,
                         uint64_t nblk_by_scheme[4]);

/* Block-nnz histogram for r = 2^kr rows x s = 2^ks cols
 * (block_statistics/block_stats.cpp:69-100).  Writes up to cap (z, count)
 * pairs sorted by z; returns the number of distinct z (may exceed cap). */
uint64_t orc_block_hist(uint64_t nnz, const uint32_t *rows, const uint32_t *cols, int kr, int ks,
                        uint64_t *z_out, uint64_t *count_out, uint64_t cap);

/* mmf row for one (r,s) (block_statistics/mmf.cpp:53-138): out[8] =
 * {coo32, csr32, coo, csr, bitmap, dense, adaptive, adaptive134}, each with the
 * value bits added, for value precision bits_single (32 real) x mult (1 single,
 * 2 double).  hist is (z, count) pairs. */
void orc_mmf_row(uint64_t m, uint64_t n, uint64_t nnz, int is_binary, uint64_t bits_single, int mult,
                 uint64_t r, uint64_t s, uint64_t nh, const uint64_t *hz, const uint64_t *hc,
                 uint64_t out[8]);

/* ---------------------------------------------------------------------------
 * ABHSF storage built from a row-sorted CSR matrix.  The layout is the one
 * the HIP library stores in HBM; see include/abhsf_c.h for the contract.
 * ------------------------------------------------------------------------- */
typedef struct {
    int64_t m, n;
    int32_t s, k;            /* s = 2^k, 1 <= k <= 8 */
    int64_t M, N;            /* ceil(m/s), ceil(n/s) */
    int32_t policy, is_binary, val_bytes; /* val_bytes 8 (f64) or 4 (f32) */
    uint64_t cost_b;         /* value bits used by the cost model */
    int32_t bcol_bytes;      /* 2 if N <= 65536 else 4 */
    int32_t coo_idx_bytes;   /* 1 if s <= 16 else 2 */
    int32_t csr_rp_bytes;    /* 1 if s <= 8 else 2 */
    int32_t bm_nw;           /* u64 words per bitmap block: max(1, s*s/64) */
    uint64_t nnz;
    uint64_t nb[4];          /* blocks per scheme */
    uint64_t ne[4];          /* stored values per scheme (dense: nb*s*s) */
    uint32_t *sptr[4];       /* [M+1] block offsets per strip */
    uint32_t *eptr[4];       /* [M+1] value offsets per strip (dense: block*s*s) */
    void *bcol[4];           /* [nb] block column, u16 or u32 */
    uint16_t *z1[4];         /* [nb] entries in block minus one */
    void *coo_idx;           /* [ne COO] (lrow<<k)|lcol, u8 or u16 */
    void *csr_rowptr;        /* [nb CSR * (s+1)] u8 or u16 */
    uint8_t *csr_col;        /* [ne CSR] lcol */
    uint64_t *bm_words;      /* [nb BM * bm_words] bit lrow*s+lcol, LSB first */
    void *val[4];            /* values per scheme, f64 or f32 */
    uint64_t smf_bits;       /* structure bits of the chosen policy */
} orc_abhsf;

/* Build from CSR (rowptr[m+1], col sorted ascending and unique within a row).
 * Returns 0 on success, nonzero on invalid input. */
int orc_build(orc_abhsf *A, int64_t m, int64_t n, const int64_t *rowptr, const int32_t *col,
              const double *val, int k, int policy, int val_bytes, int is_binary);
void orc_free(orc_abhsf *A);

/* y = alpha*A*x + beta*y from the ABHSF storage (beta == 0 does not read y).
 * X is n x nrhs row-major with leading dim ldx, Y is m x nrhs with ldy.
 * OpenMP over strips when compiled with -fopenmp. */
void orc_spmm(const orc_abhsf *A, int nrhs, double alpha, const double *X, int64_t ldx, double beta,
              double *Y, int64_t ldy);

/* Textbook CSR FP64 SpMV y = alpha*A*x + beta*y (row-parallel). */
void orc_csr_spmv(int64_t m, const int64_t *rowptr, const int32_t *col, const double *val,
                  double alpha, const double *x, double beta, double *y);

/* Decode the storage back to triplets in (I,J,lrow,lcol) order; returns count.
 * Dense blocks yield all s*s positions inside the matrix whose value != 0. */
uint64_t orc_decode(const orc_abhsf *A, int64_t *rows, int64_t *cols, double *vals, uint64_t cap);

/* Population standard deviation of per-row counts
 * (block_statistics/rows-nnz-stdev.cpp:11-30). */
double orc_rows_stdev(uint64_t m, const uint64_t *rows_nnz);

/* Tagged varint delta stream of linearised keys (spmv/unified/spmv.cpp:
 * 859-896, one thread): keys[0] is the start value (not encoded); element
 * k >= 1 stores delta = keys[k] - keys[k-1] little-endian as
 * (delta << 2) | tag in 1 byte (delta < 2^6, tag 0), 2 bytes (< 2^14, tag 1),
 * 4 bytes (< 2^30, tag 2) or 8 bytes (tag 3).  out may be NULL (size only);
 * returns the stream length in bytes. */
uint64_t orc_varint_encode(uint64_t nnz, const uint64_t *keys, uint8_t *out);

/* Decode of the same stream with a running sum (spmv/unified/spmv.cpp:
 * 923-944): keys[0] = first, keys[k] = keys[k-1] + delta.  Returns the bytes
 * consumed. */
uint64_t orc_varint_decode(uint64_t nnz, uint64_t first, const uint8_t *in, uint64_t *keys);

/* Fixed-width u32 delta codecs of linearised keys.
 * layout 1 (spmv/orig/delta.cpp:124-151): row-major order,
 *   key = (row << 31) + col;
 * layout 2 (spmv/unified/spmv.cpp:636-689, 8x8 cache blocks): order by
 *   (row / 8, col / 8, row % 8, col % 8), key = ((row / 8) << 31) +
 *   ((col / 8) << 6) + ((row % 8) << 3) + col % 8 (needs n <= 2^28).
 * Element e (in key order) stores da[e] = key_e - key_{e-1} as a u32, with
 * key_{-1} = 0.  Defined by this build (the reference asserts instead):
 * elements come in chunks of `chunk`; base[c] is the key before chunk c's
 * first element, or that element's own key (and its da 0) when its delta
 * needs more than 32 bits.  Any other element whose delta needs more than 32
 * bits stores 0 -- never a real delta there, keys strictly increase -- and
 * its key goes to the exception list exc, in element order.  xstart[c] is the
 * chunk's first exception (xstart[nchunk] =
 * exceptions in all).  perm[e] = CSR position of element e.  Inputs: CSR
 * with columns ascending within rows.  Arrays sized nnz / nnz / nchunk /
 * nchunk + 1 (exc: nnz at most).  Returns the number of exceptions. */
uint64_t orc_delta32_encode(int layout, int64_t m, const int64_t *rowptr, const int32_t *col, uint64_t chunk,
                            uint32_t *da, uint64_t *perm, uint64_t *base, uint32_t *xstart, uint64_t *exc);

/* Running-sum decode of the same stream: rows[e], cols[e] of element e. */
void orc_delta32_decode(int layout, uint64_t nnz, uint64_t chunk, const uint32_t *da, const uint64_t *base,
                        const uint32_t *xstart, const uint64_t *exc, int64_t *rows, int64_t *cols);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
