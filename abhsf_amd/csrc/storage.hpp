// storage.hpp -- the ABHSF storage owned by a handle, and its GPU builder.
#pragma once

#include "prims.hpp"
#include "spmv.hpp"

namespace abhsf {

struct Storage {
    // shape and options
    int64_t m = 0, n = 0, nnz = 0, M = 0, N = 0;
    int s = 0, k = 0, W = 0;
    int policy = ABHSF_POLICY_PREPROC, is_binary = 0, val_bytes = 8;
    uint64_t cost_b = 64;
    int bcol_bytes = 4, coo_idx_bytes = 2, csr_rp_bytes = 2;
    bool built = false;
    // per scheme
    uint64_t nb[4] = {0, 0, 0, 0}, ne[4] = {0, 0, 0, 0};
    DevBuf sptr[4], eptr[4], bcol[4], z1[4], val[4];
    DevBuf coo_idx, csr_rowptr, csr_col, bm_words;
    // work decomposition of the single-RHS kernel (heavy strips split in units;
    // with column panels, the slot table places each item on its panel's XCD;
    // bal_heavy holds the heavy strips, then the fix-up workgroups)
    DevBuf bal_light, bal_pack, bal_units, bal_heavy, bal_part, bal_order;
    int64_t n_light = 0, n_units = 0, n_heavy = 0, n_slots = 0, n_fix = 0;
    // tiled variant: light items are tiles of consecutive strips whose COO
    // blocks are stored in (block column, strip) order; COO z1 holds z - 1 in
    // its low coo_zb bits and the strip within the tile above them
    bool tiled = false;
    bool tile_zs1 = false;  // tiles decode one COO entry per block in phase A (see Balance)
    int coo_zb = 16;
    DevBuf bal_coo_rng;  // tiled: [n_light][3] COO {first block, end block, first value} of each item
    // entries of the largest strip (dense blocks: positions), for the SpMM
    // dispatch: its kernel gives each strip to one wave
    int64_t max_strip_work = 0;
    // footprints (bits) under both cost models
    uint64_t smf_preproc = 0, smf_msmf = 0;
    double build_ms = 0.0;

    DevMat view() const;
    SpmvLaunch launch() const;
    // Split strips whose work exceeds 2*unit entries into units of about
    // `unit` entries, and pack runs of consecutive light strips into items of
    // at most unit/2 entries (unit <= 0: neither).  With panels > 1 the block
    // columns are also cut into `panels` column panels of equal work: a strip
    // whose blocks span several panels is split at the panel boundaries, and
    // every work item is placed on the XCD of its panel (panels must equal the
    // XCD count, kXcdCount).  With tile_rows > 0 (tiled variant; panels must
    // be 1) light strips whose entries are all COO are gathered, in strip
    // order, into tiles of at most tile_rows rows and kTileWork entries (other
    // light strips stay items of their own), and the COO group is reordered so
    // that every item's blocks are contiguous, a tile's sorted by (block
    // column, strip).
    void build_balance(int64_t unit, hipStream_t st, int panels = 1, int tile_rows = 0);
    uint64_t stored_bytes() const;
    uint64_t value_bytes() const;
    void set_shape(int64_t m_, int64_t n_, int s_);
};

// Build S from a device CSR (rowptr int64[m+1], col int32[nnz], val f64[nnz]).
// S's shape/options must be set; throws abhsf::Error on invalid input.
void build_from_device_csr(Storage &S, const int64_t *rowptr, const int32_t *col, const double *val,
                           hipStream_t st);

// Triangle flags of a device CSR matrix (sorted columns): bit 0 an entry
// below the diagonal, bit 1 above, bit 2 on it.
uint32_t csr_triangle_flags(const int64_t *rowptr, const int32_t *col, int64_t m, hipStream_t st);

// Recreate S (shape and options already set) from exported arrays in host
// memory; throws abhsf::Error if the arrays are not a valid storage.
void import_storage(Storage &S, int32_t count, const char *const *names, const void *const *data,
                    const size_t *bytes, hipStream_t st);

}  // namespace abhsf
