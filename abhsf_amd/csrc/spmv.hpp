// spmv.hpp -- launch interface of the ABHSF SpMV / SpMM kernels.
#pragma once

#include "common.hpp"

namespace abhsf {

// Work decomposition of the single-RHS kernel: light strips are processed
// whole, several consecutive ones per work item when they are small; heavy
// strips are cut into units (one scheme group's block range each) whose
// partial row sums are combined by a fix-up pass.
struct Balance {
    const uint32_t *light = nullptr;  // light strip ids (null: every strip, in order)
    const uint32_t *pack = nullptr;   // [n_light + 1] light item k = light[pack[k] .. pack[k+1])
                                      // (null: one strip per light item)
    const uint32_t *units = nullptr;  // [n_units][4] {scheme | strip << 2, b0, b1, value offset of b0}
    const uint32_t *heavy = nullptr;  // [n_heavy][3] {strip, first unit, unit count}
    const uint32_t *fix = nullptr;    // [n_fix][2] fix-up workgroups {heavy index, first row | rows << 16}
    double *part = nullptr;           // [n_units][s] (x2 when compensated)
    // Column-panel variant: slot table, kStripWaves slots (waves) per
    // workgroup; slot kStripWaves * b + w holds the item of wave w of
    // workgroup b, or kNoItem.  Workgroup b runs on XCD b % kXcdCount, so
    // items are grouped by panel = XCD.
    const uint32_t *order = nullptr;  // [n_slots]
    int64_t n_light = 0, n_units = 0, n_heavy = 0, n_slots = 0, n_fix = 0;
    // Tiled variant: light item k holds the strips light[pack[k] .. pack[k+1])
    // (a tile of COO-only strips, or one strip), accumulated together (their
    // rows (position << k) + lrow; kTileRows rows at most, kTileRowsComp with
    // FP32 values); its COO blocks are coo_rng[k] = {b0, b1, first value},
    // in (block column, strip) order.
    bool tiled = false;
    const uint32_t *coo_rng = nullptr;  // [n_light][3]
    // the tiles' COO blocks hold 3 or more entries often enough (one in 64
    // blocks or more) that phase B runs in most 64-block sub-chunks anyway:
    // phase A then takes one entry per block instead of two
    bool tile_zs1 = false;
};

#ifndef ABHSF_TILE_WORK
#define ABHSF_TILE_WORK 8192
#endif
#ifndef ABHSF_TILE_ROWS
#define ABHSF_TILE_ROWS 512
#endif
constexpr int kTileRows = ABHSF_TILE_ROWS;      // rows per tile at most (FP64 values)
constexpr int kTileRowsComp = 256;              // with FP32 values (two accumulators per row)
constexpr int64_t kTileWork = ABHSF_TILE_WORK;  // entries per tile at most

constexpr int kXcdCount = 8;            // XCDs of an MI355X; workgroups go round-robin
// Waves per workgroup of the single-RHS strip kernel.  Each wave runs its own
// items and holds its own LDS, but a workgroup's LDS is freed only when its
// last wave ends: with 4 waves R-MAT scale 22 kept about 5,000 of 8,192 wave
// slots busy.  2 waves: R-MAT (256, tiled) 490 -> 472 us, block_mix (32,
// tiled) 455 -> 448 us; 1 wave is no faster and loses the column panels'
// XCD grouping of 4-slot tables.
constexpr int kStripWaves = 2;
constexpr uint32_t kNoItem = 0xFFFFFFFFu;

struct SpmvLaunch {
    int bcol_bytes;          // 2 or 4
    int val_bytes;           // 8 (f64) or 4 (f32, compensated accumulation)
    bool coo_single = false; // COO blocks hold about one entry on average
    bool has_dense = false;  // some dense-scheme blocks (SpMM: matrix-core tiles at R >= 16)
    Balance bal;
};

// One matrix of a batched launch (device-resident table entry).
struct BatchEntry {
    DevMat A;
    Balance B;
    const double *x;
    double *y;
    int64_t item0;  // first global work item of this matrix
    int32_t code;   // index-type combination, batch_code()
    int32_t comp;   // 1: FP32 values with compensated accumulation (its items follow all FP64 ones)
};

// Index-type combination of a batch member (block-column width x in-block
// index widths), 0 .. kBatchCodes - 1.
constexpr int kBatchCodes = 6;
int batch_code(int bcol_bytes, int s);

// All matrices' work items in grid-stride launches (each member's index-type
// combination, BatchEntry::code, selects its decode): the FP64 members' items
// [0, G64) in one launch, the FP32 members' items [G64, G) in a second one
// (compensated accumulation), then the heavy-strip fix-up of every member
// (dheavy: pairs {matrix, heavy index}).  item_mat[g]: matrix of global item
// g.  With cycles (device, [nmat], zeroed by the caller) every wave adds its
// shader-clock cycles to its matrix's counter.
void batch_spmv(const BatchEntry *dE, const uint32_t *item_mat, int64_t G64, int64_t G, const uint32_t *dheavy,
                int64_t nheavy, double alpha, double beta, hipStream_t st, unsigned long long *cycles = nullptr);

// Symmetric storage, second pass: Y[j] += scale * T_off^T * X[j] for the
// stored triangle T without its diagonal (global atomics; FP64 values).
void spmv_transposed(const SpmvLaunch &L, const DevMat &A, int nrhs, double scale, const double *X, int64_t ldx,
                     double *Y, int64_t ldy, hipStream_t st);

// Columns [0, cols) of a row-major rows x ld array to contiguous columns (column
// c at dst + c * rows) and back; cols <= col_block().
int col_block();
void gather_cols(const double *src, int64_t lds, int64_t rows, int cols, double *dst, hipStream_t st);
void scatter_cols(const double *src, int64_t rows, int cols, double *dst, int64_t ldd, hipStream_t st);

// Rows per fix-up workgroup for a heavy strip of nu units (s-row strips;
// comp: FP32 values, compensated partials): a power of two that divides s.
int fix_rows(int s, uint32_t nu, bool comp);

// Largest power-of-two rhs tile whose s x R accumulator fits the LDS budget.
int max_rhs_tile(int s, bool comp);

// Y = alpha*A*X + beta*Y, X n x nrhs (ldx), Y m x nrhs (ldy), row-major, device.
void spmm(const SpmvLaunch &L, const DevMat &A, int nrhs, double alpha, const double *X, int64_t ldx, double beta,
          double *Y, int64_t ldy, hipStream_t st);

}  // namespace abhsf
