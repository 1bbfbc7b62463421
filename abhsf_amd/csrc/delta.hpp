// delta.hpp -- delta-coded coordinate storage (tagged varint deltas of
// linearised keys) and its SpMV.
#pragma once

#include "prims.hpp"

namespace abhsf {

// Elements in row-major order, key = (row << K) + col (K = ceil_log2(n),
// at least 1).  Element e stores delta_e = key_e - key_{e-1} (key_{-1} = 0)
// as the reference's tagged word (delta << 2) | tag with tag 0/1/2/3 for
// delta < 2^6 / 2^14 / 2^30 / otherwise, 1/2/4/8 bytes wide
// (spmv/unified/spmv.cpp:859-896).  Layout for parallel decode: the 2-bit
// tags are also kept apart (16 per u32 word), and the words go to one
// naturally aligned stream per width, in element order within a width;
// chunks of kDeltaChunk elements carry the key before their first element
// and their start index in each width stream, so every wavefront starts
// decoding without looking at other chunks.
constexpr int kDeltaChunk = 512;

// Codecs (ABHSF_DELTA_* in the C ABI):
//   kDeltaVarint     the tagged varint words above (spmv/unified/spmv.cpp:859-944);
//   kDeltaU32Row     u32 deltas of key = (row << 31) + col in row-major order
//                    (spmv/orig/delta.cpp:124-151);
//   kDeltaU32Block   u32 deltas of key = ((row/8) << 31) + ((col/8) << 6) +
//                    ((row%8) << 3) + col%8, elements ordered by (row/8, col/8,
//                    row%8, col%8): 8 x 8 cache blocks of x and y
//                    (spmv/unified/spmv.cpp:636-689); needs n <= 2^28.
// The u32 codecs keep the reference's fixed-width words; where the reference
// asserts that a delta fits 32 bits, a chunk's first element instead moves
// its key into the chunk base (delta 0), and any other element stores 0 --
// never a real delta there, keys strictly increase -- with its key in the
// exception list exc (element order; chunk c's first exception xstart[c]).
enum { kDeltaVarint = 0, kDeltaU32Row = 1, kDeltaU32Block = 2 };

struct DeltaStorage {
    int64_t m = 0, n = 0, nnz = 0, nchunk = 0;
    int K = 1;
    int codec = kDeltaVarint;
    uint64_t count[4] = {0, 0, 0, 0};  // elements per width class (u32 codecs: [2] words, [3] exceptions)
    DevBuf tags;                       // varint: u32[(nnz + 15) / 16]
    DevBuf words[4];                   // varint: u8 / u16 / u32 / u64 tagged words
    DevBuf base;                       // u64[nchunk]
    DevBuf start;                      // varint: u32[nchunk][4]
    DevBuf da;                         // u32 codecs: u32[nnz]
    DevBuf xstart;                     // u32 codecs: u32[nchunk + 1]
    DevBuf exc;                        // u32 codecs: u64[exceptions]
    DevBuf val;                        // f64[nnz], in element order
    bool built = false;
    double build_ms = 0.0;

    uint64_t stored_bytes() const;
};

// Build from a host CSR (rowptr int64[m+1], col int32[nnz] ascending within
// rows, val f64[nnz] or null = all ones) with the given codec; throws Error
// on invalid input.
void delta_build(DeltaStorage &D, int64_t m, int64_t n, const int64_t *rowptr, const int32_t *col,
                 const double *val, hipStream_t st, int codec = kDeltaVarint);

// y = alpha*A*x + beta*y (device vectors).
void delta_spmv(const DeltaStorage &D, double alpha, const double *x, double beta, double *y, hipStream_t st);

}  // namespace abhsf
