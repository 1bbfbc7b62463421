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

struct DeltaStorage {
    int64_t m = 0, n = 0, nnz = 0, nchunk = 0;
    int K = 1;
    uint64_t count[4] = {0, 0, 0, 0};  // elements per width class
    DevBuf tags;                       // u32[(nnz + 15) / 16]
    DevBuf words[4];                   // u8 / u16 / u32 / u64 tagged words
    DevBuf base;                       // u64[nchunk]
    DevBuf start;                      // u32[nchunk][4]
    DevBuf val;                        // f64[nnz]
    bool built = false;
    double build_ms = 0.0;

    uint64_t stored_bytes() const;
};

// Build from a host CSR (rowptr int64[m+1], col int32[nnz] ascending within
// rows, val f64[nnz] or null = all ones); throws Error on invalid input.
void delta_build(DeltaStorage &D, int64_t m, int64_t n, const int64_t *rowptr, const int32_t *col,
                 const double *val, hipStream_t st);

// y = alpha*A*x + beta*y (device vectors).
void delta_spmv(const DeltaStorage &D, double alpha, const double *x, double beta, double *y, hipStream_t st);

}  // namespace abhsf
