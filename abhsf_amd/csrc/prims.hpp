// prims.hpp -- device buffers, scans and a stable LSD radix sort (gfx950).
#pragma once

#include "common.hpp"

#include <vector>

namespace abhsf {

// RAII device buffer.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t bytes)
    {
        release();
        n_ = bytes;
        if (bytes == 0) return;
        hipError_t e = hipMalloc(&p_, bytes);
        if (e != hipSuccess) {
            p_ = nullptr;
            n_ = 0;
            throw Error(ABHSF_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                             hipGetErrorString(e));
        }
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    template <typename T = void>
    T *get() const { return static_cast<T *>(p_); }
    size_t bytes() const { return n_; }

  private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

// Exclusive scan out[i] = sum_{j<i} in[j]; out may alias in.  If total is
// non-null the sum of all n inputs is written there (device pointer).
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total, hipStream_t st);
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total, hipStream_t st);

// Stable LSD radix sort of (key, value) pairs by the low key_bits of key.
// keys/vals are sorted in place (scratch of the same sizes is allocated).
void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, hipStream_t st);

}  // namespace abhsf
