// abhsf.hpp -- header-only C++ host interface over the C ABI (abhsf_c.h).
//
// Mirrors the shape of the reference's kernel interface, the matrix classes
// of spmv/unified/spmv.cpp:71-242:
//   from_elements(elements, props)  build from (row, col, val) tuples
//   spmv(x, y)                      y += A*x, returns when done
//   release()                       free the storage
// Errors are thrown as std::runtime_error("Error running <fn> function! ..."),
// the reference's convention (spmv/unified/spmv.cpp:109-136).  x and y are
// device pointers, as for the reference's GPU variant spmv_cusparse_hyb
// (spmv/unified/spmv.cpp:199-207).
#ifndef ABHSF_ABHSF_HPP
#define ABHSF_ABHSF_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../abhsf_c.h"

namespace abhsf {

using real_type = double;  // spmv/unified/spmv.cpp:59
using index_type = int;    // spmv/unified/spmv.cpp:65
using element_t = std::tuple<index_type, index_type, real_type>;  // spmv/unified/spmv.cpp:68
using elements_t = std::vector<element_t>;

// include/abhsf/utils/matrix_properties.h:8-29
enum class matrix_type_t { BINARY, INTEGER, REAL, COMPLEX };
enum class matrix_symmetry_t { SYMMETRIC, HERMITIAN, SKEW_SYMMETRIC, UNSYMMETRIC };
struct matrix_properties {
    matrix_type_t type = matrix_type_t::REAL;
    matrix_symmetry_t symmetry = matrix_symmetry_t::UNSYMMETRIC;
    uintmax_t m = 0, n = 0, nnz = 0;
};

inline void check(abhsf_status st, const char *fn)
{
    if (st != ABHSF_OK)
        throw std::runtime_error(std::string("Error running ") + fn + " function! " + abhsf_last_error());
}

class abhsf_matrix {
  public:
    explicit abhsf_matrix(int block_size = 64, int policy = ABHSF_POLICY_PREPROC, int precision = ABHSF_F64,
                          int variant = ABHSF_VARIANT_AUTO, void *stream = nullptr)
    {
        abhsf_options o{};
        o.block_size = block_size;
        o.policy = policy;
        o.precision = precision;
        o.variant = variant;
        o.stream = stream;
        check(abhsf_create(&o, &h_), "abhsf_create");
    }
    ~abhsf_matrix() { release(); }
    abhsf_matrix(const abhsf_matrix &) = delete;
    abhsf_matrix &operator=(const abhsf_matrix &) = delete;
    abhsf_matrix(abhsf_matrix &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }

    // The caller's vector is left as it is (a private copy is sorted).
    void from_elements(const elements_t &elements, const matrix_properties &props)
    {
        std::vector<int32_t> r(elements.size()), c(elements.size());
        std::vector<double> v(elements.size());
        for (size_t k = 0; k < elements.size(); k++) {
            r[k] = std::get<0>(elements[k]);
            c[k] = std::get<1>(elements[k]);
            v[k] = std::get<2>(elements[k]);
        }
        // symmetric / skew-symmetric: one stored triangle, SpMV applies both
        const int32_t sym = props.symmetry == matrix_symmetry_t::SKEW_SYMMETRIC ? ABHSF_SKEW_SYMMETRIC
                            : props.symmetry == matrix_symmetry_t::UNSYMMETRIC  ? ABHSF_GENERAL
                                                                                : ABHSF_SYMMETRIC;
        check(abhsf_set_symmetry(h_, sym), "abhsf_set_symmetry");
        check(abhsf_build_coo(h_, (int64_t)props.m, (int64_t)props.n, (int64_t)elements.size(), r.data(), c.data(),
                              v.data()),
              "abhsf_build_coo");
    }

    void from_csr(int64_t m, int64_t n, int64_t nnz, const int64_t *rowptr, const int32_t *col, const double *val,
                  bool on_device)
    {
        check(abhsf_build_csr(h_, m, n, nnz, rowptr, col, val, on_device ? 1 : 0), "abhsf_build_csr");
    }

    // y += A*x and wait (device pointers).
    void spmv(const real_type *x, real_type *y)
    {
        check(abhsf_spmv(h_, 1.0, x, 1.0, y), "abhsf_spmv");
        check(abhsf_synchronize(h_), "abhsf_synchronize");
    }

    // y = alpha*A*x + beta*y, enqueued on the handle's stream.
    void spmv(real_type alpha, const real_type *x, real_type beta, real_type *y)
    {
        check(abhsf_spmv(h_, alpha, x, beta, y), "abhsf_spmv");
    }

    void spmm(int nrhs, real_type alpha, const real_type *X, int64_t ldx, real_type beta, real_type *Y, int64_t ldy)
    {
        check(abhsf_spmm(h_, nrhs, alpha, X, ldx, beta, Y, ldy), "abhsf_spmm");
    }

    void synchronize() { check(abhsf_synchronize(h_), "abhsf_synchronize"); }

    abhsf_stats stats() const
    {
        abhsf_stats s{};
        check(abhsf_get_stats(h_, &s), "abhsf_get_stats");
        return s;
    }

    void release()
    {
        if (h_) {
            abhsf_destroy(h_);
            h_ = nullptr;
        }
    }

    abhsf_handle *handle() const { return h_; }

  private:
    abhsf_handle *h_ = nullptr;
};

}  // namespace abhsf

#endif
