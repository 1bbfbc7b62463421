// spmv_unified.cpp -- the reference's SpMV harness (spmv/unified/spmv.cpp main,
// lines 1072-1283) on top of the ABHSF library: same matrix families, same
// x = 0.5 / y = 0 set-up, same 5 warm-up + 20 timed accumulating spmv calls,
// same "Measured MFLOP/s ... result = ||y||/25" line next to the expected
// ||A*0.5||.  Differences: the random-column family is seeded; x and y live in
// device memory and the matrix is ABHSF on the GPU.
//
// usage: spmv_unified <dense|diagonal|single-column|random-column|
//                      single-row-column|single-row|file.mtx> [n] [block_size]
//                      [strip|balanced|panel|tiled]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>

#include "../../include/abhsf/abhsf.hpp"

using namespace abhsf;

static const int warmup_iters = 5;  // spmv/unified/spmv.cpp:54-55
static const int n_iters = 20;

#define HIPCHECK(x)                                                                       \
    do {                                                                                  \
        if ((x) != hipSuccess) throw std::runtime_error("Error running " #x " function!"); \
    } while (0)

// Coordinate Matrix Market reader: real or pattern, general or symmetric
// (only the stored triangle is kept, as in the reference); explicit zeros are
// dropped (include/abhsf/utils/matrix_market_reader.h:145-200).
static void read_mtx(const std::string &fn, elements_t &el, matrix_properties &props)
{
    std::ifstream f(fn);
    if (!f) throw std::runtime_error("Cannot open input file: " + fn);
    std::string line;
    std::getline(f, line);
    if (line.rfind("%%MatrixMarket", 0) != 0) throw std::runtime_error("Could not process Matrix Market banner.");
    if (line.find("coordinate") == std::string::npos)
        throw std::runtime_error("Only sparse matrices in coordinate format are supported.");
    const bool pattern = line.find("pattern") != std::string::npos;
    if (!pattern && line.find("real") == std::string::npos && line.find("integer") == std::string::npos)
        throw std::runtime_error("read_matrix_real_or_binary() invoked for non-real-non-binary matrix.");
    props.type = pattern ? matrix_type_t::BINARY : matrix_type_t::REAL;
    props.symmetry = line.find("general") != std::string::npos          ? matrix_symmetry_t::UNSYMMETRIC
                     : line.find("skew-symmetric") != std::string::npos ? matrix_symmetry_t::SKEW_SYMMETRIC
                                                                        : matrix_symmetry_t::SYMMETRIC;
    while (std::getline(f, line))
        if (!line.empty() && line[0] != '%') break;
    std::istringstream hs(line);
    uintmax_t nnz;
    hs >> props.m >> props.n >> nnz;
    el.reserve(nnz);
    uintmax_t zeros = 0;
    for (uintmax_t k = 0; k < nnz; k++) {
        long r, c;
        double v = 1.0;
        f >> r >> c;
        if (!pattern) f >> v;
        if (!f) throw std::runtime_error("Could not read matrix element.");
        if (!pattern && v == 0.0) { zeros++; continue; }
        el.emplace_back((index_type)(r - 1), (index_type)(c - 1), v);
    }
    props.nnz = nnz - zeros;
}

int main(int argc, char *argv[])
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <family|file.mtx> [n] [block_size] [strip|balanced|panel|tiled]\n",
                     argv[0]);
        return 2;
    }
    try {
        elements_t elements;
        matrix_properties props;
        const std::string kind = argv[1];
        const uintmax_t n = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 0;
        const int s = argc > 3 ? std::atoi(argv[3]) : 64;
        int variant = ABHSF_VARIANT_AUTO;
        if (argc > 4) {
            const std::string v = argv[4];
            if (v == "strip") variant = ABHSF_VARIANT_STRIP;
            else if (v == "balanced") variant = ABHSF_VARIANT_BALANCED;
            else if (v == "panel") variant = ABHSF_VARIANT_PANEL;
            else if (v == "tiled") variant = ABHSF_VARIANT_TILED;
            else throw std::runtime_error("unknown variant " + v);
        }
        props.type = matrix_type_t::REAL;
        props.symmetry = matrix_symmetry_t::UNSYMMETRIC;
        // matrix families of spmv/unified/spmv.cpp:1078-1156 (sizes default smaller)
        if (kind == "dense") {
            props.m = props.n = n ? n : 2000;
            for (uintmax_t i = 0; i < props.m; i++)
                for (uintmax_t j = 0; j < props.n; j++) elements.emplace_back(i, j, 1.0);
        } else if (kind == "diagonal") {
            props.m = props.n = n ? n : 2000000;
            for (uintmax_t i = 0; i < props.m; i++) elements.emplace_back(i, i, 1.0);
        } else if (kind == "single-column") {
            props.m = props.n = n ? n : 2000000;
            for (uintmax_t i = 0; i < props.m; i++) elements.emplace_back(i, 0, 1.0);
        } else if (kind == "random-column") {
            props.m = props.n = n ? n : 2000000;
            std::mt19937 mt(0x0498bc04);
            std::uniform_int_distribution<int> dist(0, (int)props.n - 1);
            for (uintmax_t i = 0; i < props.m; i++) elements.emplace_back(i, dist(mt), 1.0);
        } else if (kind == "single-row-column") {
            props.m = props.n = n ? n : 1000000;
            for (uintmax_t j = 0; j < props.n; j++) elements.emplace_back(0, j, 1.0);
            for (uintmax_t i = 1; i < props.m; i++) elements.emplace_back(i, 0, 1.0);
        } else if (kind == "single-row") {
            props.m = props.n = n ? n : 2000000;
            for (uintmax_t i = 0; i < props.n; i++) elements.emplace_back(props.m - 1, i, 1.0);
        } else {
            read_mtx(kind, elements, props);
        }
        props.nnz = elements.size();

        uintmax_t nnz_all = 0;  // spmv/unified/spmv.cpp:1160-1166
        for (const auto &e : elements) {
            nnz_all++;
            if (props.symmetry != matrix_symmetry_t::UNSYMMETRIC && std::get<0>(e) != std::get<1>(e)) nnz_all++;
        }
        std::cout << "Nonzeros (stored): " << elements.size() << "\nNonzeros (all):    " << nnz_all << std::endl;

        const double x_element = 0.5;
        std::vector<double> hx(props.n, x_element), hy(props.m, 0.0);
        double *x, *y;
        HIPCHECK(hipMalloc(&x, props.n * sizeof(double)));
        HIPCHECK(hipMalloc(&y, props.m * sizeof(double)));
        HIPCHECK(hipMemcpy(x, hx.data(), props.n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(y, hy.data(), props.m * sizeof(double), hipMemcpyHostToDevice));

        // verification (spmv/unified/spmv.cpp:1195-1199) over all nnz_all
        // entries: symmetric storage applies the mirrored triangle as well
        // (the reference's kernels use the stored triangle only)
        const bool sym = props.symmetry != matrix_symmetry_t::UNSYMMETRIC;
        const double mirror = props.symmetry == matrix_symmetry_t::SKEW_SYMMETRIC ? -1.0 : 1.0;
        std::vector<double> yt(props.m, 0.0);
        for (const auto &e : elements) {
            yt[std::get<0>(e)] += std::get<2>(e) * x_element;
            if (sym && std::get<0>(e) != std::get<1>(e)) yt[std::get<1>(e)] += mirror * std::get<2>(e) * x_element;
        }
        double expected = 0.0;
        for (double v : yt) expected += v * v;
        std::cout << "Expected result = " << std::sqrt(expected) << std::endl;

        const double n_mflops = (double)(nnz_all * 2 * n_iters) / 1.0e6;
        abhsf_matrix A(s, ABHSF_POLICY_PREPROC, ABHSF_F64, variant);
        A.from_elements(elements, props);
        const abhsf_stats st = A.stats();
        std::cout << "Block size: " << s << ", blocks COO/CSR/bitmap/dense: " << st.blocks[0] << "/" << st.blocks[1]
                  << "/" << st.blocks[2] << "/" << st.blocks[3] << ", structure bits: " << st.smf_bits_preproc
                  << std::endl;

        for (int iter = 0; iter < warmup_iters; iter++) A.spmv(x, y);
        const auto t0 = std::chrono::high_resolution_clock::now();
        for (int iter = 0; iter < n_iters; iter++) A.spmv(x, y);
        const double sec =
            std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        HIPCHECK(hipMemcpy(hy.data(), y, props.m * sizeof(double), hipMemcpyDeviceToHost));
        double res = 0.0;
        for (double v : hy) res += v * v;
        std::printf("%-36s%10.1f, result = %.6f\n", "Measured MFLOP/s ABHSF:", n_mflops / sec,
                    std::sqrt(res) / (double)(warmup_iters + n_iters));
        A.release();
        HIPCHECK(hipFree(x));
        HIPCHECK(hipFree(y));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
