// spc5_compare.cpp -- a reference-style driver built against the drop-in
// header include/spc5_hip.hpp: the flow of the reference's
// examples/load_mm_and_compare.cpp test<T> (CSR product 16 times, then
// every beta(r,VS) format converted from COO and applied 16 times, then the
// prepared split/Spmv_omp pair), with the accuracy printed in the
// reference's words plus a normwise error used by tests/test_gpu_cpp_dropin.py.
//
// usage: spc5_compare <dense N | lap N | mm FILE | api FILE> <double|float>
#include "spc5_hip.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

template <class T>
static std::vector<Ijv<T>> build_dense(int n) {  // examples/load_mm_and_compare.cpp:431-449
    std::vector<Ijv<T>> v;
    v.reserve((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) v.push_back({i, j, T(i + j)});
    return v;
}

template <class T>
static std::vector<Ijv<T>> build_lap(int n) {
    std::vector<Ijv<T>> v;
    for (int i = 0; i < n * n; ++i) {
        const int gi = i / n, gj = i % n;
        if (gi > 0) v.push_back({i, i - n, T(-1)});
        if (gj > 0) v.push_back({i, i - 1, T(-1)});
        v.push_back({i, i, T(4)});
        if (gj < n - 1) v.push_back({i, i + 1, T(-1)});
        if (gi < n - 1) v.push_back({i, i + n, T(-1)});
    }
    return v;
}

// Coordinate real general/symmetric MatrixMarket (1-based), as loadMM reads.
template <class T>
static std::vector<Ijv<T>> load_mm(const std::string& path, int& nrows, int& ncols) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line;
    std::getline(f, line);
    const bool symmetric = line.find("symmetric") != std::string::npos;
    while (std::getline(f, line) && !line.empty() && line[0] == '%') {
    }
    long long nnz;
    std::istringstream(line) >> nrows >> ncols >> nnz;
    std::vector<Ijv<T>> v;
    for (long long k = 0; k < nnz; ++k) {
        int i, j;
        double val;
        f >> i >> j >> val;
        v.push_back({i - 1, j - 1, T(val)});
        if (symmetric && i != j) v.push_back({j - 1, i - 1, T(val)});
    }
    return v;
}

template <class T>
static T check_accuracy(const T* y1, const T* y2, int n) {  // examples/load_mm_and_compare.cpp:262-274
    T m = 0;
    for (int i = 0; i < n; ++i) {
        if (y1[i] == 0.0 || y2[i] == 0.0)
            m = std::max(m, std::abs(y1[i] + y2[i]));
        else
            m = std::max(m, std::abs((y1[i] - y2[i]) / y1[i]));
    }
    return m;
}

template <class T>
static int run(std::vector<Ijv<T>> values, int nrows, int ncols) {
    const int nnz = (int)values.size();
    const int loops = 16;
    std::cout << "-> number of rows = " << nrows << "\n-> number of values = " << nnz << "\n";
    std::vector<T> x((size_t)ncols + SPC5_VEC_PADDING_X, T(0));
    for (int c = 0; c < ncols; ++c) x[c] = T(c % 98);
    std::vector<T> ycsr((size_t)nrows + SPC5_VEC_PADDING_Y, T(0));
    double anorm = 0;
    {
        SPC5Mat<T> csr = COO_unsorted_to_CSR<T>(nrows, ncols, values.data(), nnz);
        for (int l = 0; l < loops; ++l) CSR_Spmv_scalar<T>(csr, x.data(), ycsr.data());
        for (int i = 0; i < nrows; ++i) {
            double s = 0;
            for (int p = csr.rowsSize[i]; p < csr.rowsSize[i + 1]; ++p) s += std::abs((double)csr.values[p]);
            anorm = std::max(anorm, s);
        }
        auto best = SPC5_find_best<T>(csr);
        std::cout << "-> Esimated performance are " << best.second << " for " << SPC5_type_to_string(best.first)
                  << "\n";
    }
    const double scale = std::max(1.0, loops * anorm * 97.0);
    int fails = 0;
    const SPC5_MATRIX_TYPE fmts[4] = {FORMAT_1rVc_WT, FORMAT_2rVc, FORMAT_4rVc, FORMAT_8rVc};
    for (int prepared = 0; prepared < 2; ++prepared) {
        for (SPC5_MATRIX_TYPE fmt : fmts) {
            std::vector<T> y((size_t)nrows + SPC5_VEC_PADDING_Y, T(0));
            SPC5Mat<T> mat = COO_unsorted_to_CSR<T>(nrows, ncols, values.data(), nnz);
            const int nb_expect = SPC5_block_count<T>(mat, fmt);
            CSR_to_SPC5<T>(&mat, fmt);
            auto t0 = std::chrono::steady_clock::now();
            if (prepared) {
                auto iv = SPC5_split_omp<T>(mat);
                for (int l = 0; l < loops; ++l) SPC5_Spmv_omp<T>(mat, x.data(), y.data(), iv);
            } else {
                for (int l = 0; l < loops; ++l) SPC5_Spmv<T>(mat, x.data(), y.data());
            }
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            double nerr = 0;
            for (int i = 0; i < nrows; ++i) nerr = std::max(nerr, std::abs((double)y[i] - (double)ycsr[i]));
            nerr /= scale;
            long long count = 0;
            SPC5_iterate<T>(mat, [&](int, int, T) { ++count; });
            std::cout << "Start " << (prepared ? "prepared " : "usual ") << SPC5_type_to_string(fmt) << ":\n"
                      << "-> Number of blocks " << mat.numberOfBlocks << "( avg. "
                      << double(mat.numberOfNNZ) / double(std::max(1, mat.numberOfBlocks)) << " values per block)\n"
                      << "-> GFlops " << 2.0 * nnz * loops / sec / 1e9 << "s\n"
                      << "-> Max Difference in Accuracy " << check_accuracy(ycsr.data(), y.data(), nrows) << "\n"
                      << "-> normwise " << nerr << "\n";
            if (mat.numberOfBlocks != nb_expect || count != nnz) {
                std::cout << "-> FAIL block count / iterate mismatch\n";
                ++fails;
            }
        }
    }
    std::cout << "RESULT " << (fails ? "FAIL" : "DONE") << std::endl;
    return fails ? 1 : 0;
}

// API check on a CSR matrix stored raw (int64 nrows, ncols, nnz; int32
// rowptr[nrows+1]; int32 col[nnz]; double values[nnz]): prints the choices
// of SPC5_find_best (and SPC5_find_best_omp<double> for 8 and 36 threads),
// and writes, for every format, the (row, column, value) triplets that
// SPC5_iterate visits to FILE.iter_r<r>.bin (int32 row, int32 column,
// double value each) for tests/test_gpu_cpp_dropin.py to compare.
template <class T>
static int api_check(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    int64_t hdr[3];
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    const int nrows = (int)hdr[0], ncols = (int)hdr[1], nnz = (int)hdr[2];
    std::vector<int> rp(nrows + 1), col(nnz);
    std::vector<double> v(nnz);
    f.read(reinterpret_cast<char*>(rp.data()), sizeof(int) * rp.size());
    f.read(reinterpret_cast<char*>(col.data()), sizeof(int) * col.size());
    f.read(reinterpret_cast<char*>(v.data()), sizeof(double) * v.size());
    if (!f) throw std::runtime_error("short file " + path);
    std::vector<Ijv<T>> ijv;
    for (int i = 0; i < nrows; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) ijv.push_back({i, col[p], T(v[p])});
    std::cout.precision(17);
    {
        SPC5Mat<T> csr = COO_sorted_to_CSR<T>(nrows, ncols, ijv.data(), nnz);
        auto best = SPC5_find_best<T>(csr);
        std::cout << "FINDBEST " << spc5_hip_detail::r_of(best.first) << " " << best.second << "\n";
        if constexpr (sizeof(T) == sizeof(double)) {
            for (int nth : {8, 36}) {
                auto bo = SPC5_find_best_omp<double>(csr, nth);
                std::cout << "FINDBEST_OMP " << nth << " " << spc5_hip_detail::r_of(bo.first) << " " << bo.second
                          << "\n";
            }
        }
    }
    const SPC5_MATRIX_TYPE fmts[4] = {FORMAT_1rVc_WT, FORMAT_2rVc, FORMAT_4rVc, FORMAT_8rVc};
    for (SPC5_MATRIX_TYPE fmt : fmts) {
        SPC5Mat<T> mat = COO_sorted_to_CSR<T>(nrows, ncols, ijv.data(), nnz);
        CSR_to_SPC5<T>(&mat, fmt);
        std::ofstream o(path + ".iter_r" + std::to_string(spc5_hip_detail::r_of(fmt)) + ".bin", std::ios::binary);
        long long count = 0;
        SPC5_iterate<T>(mat, [&](int row, int c, T val) {
            const double d = (double)val;
            o.write(reinterpret_cast<const char*>(&row), sizeof(int));
            o.write(reinterpret_cast<const char*>(&c), sizeof(int));
            o.write(reinterpret_cast<const char*>(&d), sizeof(double));
            ++count;
        });
        std::cout << "ITERATE " << spc5_hip_detail::r_of(fmt) << " " << count << "\n";
    }
    std::cout << "RESULT DONE" << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "usage: " << argv[0] << " <dense N | lap N | mm FILE | api FILE> <double|float>\n";
        return 2;
    }
    const std::string kind = argv[1], arg = argv[2], real = argv[3];
    try {
        if (kind == "api") return real == "double" ? api_check<double>(arg) : api_check<float>(arg);
        int nrows = 0, ncols = 0;
        if (real == "double") {
            std::vector<Ijv<double>> v;
            if (kind == "dense") {
                nrows = ncols = std::atoi(arg.c_str());
                v = build_dense<double>(nrows);
            } else if (kind == "lap") {
                const int n = std::atoi(arg.c_str());
                nrows = ncols = n * n;
                v = build_lap<double>(n);
            } else {
                v = load_mm<double>(arg, nrows, ncols);
            }
            return run<double>(std::move(v), nrows, ncols);
        }
        std::vector<Ijv<float>> v;
        if (kind == "dense") {
            nrows = ncols = std::atoi(arg.c_str());
            v = build_dense<float>(nrows);
        } else if (kind == "lap") {
            const int n = std::atoi(arg.c_str());
            nrows = ncols = n * n;
            v = build_lap<float>(n);
        } else {
            v = load_mm<float>(arg, nrows, ncols);
        }
        return run<float>(std::move(v), nrows, ncols);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
}
