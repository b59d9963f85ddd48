// ref_shim.cpp -- C entry points over the compiled reference header.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference's own src/spc5.hpp (included from where it lies, never
// copied).  Used by tests/ to pin the oracle restatement (converter bytes,
// SpMV bits, block counts, split rule, format model) and by bench.py's
// cpu_baseline leg to time the reference OpenMP AVX-512 path.
//
// The reference header defines its kernels and merge helpers with external
// linkage, so it is included in exactly this one translation unit
// (src/spc5.hpp:1495, 2157, 2309).  The library is built with hidden
// visibility: only the ref_* functions below are exported, so the
// reference's core_SPC5_* symbols never clash with the product library's.
#include "spc5.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <omp.h>

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

template <class T>
SPC5Mat<T> make_csr(int nrows, int ncols, int nnz, const int* rowptr, const int* col, const T* val) {
    SPC5Mat<T> m;
    m.format = SPC5_MATRIX_TYPE::FORMAT_CSR;
    m.numberOfRows = nrows;
    m.numberOfColumns = ncols;
    m.numberOfNNZ = nnz;
    m.values.reset(new T[nnz > 0 ? nnz : 1]);
    m.rowsSize.reset(new int[nrows + 1]);
    m.valuesColumnIndexes.reset(new int[nnz > 0 ? nnz : 1]);
    if (nnz > 0) {
        std::memcpy(m.values.get(), val, sizeof(T) * (size_t)nnz);
        std::memcpy(m.valuesColumnIndexes.get(), col, sizeof(int) * (size_t)nnz);
    }
    std::memcpy(m.rowsSize.get(), rowptr, sizeof(int) * (size_t)(nrows + 1));
    return m;
}

SPC5_MATRIX_TYPE fmt_of(int r) {
    switch (r) {
    case 1: return SPC5_MATRIX_TYPE::FORMAT_1rVc_WT;
    case 2: return SPC5_MATRIX_TYPE::FORMAT_2rVc;
    case 4: return SPC5_MATRIX_TYPE::FORMAT_4rVc;
    case 8: return SPC5_MATRIX_TYPE::FORMAT_8rVc;
    default: return SPC5_MATRIX_TYPE::UNDEFINED_FORMAT;
    }
}

template <class T>
int64_t convert(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col, const T* val,
                int* out_rows_size, unsigned char* out_hdr, int64_t hdr_cap, T* out_val) {
    SPC5Mat<T> m = make_csr<T>(nrows, ncols, nnz, rowptr, col, val);
    CSR_to_SPC5<T>(&m, fmt_of(r));
    const int64_t hsz = 4 + (int64_t)r * (int64_t)sizeof(typename SPC5Mat_Mask<T>::type);
    const int64_t nbytes = (int64_t)m.numberOfBlocks * hsz;
    if (nbytes > hdr_cap) return -1;
    std::memcpy(out_rows_size, m.rowsSize.get(), sizeof(int) * (size_t)(nrows + 1));
    if (nbytes > 0) std::memcpy(out_hdr, m.blocksColumnIndexesWithMasks.get(), (size_t)nbytes);
    if (nnz > 0) std::memcpy(out_val, m.values.get(), sizeof(T) * (size_t)nnz);
    return m.numberOfBlocks;
}

template <class T>
void spmv(int r, long nrows, const int* bptr, const unsigned char* hdr, const T* val, const T* x, T* y);

template <>
void spmv<double>(int r, long nrows, const int* bptr, const unsigned char* hdr, const double* val,
                  const double* x, double* y) {
    switch (r) {
    case 1: core_SPC5_1rVc_Spmv_double(nrows, bptr, hdr, val, x, y); break;
    case 2: core_SPC5_2rVc_Spmv_double(nrows, bptr, hdr, val, x, y); break;
    case 4: core_SPC5_4rVc_Spmv_double(nrows, bptr, hdr, val, x, y); break;
    case 8: core_SPC5_8rVc_Spmv_double(nrows, bptr, hdr, val, x, y); break;
    }
}

template <>
void spmv<float>(int r, long nrows, const int* bptr, const unsigned char* hdr, const float* val,
                 const float* x, float* y) {
    switch (r) {
    case 1: core_SPC5_1rVc_Spmv_float(nrows, bptr, hdr, val, x, y); break;
    case 2: core_SPC5_2rVc_Spmv_float(nrows, bptr, hdr, val, x, y); break;
    case 4: core_SPC5_4rVc_Spmv_float(nrows, bptr, hdr, val, x, y); break;
    case 8: core_SPC5_8rVc_Spmv_float(nrows, bptr, hdr, val, x, y); break;
    }
}

template <class T>
int block_count(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col, const T* val) {
    SPC5Mat<T> m = make_csr<T>(nrows, ncols, nnz, rowptr, col, val);
    switch (r) {
    case 1: return SPC5_1rVc_block_count<T>(m);
    case 2: return SPC5_2rVc_block_count<T>(m);
    case 4: return SPC5_4rVc_block_count<T>(m);
    case 8: return SPC5_8rVc_block_count<T>(m);
    }
    return -1;
}

// Split + OpenMP SpMV timing, as examples/load_mm_and_compare.cpp:656-729 does:
// convert, split once (untimed), then time nloops SPC5_Spmv_omp calls;
// repeated nsamples times, t_spmv[s] receiving sample s's seconds.
template <class T>
int omp_bench(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col, const T* val,
              const T* x, T* y, int nloops, int nsamples, int nthreads, double* t_convert, double* t_spmv) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    SPC5Mat<T> m = make_csr<T>(nrows, ncols, nnz, rowptr, col, val);
    auto t0 = std::chrono::steady_clock::now();
    CSR_to_SPC5<T>(&m, fmt_of(r));
    auto t1 = std::chrono::steady_clock::now();
    std::vector<ThreadInterval<T>> iv = SPC5_split_omp<T>(m);
    for (int s = 0; s < nsamples; ++s) {
        auto t2 = std::chrono::steady_clock::now();
        for (int l = 0; l < nloops; ++l) SPC5_Spmv_omp<T>(m, x, y, iv);
        auto t3 = std::chrono::steady_clock::now();
        t_spmv[s] = std::chrono::duration<double>(t3 - t2).count();
    }
    *t_convert = std::chrono::duration<double>(t1 - t0).count();
    return omp_get_max_threads();
}

}  // namespace

REF_EXPORT int64_t ref_convert_f64(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                   const double* val, int* out_rows_size, unsigned char* out_hdr,
                                   int64_t hdr_cap, double* out_val) {
    return convert<double>(r, nrows, ncols, nnz, rowptr, col, val, out_rows_size, out_hdr, hdr_cap, out_val);
}

REF_EXPORT int64_t ref_convert_f32(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                   const float* val, int* out_rows_size, unsigned char* out_hdr,
                                   int64_t hdr_cap, float* out_val) {
    return convert<float>(r, nrows, ncols, nnz, rowptr, col, val, out_rows_size, out_hdr, hdr_cap, out_val);
}

REF_EXPORT void ref_spmv_f64(int r, long nrows, const int* bptr, const unsigned char* hdr, const double* val,
                             const double* x, double* y) {
    spmv<double>(r, nrows, bptr, hdr, val, x, y);
}

REF_EXPORT void ref_spmv_f32(int r, long nrows, const int* bptr, const unsigned char* hdr, const float* val,
                             const float* x, float* y) {
    spmv<float>(r, nrows, bptr, hdr, val, x, y);
}

REF_EXPORT void ref_csr_spmv_f64(int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                 const double* val, const double* x, double* y) {
    SPC5Mat<double> m = make_csr<double>(nrows, ncols, nnz, rowptr, col, val);
    CSR_Spmv_scalar<double>(m, x, y);
}

REF_EXPORT void ref_csr_spmv_f32(int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                 const float* val, const float* x, float* y) {
    SPC5Mat<float> m = make_csr<float>(nrows, ncols, nnz, rowptr, col, val);
    CSR_Spmv_scalar<float>(m, x, y);
}

REF_EXPORT int ref_block_count_f64(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                   const double* val) {
    return block_count<double>(r, nrows, ncols, nnz, rowptr, col, val);
}

REF_EXPORT int ref_block_count_f32(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                   const float* val) {
    return block_count<float>(r, nrows, ncols, nnz, rowptr, col, val);
}

// SPC5_find_best<T> (src/spc5.hpp:1151) and SPC5_find_best_omp<double> (1264).
// Returns r of the chosen format (0 for UNDEFINED) and writes the estimate.
REF_EXPORT int ref_find_best(int is_double, int nthreads, int nrows, int ncols, int nnz, const int* rowptr,
                             const int* col, const double* val64, const float* val32, double* estimate) {
    std::pair<SPC5_MATRIX_TYPE, double> res;
    if (is_double) {
        SPC5Mat<double> m = make_csr<double>(nrows, ncols, nnz, rowptr, col, val64);
        res = nthreads > 0 ? SPC5_find_best_omp<double>(m, nthreads) : SPC5_find_best<double>(m);
    } else {
        SPC5Mat<float> m = make_csr<float>(nrows, ncols, nnz, rowptr, col, val32);
        res = SPC5_find_best<float>(m);
    }
    *estimate = res.second;
    switch (res.first) {
    case SPC5_MATRIX_TYPE::FORMAT_1rVc_WT: return 1;
    case SPC5_MATRIX_TYPE::FORMAT_2rVc: return 2;
    case SPC5_MATRIX_TYPE::FORMAT_4rVc: return 4;
    case SPC5_MATRIX_TYPE::FORMAT_8rVc: return 8;
    default: return 0;
    }
}

// core_SPC5_rVc_threadsplit (src/spc5.hpp:346-427) on a converted FP64 matrix.
REF_EXPORT void ref_thread_split_f64(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                     const double* val, int nthreads, int* start_row, int* nrows_out,
                                     int* value_offset) {
    SPC5Mat<double> m = make_csr<double>(nrows, ncols, nnz, rowptr, col, val);
    CSR_to_SPC5<double>(&m, fmt_of(r));
    std::vector<ThreadInterval<double>> iv;
    switch (r) {
    case 1: iv = SPC5_1rVc_split_omp<double>(m, nthreads); break;
    case 2: iv = SPC5_2rVc_split_omp<double>(m, nthreads); break;
    case 4: iv = SPC5_4rVc_split_omp<double>(m, nthreads); break;
    case 8: iv = SPC5_8rVc_split_omp<double>(m, nthreads); break;
    }
    for (int t = 0; t < nthreads; ++t) {
        start_row[t] = iv[t].startingRow;
        nrows_out[t] = iv[t].numberOfRows;
        value_offset[t] = iv[t].valuesOffset;
    }
}

REF_EXPORT int ref_omp_bench_f64(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                 const double* val, const double* x, double* y, int nloops, int nsamples, int nthreads,
                                 double* t_convert, double* t_spmv) {
    return omp_bench<double>(r, nrows, ncols, nnz, rowptr, col, val, x, y, nloops, nsamples, nthreads, t_convert,
                          t_spmv);
}

REF_EXPORT int ref_omp_bench_f32(int r, int nrows, int ncols, int nnz, const int* rowptr, const int* col,
                                 const float* val, const float* x, float* y, int nloops, int nsamples, int nthreads,
                                 double* t_convert, double* t_spmv) {
    return omp_bench<float>(r, nrows, ncols, nnz, rowptr, col, val, x, y, nloops, nsamples, nthreads, t_convert,
                          t_spmv);
}
