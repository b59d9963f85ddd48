// spc5_hip.hpp -- drop-in replacement for the reference's spc5.hpp
// (berenger-eu/spc5 src/spc5.hpp) backed by libspc5_hip.so on an MI355X.
//
// Same types, names, arguments, in-place semantics and exceptions:
//   * SPC5Mat<T>, SPC5_MATRIX_TYPE, SPC5_VEC_PADDING, SPC5Mat_Mask, Ijv,
//     ValPerVec, ToUniquePtr                         (src/spc5.hpp:17-62, 145-157)
//   * COO_sorted_to_CSR / COO_unsorted_to_CSR       (src/spc5.hpp:77-114)
//   * CSR_to_SPC5_<r>rVc, COO_to_SPC5_<r>rVc, CSR_to_SPC5 -- the conversion
//     runs on the GPU and leaves the matrix exactly as the reference does
//     (rowsSize keeps N+1 entries with its stale tail, values replaced,
//     valuesColumnIndexes untouched)                (src/spc5.hpp:165-228, 455-488, 989-1018)
//   * SPC5_<r>rVc_block_count, SPC5_block_count    (src/spc5.hpp:276-317, 1120-1148)
//   * SPC5_<r>rVc_Spmv<T>, SPC5_Spmv: y += A x on the GPU through the
//     reference C ABI exported by the library       (src/spc5.hpp:525-542, 1090-1118)
//   * SPC5_<r>rVc_split_omp / SPC5_<r>rVc_Spmv_omp / SPC5_split_omp /
//     SPC5_Spmv_omp: the "prepare once, multiply many times" pair; on the
//     GPU the split uploads the matrix and builds its tile plan, the product
//     runs on the device copy                         (src/spc5.hpp:544-602, 1307-1367)
//   * SPC5_find_best, SPC5_find_best_omp<double>, SPC5_type_to_string
//                                                     (src/spc5.hpp:1151-1304)
//   * SPC5_<r>rVc_Spmv_scalar / SPC5_<r>rVc_iterate / SPC5_Spmv_scalar /
//     SPC5_iterate with the semantics the reference intends (its versions
//     index rowsSize and the masks wrongly for r > 1): the scalar product is
//     served by the GPU kernel, iterate walks the host arrays.
//
// Differences a caller can observe: SPC5_block_count returns the count (the
// reference discards it); GPU failures throw std::runtime_error carrying
// spc5_hip_last_error(); the header may be included in any number of
// translation units.  Link with -lspc5_hip.
#ifndef SPC5_HIP_HPP
#define SPC5_HIP_HPP

#include <algorithm>
#include <cassert>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "spc5_hip.h"

//////////////////////////////////////////////////////////////////////////////
/// Matrix structure
//////////////////////////////////////////////////////////////////////////////

enum SPC5_MATRIX_TYPE { UNDEFINED_FORMAT, FORMAT_CSR, FORMAT_1rVc_WT, FORMAT_2rVc, FORMAT_4rVc, FORMAT_8rVc };

enum SPC5_VEC_PADDING { SPC5_VEC_PADDING_X = 15, SPC5_VEC_PADDING_Y = 7 };

template <class ValueType>
struct SPC5Mat_Mask;
template <>
struct SPC5Mat_Mask<double> {
    using type = unsigned char;
};
template <>
struct SPC5Mat_Mask<float> {
    using type = unsigned short;
};

template <class ValueType>
struct SPC5Mat {
    SPC5_MATRIX_TYPE format;
    int numberOfRows;
    int numberOfColumns;
    int numberOfNNZ;
    std::unique_ptr<ValueType[]> values;
    std::unique_ptr<int[]> rowsSize;
    std::unique_ptr<int[]> valuesColumnIndexes;
    std::unique_ptr<unsigned char[]> blocksColumnIndexesWithMasks;
    int numberOfBlocks;
    SPC5Mat() : format(UNDEFINED_FORMAT), numberOfRows(0), numberOfColumns(0), numberOfNNZ(0), numberOfBlocks(0) {}
};

template <class ValueType>
struct Ijv {
    int i;
    int j;
    ValueType v;
};

template <class ValueType>
struct ValPerVec {
    static const int size = int(512 / (8 * sizeof(ValueType)));
};

template <class ValueType>
std::unique_ptr<ValueType[]> ToUniquePtr(const std::vector<ValueType>& values) {
    std::unique_ptr<ValueType[]> ptr(new ValueType[values.size()]);
    std::copy(values.begin(), values.end(), ptr.get());
    return ptr;
}

namespace spc5_hip_detail {

inline void check(int rc, const char* what) {
    if (rc != SPC5_HIP_OK) throw std::runtime_error(std::string(what) + ": " + spc5_hip_last_error());
}

inline int r_of(SPC5_MATRIX_TYPE t) {
    switch (t) {
    case FORMAT_1rVc_WT: return 1;
    case FORMAT_2rVc: return 2;
    case FORMAT_4rVc: return 4;
    case FORMAT_8rVc: return 8;
    default: return 0;
    }
}

inline SPC5_MATRIX_TYPE format_of(int r) {
    switch (r) {
    case 1: return FORMAT_1rVc_WT;
    case 2: return FORMAT_2rVc;
    case 4: return FORMAT_4rVc;
    case 8: return FORMAT_8rVc;
    default: return UNDEFINED_FORMAT;
    }
}

template <class ValueType>
constexpr int mask_bytes() {
    return (int)sizeof(typename SPC5Mat_Mask<ValueType>::type);
}

template <class ValueType>
constexpr int value_type_code() {
    return sizeof(ValueType) == 8 ? SPC5_HIP_F64 : SPC5_HIP_F32;
}

// In-place CSR -> beta(r,VS) conversion on the GPU (src/spc5.hpp:165-228).
template <class ValueType>
void convert(SPC5Mat<ValueType>* csr, int r) {
    assert(csr->format == FORMAT_CSR);
    const int vs = ValPerVec<ValueType>::size;
    const int mb = mask_bytes<ValueType>();
    int64_t nb = 0;
    check(spc5_hip_host_block_count(vs, r, csr->numberOfRows, csr->numberOfNNZ, csr->rowsSize.get(),
                                    csr->valuesColumnIndexes.get(), &nb),
          "CSR_to_SPC5 (block count)");
    const int64_t hbytes = nb * (4 + (int64_t)r * mb);
    std::unique_ptr<unsigned char[]> hdr(new unsigned char[hbytes > 0 ? hbytes : 1]);
    std::unique_ptr<ValueType[]> vals(new ValueType[csr->numberOfNNZ > 0 ? csr->numberOfNNZ : 1]);
    const int nbr = (csr->numberOfRows + r - 1) / r;
    std::vector<int> bptr((size_t)nbr + 1);
    int64_t nb2 = 0;
    check(spc5_hip_host_convert((int)sizeof(ValueType), vs, mb, r, csr->numberOfRows, csr->numberOfNNZ,
                                csr->rowsSize.get(), csr->valuesColumnIndexes.get(), csr->values.get(), bptr.data(),
                                hdr.get(), hbytes, vals.get(), &nb2),
          "CSR_to_SPC5");
    std::copy(bptr.begin(), bptr.end(), csr->rowsSize.get());  // the tail past ceil(N/r) stays as it was
    csr->values = std::move(vals);
    csr->blocksColumnIndexesWithMasks = std::move(hdr);
    csr->numberOfBlocks = (int)nb2;
    csr->format = format_of(r);
}

template <class ValueType>
int block_count(const SPC5Mat<ValueType>& csr, int r) {
    int64_t nb = 0;
    check(spc5_hip_host_block_count(ValPerVec<ValueType>::size, r, csr.numberOfRows, csr.numberOfNNZ,
                                    csr.rowsSize.get(), csr.valuesColumnIndexes.get(), &nb),
          "SPC5_block_count");
    return (int)nb;
}

template <class ValueType>
void spmv(const SPC5Mat<ValueType>& mat, int r, const ValueType x[], ValueType y[]);

template <>
inline void spmv<double>(const SPC5Mat<double>& m, int r, const double x[], double y[]) {
    const int* rs = m.rowsSize.get();
    const unsigned char* h = m.blocksColumnIndexesWithMasks.get();
    switch (r) {
    case 1: core_SPC5_1rVc_Spmv_double(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    case 2: core_SPC5_2rVc_Spmv_double(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    case 4: core_SPC5_4rVc_Spmv_double(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    case 8: core_SPC5_8rVc_Spmv_double(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    }
}

template <>
inline void spmv<float>(const SPC5Mat<float>& m, int r, const float x[], float y[]) {
    const int* rs = m.rowsSize.get();
    const unsigned char* h = m.blocksColumnIndexesWithMasks.get();
    switch (r) {
    case 1: core_SPC5_1rVc_Spmv_float(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    case 2: core_SPC5_2rVc_Spmv_float(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    case 4: core_SPC5_4rVc_Spmv_float(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    case 8: core_SPC5_8rVc_Spmv_float(m.numberOfRows, rs, h, m.values.get(), x, y); break;
    }
}

// Walks a beta(r,VS) matrix in stream order: func(row, column, value).
template <class ValueType, int R, class FuncType>
void iterate(const SPC5Mat<ValueType>& mat, FuncType&& func) {
    using Mask = typename SPC5Mat_Mask<ValueType>::type;
    const int hsz = 4 + R * (int)sizeof(Mask);
    int idxVal = 0;
    for (int row0 = 0; row0 < mat.numberOfRows; row0 += R) {
        const int br = row0 / R;
        for (int b = mat.rowsSize[br]; b < mat.rowsSize[br + 1]; ++b) {
            const unsigned char* h = &mat.blocksColumnIndexesWithMasks[(size_t)b * hsz];
            int col;
            std::memcpy(&col, h, sizeof(int));
            for (int s = 0; s < R; ++s) {
                Mask m;
                std::memcpy(&m, h + sizeof(int) + s * sizeof(Mask), sizeof(Mask));
                for (int k = 0; k < ValPerVec<ValueType>::size; ++k)
                    if ((1u << k) & m) func(row0 + s, col + k, mat.values[idxVal++]);
            }
        }
    }
}

}  // namespace spc5_hip_detail

//////////////////////////////////////////////////////////////////////////////
/// CSR
//////////////////////////////////////////////////////////////////////////////

template <class ValueType>
inline SPC5Mat<ValueType> COO_sorted_to_CSR(const int nbRows, const int nbCols, const Ijv<ValueType> values[],
                                            int nbValues) {
    SPC5Mat<ValueType> csr;
    csr.format = FORMAT_CSR;
    csr.numberOfRows = nbRows;
    csr.numberOfColumns = nbCols;
    csr.numberOfNNZ = nbValues;
    csr.values.reset(new ValueType[nbValues > 0 ? nbValues : 1]());
    csr.rowsSize.reset(new int[nbRows + 1]());
    csr.valuesColumnIndexes.reset(new int[nbValues > 0 ? nbValues : 1]());
    for (int e = 0; e < nbValues; ++e) {
        csr.values[e] = values[e].v;
        csr.valuesColumnIndexes[e] = values[e].j;
        csr.rowsSize[values[e].i + 1] += 1;
    }
    for (int i = 2; i <= nbRows; ++i) csr.rowsSize[i] += csr.rowsSize[i - 1];
    return csr;
}

template <class ValueType>
inline SPC5Mat<ValueType> COO_unsorted_to_CSR(const int nbRows, const int nbCols, Ijv<ValueType> values[],
                                              int nbValues) {
    std::sort(&values[0], &values[nbValues], [](const Ijv<ValueType>& a, const Ijv<ValueType>& b) {
        return a.i < b.i || (a.i == b.i && a.j < b.j);
    });
    return COO_sorted_to_CSR(nbRows, nbCols, values, nbValues);
}

// The reference's host CSR product (src/spc5.hpp:116-127), kept as the CSR
// check routine callers use next to the beta(r,VS) formats; it is not a
// path of the beta(r,VS) product.
template <class ValueType>
inline void CSR_Spmv_scalar(const SPC5Mat<ValueType>& csr, const ValueType x[], ValueType y[]) {
    for (int i = 0; i < csr.numberOfRows; ++i) {
        ValueType sum = 0;
        for (int p = csr.rowsSize[i]; p < csr.rowsSize[i + 1]; ++p) sum += x[csr.valuesColumnIndexes[p]] * csr.values[p];
        y[i] += sum;
    }
}

template <class ValueType, class FuncType>
inline void CSR_iterate(SPC5Mat<ValueType>& csr, const FuncType&& func) {
    for (int i = 0; i < csr.numberOfRows; ++i)
        for (int p = csr.rowsSize[i]; p < csr.rowsSize[i + 1]; ++p) func(i, csr.valuesColumnIndexes[p], csr.values[p]);
}

//////////////////////////////////////////////////////////////////////////////
/// Per-format API (1, 2, 4, 8 rows per block)
//////////////////////////////////////////////////////////////////////////////

#define SPC5_HIP_FORMAT_API(R)                                                                               \
    template <class ValueType>                                                                               \
    inline void CSR_to_SPC5_##R##rVc(SPC5Mat<ValueType>* csr) {                                              \
        spc5_hip_detail::convert<ValueType>(csr, R);                                                         \
    }                                                                                                        \
    template <class ValueType>                                                                               \
    inline SPC5Mat<ValueType> COO_to_SPC5_##R##rVc(const int nbRows, const int nbCols, Ijv<ValueType> values[], \
                                                   int nbValues) {                                           \
        SPC5Mat<ValueType> mat = COO_unsorted_to_CSR(nbRows, nbCols, values, nbValues);                      \
        CSR_to_SPC5_##R##rVc(&mat);                                                                          \
        return mat;                                                                                          \
    }                                                                                                        \
    template <class ValueType>                                                                               \
    inline int SPC5_##R##rVc_block_count(const SPC5Mat<ValueType>& csr) {                                    \
        return spc5_hip_detail::block_count<ValueType>(csr, R);                                              \
    }                                                                                                        \
    template <class ValueType>                                                                               \
    inline void SPC5_##R##rVc_Spmv(const SPC5Mat<ValueType>& mat, const ValueType x[], ValueType y[]) {      \
        assert(mat.format == spc5_hip_detail::format_of(R));                                                 \
        spc5_hip_detail::spmv<ValueType>(mat, R, x, y);                                                      \
    }                                                                                                        \
    template <class ValueType>                                                                               \
    inline void SPC5_##R##rVc_Spmv_scalar(const SPC5Mat<ValueType>& mat, const ValueType x[], ValueType y[]) { \
        SPC5_##R##rVc_Spmv<ValueType>(mat, x, y);                                                            \
    }                                                                                                        \
    template <class ValueType, class FuncType>                                                               \
    inline void SPC5_##R##rVc_iterate(const SPC5Mat<ValueType>& mat, FuncType&& func) {                      \
        spc5_hip_detail::iterate<ValueType, R>(mat, std::forward<FuncType>(func));                           \
    }

SPC5_HIP_FORMAT_API(1)
SPC5_HIP_FORMAT_API(2)
SPC5_HIP_FORMAT_API(4)
SPC5_HIP_FORMAT_API(8)
#undef SPC5_HIP_FORMAT_API

//////////////////////////////////////////////////////////////////////////////
/// Wrappers
//////////////////////////////////////////////////////////////////////////////

template <class ValueType>
inline void CSR_to_SPC5(SPC5Mat<ValueType>* mat, const SPC5_MATRIX_TYPE matType) {
    const int r = spc5_hip_detail::r_of(matType);
    if (r == 0) throw std::invalid_argument("CSR_to_SPC5 : Unknown format type");
    spc5_hip_detail::convert<ValueType>(mat, r);
}

template <class ValueType>
inline void SPC5_Spmv(const SPC5Mat<ValueType>& mat, const ValueType x[], ValueType y[]) {
    const int r = spc5_hip_detail::r_of(mat.format);
    if (r == 0) throw std::invalid_argument("SPC5_Spmv : Unknown format type");
    spc5_hip_detail::spmv<ValueType>(mat, r, x, y);
}

template <class ValueType>
inline void SPC5_Spmv_scalar(const SPC5Mat<ValueType>& mat, const ValueType x[], ValueType y[]) {
    if (mat.format == FORMAT_CSR) {
        CSR_Spmv_scalar<ValueType>(mat, x, y);
        return;
    }
    const int r = spc5_hip_detail::r_of(mat.format);
    if (r == 0) throw std::invalid_argument("SPC5_Spmv_scalar : Unknown format type");
    spc5_hip_detail::spmv<ValueType>(mat, r, x, y);
}

template <class ValueType, class FuncType>
inline void SPC5_iterate(const SPC5Mat<ValueType>& mat, FuncType&& func) {
    switch (mat.format) {
    case FORMAT_1rVc_WT: spc5_hip_detail::iterate<ValueType, 1>(mat, std::forward<FuncType>(func)); break;
    case FORMAT_2rVc: spc5_hip_detail::iterate<ValueType, 2>(mat, std::forward<FuncType>(func)); break;
    case FORMAT_4rVc: spc5_hip_detail::iterate<ValueType, 4>(mat, std::forward<FuncType>(func)); break;
    case FORMAT_8rVc: spc5_hip_detail::iterate<ValueType, 8>(mat, std::forward<FuncType>(func)); break;
    case FORMAT_CSR:
        for (int i = 0; i < mat.numberOfRows; ++i)
            for (int p = mat.rowsSize[i]; p < mat.rowsSize[i + 1]; ++p)
                func(i, mat.valuesColumnIndexes[p], mat.values[p]);
        break;
    default: throw std::invalid_argument("SPC5_iterate : Unknown format type");
    }
}

template <class ValueType>
inline int SPC5_block_count(const SPC5Mat<ValueType>& csr, const SPC5_MATRIX_TYPE matType) {
    const int r = spc5_hip_detail::r_of(matType);
    if (r == 0) throw std::invalid_argument("SPC5_Spmv : Unknown format type");
    return spc5_hip_detail::block_count<ValueType>(csr, r);
}

// Format choice from the reference's CPU performance model (cubic in the
// mean number of values per block, r = 1, 2, 4; src/spc5.hpp:1151-1223).
template <class ValueType>
std::pair<SPC5_MATRIX_TYPE, double> SPC5_find_best(const SPC5Mat<ValueType>& csr) {
    static const double kD[3][4] = {{2.186490e-03, -7.931533e-02, 8.719444e-01, 3.712380e-01},
                                    {9.822706e-04, -4.500002e-02, 6.714119e-01, 2.413985e-01},
                                    {1.320496e-04, -1.222130e-02, 3.520409e-01, 3.603195e-01}};
    static const double kF[3][4] = {{-4.039852e-03, 7.997606e-02, 6.892119e-02, 1.508645e+00},
                                    {-9.195940e-04, 4.442050e-03, 5.273045e-01, 1.101610e+00},
                                    {-4.464933e-05, 1.867266e-03, 1.520910e-01, 1.186131e+00}};
    const bool dbl = sizeof(ValueType) == sizeof(double);
    SPC5_MATRIX_TYPE best = UNDEFINED_FORMAT;
    double est = 0;
    const int rs[3] = {1, 2, 4};
    for (int i = 0; i < 3; ++i) {
        const double a = double(csr.numberOfNNZ) / double(spc5_hip_detail::block_count<ValueType>(csr, rs[i]));
        const double* c = dbl ? kD[i] : kF[i];
        const double s = c[0] * (a * a * a) + c[1] * (a * a) + c[2] * a + c[3];
        if (s > est) {
            est = s;
            best = spc5_hip_detail::format_of(rs[i]);
        }
    }
    return {best, est};
}

inline const char* SPC5_type_to_string(const SPC5_MATRIX_TYPE matType) {
    switch (matType) {
    case FORMAT_1rVc_WT: return "1rVc_WT";
    case FORMAT_2rVc: return "2rVc";
    case FORMAT_4rVc: return "4rVc";
    case FORMAT_8rVc: return "8rVc";
    default: return "undefined";
    }
}

//////////////////////////////////////////////////////////////////////////////
/// Prepared repeated products (the reference's OpenMP API)
//////////////////////////////////////////////////////////////////////////////

// Same fields as the reference (src/spc5.hpp:331-343); the per-thread
// copies stay empty: the prepared state is the device copy in `device`.
template <class ValueType>
struct ThreadInterval {
    int startingRow = 0;
    int numberOfRows = 0;
    std::unique_ptr<ValueType[]> threadY;
    std::unique_ptr<ValueType[]> threadValues;
    std::unique_ptr<int[]> threadRowsSize;
    std::unique_ptr<int[]> threadValuesColumnIndexes;
    std::unique_ptr<unsigned char[]> threadBlocksColumnIndexesWithMasks;
    int valuesOffset = 0;
    std::shared_ptr<spc5_hip_mat> device;
};

namespace spc5_hip_detail {

template <class ValueType>
std::vector<ThreadInterval<ValueType>> prepare(const SPC5Mat<ValueType>& mat, int r) {
    spc5_hip_mat_t h = nullptr;
    check(spc5_hip_mat_create_from_host(value_type_code<ValueType>(), ValPerVec<ValueType>::size, r,
                                        mat.numberOfRows, mat.numberOfColumns, mat.numberOfNNZ, mat.numberOfBlocks,
                                        mat.rowsSize.get(), mat.blocksColumnIndexesWithMasks.get(),
                                        mat.values.get(), &h),
          "SPC5_split_omp");
    std::vector<ThreadInterval<ValueType>> iv(1);
    iv[0].startingRow = 0;
    iv[0].numberOfRows = mat.numberOfRows;
    iv[0].device.reset(h, [](spc5_hip_mat* p) { spc5_hip_mat_destroy(p); });
    return iv;
}

template <class ValueType>
void run_prepared(const std::vector<ThreadInterval<ValueType>>& iv, const ValueType x[], ValueType y[]) {
    if (iv.empty() || !iv[0].device) throw std::invalid_argument("SPC5_Spmv_omp : intervals were not prepared");
    check(spc5_hip_spmv_host(iv[0].device.get(), x, y), "SPC5_Spmv_omp");
}

}  // namespace spc5_hip_detail

#define SPC5_HIP_PREPARED_API(R)                                                                              \
    template <class ValueType>                                                                                \
    inline std::vector<ThreadInterval<ValueType>> SPC5_##R##rVc_split_omp(const SPC5Mat<ValueType>& mat,      \
                                                                           const int /*numThreads*/) {        \
        return spc5_hip_detail::prepare<ValueType>(mat, R);                                                   \
    }                                                                                                         \
    template <class ValueType>                                                                                \
    inline std::vector<ThreadInterval<ValueType>> SPC5_##R##rVc_split_omp(const SPC5Mat<ValueType>& mat) {    \
        return spc5_hip_detail::prepare<ValueType>(mat, R);                                                   \
    }                                                                                                         \
    template <class ValueType>                                                                                \
    inline void SPC5_##R##rVc_Spmv_omp(const SPC5Mat<ValueType>& mat, const ValueType x[], ValueType y[],     \
                                       const std::vector<ThreadInterval<ValueType>>& threadsVecs) {           \
        assert(mat.format == spc5_hip_detail::format_of(R));                                                  \
        (void)mat;                                                                                            \
        spc5_hip_detail::run_prepared<ValueType>(threadsVecs, x, y);                                          \
    }

SPC5_HIP_PREPARED_API(1)
SPC5_HIP_PREPARED_API(2)
SPC5_HIP_PREPARED_API(4)
SPC5_HIP_PREPARED_API(8)
#undef SPC5_HIP_PREPARED_API

template <class ValueType>
inline std::vector<ThreadInterval<ValueType>> SPC5_split_omp(const SPC5Mat<ValueType>& mat) {
    const int r = spc5_hip_detail::r_of(mat.format);
    if (r == 0) throw std::invalid_argument("SPC5_Spmv : Unknown format type");
    return spc5_hip_detail::prepare<ValueType>(mat, r);
}

template <class ValueType>
inline void SPC5_Spmv_omp(const SPC5Mat<ValueType>& mat, const ValueType x[], ValueType y[],
                          const std::vector<ThreadInterval<ValueType>>& threadsVecs) {
    if (spc5_hip_detail::r_of(mat.format) == 0) throw std::invalid_argument("SPC5_Spmv : Unknown format type");
    spc5_hip_detail::run_prepared<ValueType>(threadsVecs, x, y);
}

// The reference's OpenMP format model (bilinear in threads and values per
// block, double only; src/spc5.hpp:1264-1304).
template <class ValueType>
std::pair<SPC5_MATRIX_TYPE, double> SPC5_find_best_omp(const SPC5Mat<ValueType>& csr, const int nbThreads);

template <>
inline std::pair<SPC5_MATRIX_TYPE, double> SPC5_find_best_omp<double>(const SPC5Mat<double>& csr,
                                                                       const int nbThreads) {
    static const double kC[3][4] = {{4.3764, 0.0981, 1.8728, 0.0721},
                                    {5.2172, 0.1435, 0.8331, 0.0556},
                                    {5.3212, 0.1921, 0.5201, 0.0263}};
    SPC5_MATRIX_TYPE best = UNDEFINED_FORMAT;
    double est = 0;
    const int rs[3] = {1, 2, 4};
    for (int i = 0; i < 3; ++i) {
        const double a = double(csr.numberOfNNZ) / double(spc5_hip_detail::block_count<double>(csr, rs[i]));
        const double s = kC[i][0] + kC[i][1] * nbThreads + kC[i][2] * a + kC[i][3] * nbThreads * a;
        if (s > est) {
            est = s;
            best = spc5_hip_detail::format_of(rs[i]);
        }
    }
    return {best, est};
}

template <class ValueType>
void SPC5_opti_merge(ValueType dest[], const ValueType src[], const int nbValues);
template <>
inline void SPC5_opti_merge<double>(double dest[], const double src[], const int nbValues) {
    SPC5_opti_merge_double(dest, src, nbValues);
}
template <>
inline void SPC5_opti_merge<float>(float dest[], const float src[], const int nbValues) {
    SPC5_opti_merge_float(dest, src, nbValues);
}

#endif  // SPC5_HIP_HPP
