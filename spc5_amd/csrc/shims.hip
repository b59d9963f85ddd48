// shims.hip -- the reference's C ABI on host pointers, executed on the GPU.
//
// core_SPC5_<r>rVc_Spmv_<T>(nbRows, rowsSizes, headers, values, x, y) keeps
// the reference contract (src/spc5.hpp:516-523, 641-648, 769-776, 895-902):
// y += A x, rowsSizes indexed by block-row with only differences used, the
// header and value streams read from their first element.  The call carries
// no sizes, so they are recovered from the arrays: the block count from
// rowsSizes, nnz from the mask popcounts and the extent of x from the largest
// column any block touches.  Inputs are copied to the GPU, the tile kernel
// runs, y is copied back.  The interface has no error channel: a GPU failure
// prints the reason and aborts.
#include "spc5_mat.hpp"

#include <cstdlib>
#include <vector>

namespace spc5 {
namespace {

[[noreturn]] void die(const char* where, const std::string& msg) {
    std::fprintf(stderr, "spc5_hip: %s failed: %s\n", where, msg.c_str());
    std::fflush(stderr);
    std::abort();
}

#define SHIM_HIP(expr)                                               \
    do {                                                             \
        hipError_t e__ = (expr);                                     \
        if (e__ != hipSuccess) die(fn, hip_msg(e__, #expr));         \
    } while (0)

template <class T>
void host_spmv(const char* fn, int r, long nbRows, const int* rowsSizes, const unsigned char* headers,
               const T* values, const T* x, T* y) {
    constexpr bool kDouble = sizeof(T) == 8;
    constexpr int vs = kDouble ? 8 : 16;
    constexpr int mb = kDouble ? 1 : 2;
    if (nbRows <= 0) return;
    const int hsz = 4 + r * mb;
    const long nbr = (nbRows + r - 1) / r;
    const int base = rowsSizes[0];
    const long nb = (long)rowsSizes[nbr] - base;
    if (nb <= 0) return;
    std::vector<int> bptr((size_t)nbr + 1);
    for (long i = 0; i <= nbr; ++i) bptr[(size_t)i] = rowsSizes[i] - base;
    long long nnz = 0, maxcol = -1;
    for (long b = 0; b < nb; ++b) {
        const unsigned char* h = headers + (size_t)b * hsz;
        int cs;
        std::memcpy(&cs, h, 4);
        unsigned rowor = 0;
        for (int s = 0; s < r; ++s) {
            unsigned m = 0;
            std::memcpy(&m, h + 4 + s * mb, (size_t)mb);
            nnz += __builtin_popcount(m);
            rowor |= m;
        }
        if (rowor) {
            const long long hi = (long long)cs + (31 - __builtin_clz(rowor));
            if (hi > maxcol) maxcol = hi;
        }
    }
    if (nnz == 0) return;
    const long long nx = maxcol + 1;
    DeviceBuffer d_bptr, d_hdr, d_val, d_x, d_y;
    SHIM_HIP(d_bptr.alloc(sizeof(int) * (size_t)(nbr + 1)));
    SHIM_HIP(d_hdr.alloc((size_t)nb * hsz));
    SHIM_HIP(d_val.alloc(sizeof(T) * (size_t)nnz));
    SHIM_HIP(d_x.alloc(sizeof(T) * (size_t)nx));
    SHIM_HIP(d_y.alloc(sizeof(T) * (size_t)nbRows));
    SHIM_HIP(hipMemcpy(d_bptr.p, bptr.data(), sizeof(int) * (size_t)(nbr + 1), hipMemcpyHostToDevice));
    SHIM_HIP(hipMemcpy(d_hdr.p, headers, (size_t)nb * hsz, hipMemcpyHostToDevice));
    SHIM_HIP(hipMemcpy(d_val.p, values, sizeof(T) * (size_t)nnz, hipMemcpyHostToDevice));
    SHIM_HIP(hipMemcpy(d_x.p, x, sizeof(T) * (size_t)nx, hipMemcpyHostToDevice));
    SHIM_HIP(hipMemcpy(d_y.p, y, sizeof(T) * (size_t)nbRows, hipMemcpyHostToDevice));
    spc5_hip_mat_t h = nullptr;
    int rc = spc5_hip_mat_create(kDouble ? SPC5_HIP_F64 : SPC5_HIP_F32, vs, r, nbRows, nx, nnz, nb, d_bptr.as<int>(),
                                 d_hdr.as<unsigned char>(), d_val.p, nullptr, &h);
    if (rc) die(fn, spc5_hip_last_error());
    rc = spc5_hip_spmv(h, d_x.p, d_y.p, nullptr);
    if (rc) die(fn, spc5_hip_last_error());
    spc5_hip_mat_destroy(h);
    SHIM_HIP(hipMemcpy(y, d_y.p, sizeof(T) * (size_t)nbRows, hipMemcpyDeviceToHost));
}

}  // namespace
}  // namespace spc5

using spc5::host_spmv;

extern "C" {

void core_SPC5_1rVc_Spmv_double(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                                const double* values, const double* x, double* y) {
    host_spmv<double>(__func__, 1, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_2rVc_Spmv_double(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                                const double* values, const double* x, double* y) {
    host_spmv<double>(__func__, 2, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_4rVc_Spmv_double(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                                const double* values, const double* x, double* y) {
    host_spmv<double>(__func__, 4, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_8rVc_Spmv_double(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                                const double* values, const double* x, double* y) {
    host_spmv<double>(__func__, 8, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_1rVc_Spmv_float(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                               const float* values, const float* x, float* y) {
    host_spmv<float>(__func__, 1, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_2rVc_Spmv_float(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                               const float* values, const float* x, float* y) {
    host_spmv<float>(__func__, 2, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_4rVc_Spmv_float(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                               const float* values, const float* x, float* y) {
    host_spmv<float>(__func__, 4, nbRows, rowsSizes, headers, values, x, y);
}
void core_SPC5_8rVc_Spmv_float(const long int nbRows, const int* rowsSizes, const unsigned char* headers,
                               const float* values, const float* x, float* y) {
    host_spmv<float>(__func__, 8, nbRows, rowsSizes, headers, values, x, y);
}

// Host helper of the reference's OpenMP path: dest += src.  There is no
// private y to merge on the GPU; the function keeps its host semantics for
// callers that use it directly (src/spc5.hpp:2157-2175).
void SPC5_opti_merge_double(double dest[], const double src[], const int nbValues) {
    for (int i = 0; i < nbValues; ++i) dest[i] += src[i];
}
void SPC5_opti_merge_float(float dest[], const float src[], const int nbValues) {
    for (int i = 0; i < nbValues; ++i) dest[i] += src[i];
}

}  // extern "C"
