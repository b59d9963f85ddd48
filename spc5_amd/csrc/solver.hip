// solver.hip -- conjugate gradient and mixed-precision iterative refinement
// on beta(r,VS) matrices (the precision form; no counterpart in the
// reference).
//
// CG iteration, scalars kept on the device (no host round trip):
//   K1  q = A p            (SpMV, y = A x epilogue; each tile writes its
//                           partial of p.q) -> pq = sum of the partials
//   K2  r -= alpha q       (alpha = rr / pq); per-workgroup partials of r.r
//                           -> rr_new
//   K3  x += alpha p, p = r + beta p   (beta = rr_new / rr), then rr = rr_new
// Reductions never use global atomics: partials are summed in a fixed order
// by one workgroup, so a solve is bit-reproducible.  The host reads rr every
// `check` iterations to stop.
//
// Iterative refinement (IR-CG): x64 = 0, r64 = b; repeat
//   r32 = r64 / ||r64||  (float);  solve A_lo d = r32 with float CG to a
//   loose tolerance (A_lo: the same block structure with float or bfloat16
//   values, float accumulation);  x64 += ||r64|| d;  r64 = b - A64 x64
//   (double SpMV) -- until ||r64|| <= rtol ||b||.
//
// Mixed CG with reliable updates: one CG iteration runs on float vectors
// with A_lo and accumulates a float correction d; whenever the iterated
// residual has dropped by a factor delta since the last update (checked
// every `check` iterations), d is folded into the double solution
// (x64 += d, d = 0), the residual is recomputed as b - A64 x64 with the
// double SpMV and replaces the iterated one, and the iteration goes on with
// its search direction p kept (beta uses the new residual).  Unlike the
// restarted refinement above, the Krylov space is never thrown away.
#include "spc5_mat.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace spc5;

namespace spc5 {
namespace {

struct CgScalars {
    double rr;
    double rr_new;
    double pq;
    double pad;
};

constexpr int kVecThreads = 256;
constexpr int kMaxVecGroups = 2048;

__device__ __forceinline__ double block_sum(double v, double* s_part) {
    v = wave_sum_f64(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s_part[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kVecThreads / 64; ++i) t += s_part[i];
    return t;  // valid in thread 0
}

// out = sum of part[0..n), fixed order (one workgroup).
__global__ __launch_bounds__(kVecThreads) void k_reduce_sum(const double* __restrict__ part, int64_t n,
                                                            double* __restrict__ out) {
    __shared__ double s_part[kVecThreads / 64];
    double p = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kVecThreads) p += part[i];
    p = block_sum(p, s_part);
    if (threadIdx.x == 0) *out = p;
}

template <class X>
__global__ __launch_bounds__(kVecThreads) void k_dot(const X* __restrict__ a, const X* __restrict__ b, int64_t n,
                                                     double* __restrict__ part) {
    __shared__ double s_part[kVecThreads / 64];
    double p = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        p += (double)a[i] * (double)b[i];
    p = block_sum(p, s_part);
    if (threadIdx.x == 0) part[blockIdx.x] = p;
}

template <class X>
__global__ __launch_bounds__(kVecThreads) void k_fill(X* __restrict__ a, int64_t n, X v) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        a[i] = v;
}

template <class X>
__global__ __launch_bounds__(kVecThreads) void k_copy(X* __restrict__ dst, const X* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        dst[i] = src[i];
}

// r32 = r64 / ||r64|| with ||r64||^2 read from the device.
__global__ __launch_bounds__(kVecThreads) void k_scale_to_f32(float* __restrict__ dst, const double* __restrict__ src,
                                                              int64_t n, const double* __restrict__ nrm2) {
    const double inv = 1.0 / sqrt(*nrm2);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        dst[i] = (float)(src[i] * inv);
}

// x64 += ||r64|| * d32
__global__ __launch_bounds__(kVecThreads) void k_add_scaled_f32(double* __restrict__ x, const float* __restrict__ d,
                                                                int64_t n, const double* __restrict__ nrm2) {
    const double s = sqrt(*nrm2);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        x[i] += s * (double)d[i];
}

// K2: r -= alpha q; partials of r.r
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_r(X* __restrict__ r, const X* __restrict__ q, int64_t n,
                                                      const CgScalars* __restrict__ sc, double* __restrict__ part) {
    __shared__ double s_part[kVecThreads / 64];
    const X alpha = (X)(sc->rr / sc->pq);
    double p = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads) {
        const X ri = r[i] - alpha * q[i];
        r[i] = ri;
        p += (double)ri * (double)ri;
    }
    p = block_sum(p, s_part);
    if (threadIdx.x == 0) part[blockIdx.x] = p;
}

// K3: x += alpha p; p = r + beta p
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_xp(X* __restrict__ x, X* __restrict__ p, const X* __restrict__ r,
                                                       int64_t n, const CgScalars* __restrict__ sc) {
    const double rr = sc->rr;
    const X alpha = (X)(rr / sc->pq);
    const X beta = (X)(sc->rr_new / rr);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads) {
        const X pi = p[i];
        x[i] += alpha * pi;
        p[i] = r[i] + beta * pi;
    }
}

// K3 split in two around a reliable update: x += alpha p ...
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_x(X* __restrict__ x, const X* __restrict__ p, int64_t n,
                                                      const CgScalars* __restrict__ sc) {
    const X alpha = (X)(sc->rr / sc->pq);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        x[i] += alpha * p[i];
}

// ... and p = r + beta p
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_p(X* __restrict__ p, const X* __restrict__ r, int64_t n,
                                                      const CgScalars* __restrict__ sc) {
    const X beta = (X)(sc->rr_new / sc->rr);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        p[i] = r[i] + beta * p[i];
}

// x64 += d; d = 0
__global__ __launch_bounds__(kVecThreads) void k_fold_f32(double* __restrict__ x, float* __restrict__ d, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads) {
        x[i] += (double)d[i];
        d[i] = 0.0f;
    }
}

// dst = (float) src
__global__ __launch_bounds__(kVecThreads) void k_to_f32(float* __restrict__ dst, const double* __restrict__ src,
                                                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        dst[i] = (float)src[i];
}

__global__ void k_cg_rotate(CgScalars* sc) { sc->rr = sc->rr_new; }

__global__ void k_cg_init_scalars(CgScalars* sc, const double* rr) {
    sc->rr = *rr;
    sc->rr_new = 0.0;
    sc->pq = 0.0;
    sc->pad = 0.0;
}

unsigned vec_grid(int64_t n) {
    int64_t g = (n + kVecThreads - 1) / kVecThreads;
    if (g > kMaxVecGroups) g = kMaxVecGroups;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// *d_out = a.b (partials in `part`, at least vec_grid(n) doubles)
template <class X>
int dot(const X* a, const X* b, int64_t n, double* part, double* d_out, hipStream_t s) {
    const unsigned g = vec_grid(n);
    hipLaunchKernelGGL(k_dot<X>, dim3(g), dim3(kVecThreads), 0, s, a, b, n, part);
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kVecThreads), 0, s, part, (int64_t)g, d_out);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

int read_scalar(const double* d, double* h, hipStream_t s) {
    SPC5_HIP_TRY(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return SPC5_HIP_OK;
}

// CG on A x = b, x holds the initial guess.  rtol is relative to ||b||.
// Vectors are of type X (double for F64 handles, float otherwise).
template <class X>
int cg_impl(const spc5_hip_mat& A, const X* b, X* x, double rtol, int maxit, int check, spc5_hip_solve_stats* st,
            hipStream_t s) {
    const int64_t n = A.nrows;
    SPC5_CHECK_ARG(A.nrows == A.ncols, "CG needs a square matrix");
    const int64_t npart = std::max<int64_t>(kMaxVecGroups, A.plan.ntiles);
    DeviceBuffer vr, vp, vq, vsc, vtmp, vpart;
    SPC5_HIP_TRY(vr.alloc(sizeof(X) * (size_t)n));
    SPC5_HIP_TRY(vp.alloc(sizeof(X) * (size_t)n));
    SPC5_HIP_TRY(vq.alloc(sizeof(X) * (size_t)n));
    SPC5_HIP_TRY(vsc.alloc(sizeof(CgScalars)));
    SPC5_HIP_TRY(vtmp.alloc(2 * sizeof(double)));
    SPC5_HIP_TRY(vpart.alloc(sizeof(double) * (size_t)npart));
    X* r = vr.as<X>();
    X* p = vp.as<X>();
    X* q = vq.as<X>();
    CgScalars* sc = vsc.as<CgScalars>();
    double* part = vpart.as<double>();
    double* d_bb = vtmp.as<double>();
    double* d_rr = d_bb + 1;
    const unsigned g = vec_grid(n);
    int rc;
    // r = b - A x, p = r
    hipLaunchKernelGGL(k_copy<X>, dim3(g), dim3(kVecThreads), 0, s, r, b, n);
    Epilogue ep;
    ep.alpha = -1.0;
    ep.beta = 1.0;
    if ((rc = launch_spmv(A, x, r, s, ep))) return rc;
    hipLaunchKernelGGL(k_copy<X>, dim3(g), dim3(kVecThreads), 0, s, p, r, n);
    if ((rc = dot<X>(b, b, n, part, d_bb, s))) return rc;
    if ((rc = dot<X>(r, r, n, part, d_rr, s))) return rc;
    hipLaunchKernelGGL(k_cg_init_scalars, dim3(1), dim3(1), 0, s, sc, d_rr);
    SPC5_HIP_TRY(hipGetLastError());
    double hbb = 0.0, hrr = 0.0;
    if ((rc = read_scalar(d_bb, &hbb, s))) return rc;
    if ((rc = read_scalar(d_rr, &hrr, s))) return rc;
    const double stop2 = rtol * rtol * (hbb > 0.0 ? hbb : 1.0);
    int it = 0;
    Epilogue ek1;
    ek1.alpha = 1.0;
    ek1.beta = 0.0;
    ek1.dv = p;
    ek1.dot = part;
    while (hrr > stop2 && it < maxit) {
        const int chunk = (maxit - it) < check ? (maxit - it) : check;
        for (int c = 0; c < chunk; ++c) {
            if ((rc = launch_spmv(A, p, q, s, ek1))) return rc;
            hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kVecThreads), 0, s, part, A.plan.ntiles, &sc->pq);
            hipLaunchKernelGGL(k_cg_r<X>, dim3(g), dim3(kVecThreads), 0, s, r, q, n, sc, part);
            hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kVecThreads), 0, s, part, (int64_t)g, &sc->rr_new);
            hipLaunchKernelGGL(k_cg_xp<X>, dim3(g), dim3(kVecThreads), 0, s, x, p, r, n, sc);
            hipLaunchKernelGGL(k_cg_rotate, dim3(1), dim3(1), 0, s, sc);
        }
        SPC5_HIP_TRY(hipGetLastError());
        it += chunk;
        if ((rc = read_scalar(&sc->rr, &hrr, s))) return rc;
        if (!std::isfinite(hrr)) break;
    }
    // true residual ||b - A x|| / ||b||
    hipLaunchKernelGGL(k_copy<X>, dim3(g), dim3(kVecThreads), 0, s, r, b, n);
    if ((rc = launch_spmv(A, x, r, s, ep))) return rc;
    if ((rc = dot<X>(r, r, n, part, d_rr, s))) return rc;
    if ((rc = read_scalar(d_rr, &hrr, s))) return rc;
    if (st) {
        st->iterations = it;
        st->outer = 0;
        st->relres = std::sqrt(hrr / (hbb > 0.0 ? hbb : 1.0));
        st->recurrence_relres = 0.0;
    }
    return SPC5_HIP_OK;
}

// double -> float (round to nearest even), or -> bfloat16 through float
// (each step round to nearest even; NaN stays NaN).
__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ __launch_bounds__(kVecThreads) void k_convert_values(int to_type, const double* __restrict__ in,
                                                                void* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads) {
        const float f = (float)in[i];
        if (to_type == SPC5_HIP_F32)
            reinterpret_cast<float*>(out)[i] = f;
        else
            reinterpret_cast<unsigned short*>(out)[i] = f32_to_bf16_rne(f);
    }
}

}  // namespace
}  // namespace spc5

extern "C" {

int spc5_hip_values_convert(int to_type, const double* d_in, void* d_out, int64_t n, void* stream) {
    SPC5_CHECK_ARG(to_type == SPC5_HIP_F32 || to_type == SPC5_HIP_BF16, "to_type must be F32 or BF16");
    SPC5_CHECK_ARG(n >= 0 && (n == 0 || (d_in != nullptr && d_out != nullptr)), "null argument");
    if (n == 0) return SPC5_HIP_OK;
    hipLaunchKernelGGL(k_convert_values, dim3(vec_grid(n)), dim3(kVecThreads), 0, as_stream(stream), to_type, d_in,
                       d_out, n);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

int spc5_hip_spmv_ex(spc5_hip_mat_t h, double alpha, const void* d_x, double beta, void* d_y, void* stream) {
    SPC5_CHECK_ARG(h != nullptr, "null handle");
    SPC5_CHECK_ARG(d_y != nullptr && (h->nnz == 0 || d_x != nullptr), "x or y is null");
    Epilogue ep;
    ep.alpha = alpha;
    ep.beta = beta;
    return launch_spmv(*h, d_x, d_y, as_stream(stream), ep);
}

int spc5_hip_cg(spc5_hip_mat_t A, const void* d_b, void* d_x, double rtol, int maxit, spc5_hip_solve_stats* st,
                void* stream) {
    SPC5_CHECK_ARG(A != nullptr && d_b != nullptr && d_x != nullptr, "null argument");
    SPC5_CHECK_ARG(rtol > 0.0 && maxit >= 0, "rtol must be > 0 and maxit >= 0");
    hipStream_t s = as_stream(stream);
    if (A->value_type == SPC5_HIP_F64)
        return cg_impl<double>(*A, (const double*)d_b, (double*)d_x, rtol, maxit, 8, st, s);
    return cg_impl<float>(*A, (const float*)d_b, (float*)d_x, rtol, maxit, 8, st, s);
}

int spc5_hip_mixed_cg(spc5_hip_mat_t A64, spc5_hip_mat_t Alo, const double* d_b, double* d_x, double rtol,
                      double delta, int maxit, spc5_hip_solve_stats* st, void* stream) {
    SPC5_CHECK_ARG(A64 != nullptr && Alo != nullptr && d_b != nullptr && d_x != nullptr, "null argument");
    SPC5_CHECK_ARG(A64->value_type == SPC5_HIP_F64, "A64 must hold double values");
    SPC5_CHECK_ARG(Alo->value_type != SPC5_HIP_F64, "Alo must hold float or bfloat16 values");
    SPC5_CHECK_ARG(A64->nrows == Alo->nrows && A64->ncols == Alo->ncols && A64->nrows == A64->ncols,
                   "A64 and Alo must be the same square matrix");
    SPC5_CHECK_ARG(rtol > 0.0 && delta > 0.0 && delta < 1.0 && maxit >= 0, "tolerances out of range");
    hipStream_t s = as_stream(stream);
    const int64_t n = A64->nrows;
    const int check = 8;
    const int64_t npart = std::max<int64_t>(kMaxVecGroups, Alo->plan.ntiles);
    DeviceBuffer vr64, vr, vp, vq, vd, vsc, vtmp, vpart;
    SPC5_HIP_TRY(vr64.alloc(sizeof(double) * (size_t)n));
    SPC5_HIP_TRY(vr.alloc(sizeof(float) * (size_t)n));
    SPC5_HIP_TRY(vp.alloc(sizeof(float) * (size_t)n));
    SPC5_HIP_TRY(vq.alloc(sizeof(float) * (size_t)n));
    SPC5_HIP_TRY(vd.alloc(sizeof(float) * (size_t)n));
    SPC5_HIP_TRY(vsc.alloc(sizeof(CgScalars)));
    SPC5_HIP_TRY(vtmp.alloc(2 * sizeof(double)));
    SPC5_HIP_TRY(vpart.alloc(sizeof(double) * (size_t)npart));
    double* r64 = vr64.as<double>();
    float* r = vr.as<float>();
    float* p = vp.as<float>();
    float* q = vq.as<float>();
    float* d = vd.as<float>();
    CgScalars* sc = vsc.as<CgScalars>();
    double* part = vpart.as<double>();
    double* d_bb = vtmp.as<double>();
    double* d_rr = d_bb + 1;
    const unsigned g = vec_grid(n);
    int rc;
    Epilogue eres;  // r64 = b - A64 x64 (r64 holds a copy of b)
    eres.alpha = -1.0;
    eres.beta = 1.0;
    Epilogue ek1;  // q = A_lo p, partials of p.q
    ek1.alpha = 1.0;
    ek1.beta = 0.0;
    ek1.dv = p;
    ek1.dot = part;

    if ((rc = dot<double>(d_b, d_b, n, part, d_bb, s))) return rc;
    hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kVecThreads), 0, s, r64, d_b, n);
    if ((rc = launch_spmv(*A64, d_x, r64, s, eres))) return rc;
    hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kVecThreads), 0, s, r, r64, n);
    hipLaunchKernelGGL(k_copy<float>, dim3(g), dim3(kVecThreads), 0, s, p, r, n);
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kVecThreads), 0, s, d, n, 0.0f);
    if ((rc = dot<double>(r64, r64, n, part, d_rr, s))) return rc;
    hipLaunchKernelGGL(k_cg_init_scalars, dim3(1), dim3(1), 0, s, sc, d_rr);
    SPC5_HIP_TRY(hipGetLastError());
    double hbb = 0.0, hrr = 0.0;
    if ((rc = read_scalar(d_bb, &hbb, s))) return rc;
    if ((rc = read_scalar(d_rr, &hrr, s))) return rc;
    if (hbb == 0.0) hbb = 1.0;
    const double stop2 = rtol * rtol * hbb;
    double rr_true = hrr;    // ||b - A64 x64||^2 at the last update
    double rr_peak = hrr;    // largest iterated ||r||^2 seen since the last update
    int it = 0, updates = 0;
    while (rr_true > stop2 && it < maxit && std::isfinite(rr_true)) {
        // `check` iterations; the last one stops before its x/p update
        const int chunk = std::min(check, maxit - it);
        for (int c = 0; c < chunk; ++c) {
            if ((rc = launch_spmv(*Alo, p, q, s, ek1))) return rc;
            hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kVecThreads), 0, s, part, Alo->plan.ntiles, &sc->pq);
            hipLaunchKernelGGL(k_cg_r<float>, dim3(g), dim3(kVecThreads), 0, s, r, q, n, sc, part);
            hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kVecThreads), 0, s, part, (int64_t)g, &sc->rr_new);
            if (c + 1 < chunk) {
                hipLaunchKernelGGL(k_cg_xp<float>, dim3(g), dim3(kVecThreads), 0, s, d, p, r, n, sc);
                hipLaunchKernelGGL(k_cg_rotate, dim3(1), dim3(1), 0, s, sc);
            }
        }
        SPC5_HIP_TRY(hipGetLastError());
        it += chunk;
        double rr_it = 0.0;
        if ((rc = read_scalar(&sc->rr_new, &rr_it, s))) return rc;
        if (!std::isfinite(rr_it)) {
            rr_true = rr_it;
            break;
        }
        rr_peak = std::max(rr_peak, rr_it);
        hipLaunchKernelGGL(k_cg_x<float>, dim3(g), dim3(kVecThreads), 0, s, d, p, n, sc);
        const bool update = rr_it <= delta * delta * rr_peak || rr_it <= stop2 || it >= maxit;
        if (update) {
            // reliable update: x64 += d, d = 0; r = b - A64 x64 (double)
            hipLaunchKernelGGL(k_fold_f32, dim3(g), dim3(kVecThreads), 0, s, d_x, d, n);
            hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kVecThreads), 0, s, r64, d_b, n);
            if ((rc = launch_spmv(*A64, d_x, r64, s, eres))) return rc;
            hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kVecThreads), 0, s, r, r64, n);
            if ((rc = dot<double>(r64, r64, n, part, &sc->rr_new, s))) return rc;
            if ((rc = read_scalar(&sc->rr_new, &rr_true, s))) return rc;
            ++updates;
            rr_peak = rr_true;
            if (rr_true <= stop2 || it >= maxit) break;
        }
        hipLaunchKernelGGL(k_cg_p<float>, dim3(g), dim3(kVecThreads), 0, s, p, r, n, sc);
        hipLaunchKernelGGL(k_cg_rotate, dim3(1), dim3(1), 0, s, sc);
        SPC5_HIP_TRY(hipGetLastError());
    }
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    if (st) {
        st->iterations = it;
        st->outer = updates;
        st->relres = std::sqrt(rr_true / hbb);
        st->recurrence_relres = 0.0;
    }
    return SPC5_HIP_OK;
}

int spc5_hip_ir_cg(spc5_hip_mat_t A64, spc5_hip_mat_t Alo, const double* d_b, double* d_x, double rtol,
                   double inner_rtol, int max_outer, int max_inner, spc5_hip_solve_stats* st, void* stream) {
    SPC5_CHECK_ARG(A64 != nullptr && Alo != nullptr && d_b != nullptr && d_x != nullptr, "null argument");
    SPC5_CHECK_ARG(A64->value_type == SPC5_HIP_F64, "A64 must hold double values");
    SPC5_CHECK_ARG(Alo->value_type != SPC5_HIP_F64, "Alo must hold float or bfloat16 values");
    SPC5_CHECK_ARG(A64->nrows == Alo->nrows && A64->ncols == Alo->ncols && A64->nrows == A64->ncols,
                   "A64 and Alo must be the same square matrix");
    SPC5_CHECK_ARG(rtol > 0.0 && inner_rtol > 0.0 && inner_rtol < 1.0, "tolerances out of range");
    hipStream_t s = as_stream(stream);
    const int64_t n = A64->nrows;
    DeviceBuffer vr, vr32, vd32, vtmp, vpart;
    SPC5_HIP_TRY(vr.alloc(sizeof(double) * (size_t)n));
    SPC5_HIP_TRY(vr32.alloc(sizeof(float) * (size_t)n));
    SPC5_HIP_TRY(vd32.alloc(sizeof(float) * (size_t)n));
    SPC5_HIP_TRY(vtmp.alloc(2 * sizeof(double)));
    SPC5_HIP_TRY(vpart.alloc(sizeof(double) * kMaxVecGroups));
    double* r = vr.as<double>();
    float* r32 = vr32.as<float>();
    float* d32 = vd32.as<float>();
    double* part = vpart.as<double>();
    double* d_bb = vtmp.as<double>();
    double* d_rr = d_bb + 1;
    const unsigned g = vec_grid(n);
    int rc;
    Epilogue ep;
    ep.alpha = -1.0;
    ep.beta = 1.0;
    hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kVecThreads), 0, s, r, d_b, n);
    if ((rc = launch_spmv(*A64, d_x, r, s, ep))) return rc;
    if ((rc = dot<double>(d_b, d_b, n, part, d_bb, s))) return rc;
    double hbb = 0.0, hrr = 0.0;
    if ((rc = read_scalar(d_bb, &hbb, s))) return rc;
    if (hbb == 0.0) hbb = 1.0;
    int outer = 0, inner_total = 0;
    for (;;) {
        if ((rc = dot<double>(r, r, n, part, d_rr, s))) return rc;
        if ((rc = read_scalar(d_rr, &hrr, s))) return rc;
        if (hrr <= rtol * rtol * hbb || outer >= max_outer || !std::isfinite(hrr)) break;
        hipLaunchKernelGGL(k_scale_to_f32, dim3(g), dim3(kVecThreads), 0, s, r32, r, n, d_rr);
        hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kVecThreads), 0, s, d32, n, 0.0f);
        SPC5_HIP_TRY(hipGetLastError());
        spc5_hip_solve_stats in{};
        if ((rc = cg_impl<float>(*Alo, r32, d32, inner_rtol, max_inner, 8, &in, s))) return rc;
        inner_total += in.iterations;
        hipLaunchKernelGGL(k_add_scaled_f32, dim3(g), dim3(kVecThreads), 0, s, d_x, d32, n, d_rr);
        hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kVecThreads), 0, s, r, d_b, n);
        if ((rc = launch_spmv(*A64, d_x, r, s, ep))) return rc;
        ++outer;
    }
    if (st) {
        st->iterations = inner_total;
        st->outer = outer;
        st->relres = std::sqrt(hrr / hbb);
        st->recurrence_relres = 0.0;
    }
    return SPC5_HIP_OK;
}

}  // extern "C"
