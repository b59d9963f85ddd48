// solver.hip -- conjugate gradient and mixed-precision iterative refinement
// on beta(r,VS) matrices (the precision form; no counterpart in the
// reference).
//
// CG iteration, scalars kept on the device (no host round trip):
//   K1  q = A p            (SpMV, y = A x epilogue; each tile writes its
//                           partial of p.q) -> pq = sum of the partials
//   K2  r -= alpha q       (alpha = rr / pq); per-workgroup partials of r.r
//                           -> rr_new
//   K3  x += alpha p, p = r + beta p   (beta = rr_new / rr)
// rr and rr_new alternate between two device slots from one iteration to
// the next.  Reductions never use global atomics: partials are summed in a
// fixed order by one workgroup, so a solve is bit-reproducible.  The host
// reads rr every `check` iterations to stop; a chunk of `check` iterations
// is recorded once into a hipGraph and replayed.
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
#include <cstdlib>
#include <vector>

using namespace spc5;

namespace spc5 {
namespace {

// Workgroups of the first stage of the p.q reduction (see k_reduce_stage).
constexpr int kSumGroups = 64;

// Device scalars of CG.  Iteration i reads r.r from rr[i & 1] and writes
// the new r.r to rr[(i + 1) & 1], so no kernel has to rotate them.  pq_stage
// holds the first-stage sums of the product's p.q partials.
struct CgScalars {
    double rr[2];
    double pq;
    double pad;
    double pq_stage[kSumGroups];
};

constexpr int kVecThreads = 256;
#ifndef SPC5_VEC_GROUPS
#define SPC5_VEC_GROUPS 2048
#endif
constexpr int kMaxVecGroups = SPC5_VEC_GROUPS;  // most workgroups of a vector kernel (and partials)

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

// out = sum of part[0..n) in a fixed order (one workgroup of kSumThreads):
// thread t loads part[t], part[t + kSumThreads], ... kSumLoads at a time
// (all in flight together: one pass covers 32K partials, one per tile of a
// 4M-row matrix) and adds load k into accumulator k % kSumLanes; then the
// accumulators, the wave and the waves are combined in a fixed order.
constexpr int kSumThreads = 1024;
constexpr int kSumLanes = 8;
constexpr int kSumLoads = 32;
__global__ __launch_bounds__(kSumThreads) void k_reduce_sum(const double* __restrict__ part, int64_t n,
                                                            double* __restrict__ out) {
    __shared__ double s_part[kSumThreads / 64];
    double acc[kSumLanes];
#pragma unroll
    for (int k = 0; k < kSumLanes; ++k) acc[k] = 0.0;
    if (n <= 0) {
        if (threadIdx.x == 0) *out = 0.0;
        return;
    }
    // partial counts (one per tile or per vector workgroup) are far below
    // 2^31, so the offsets from the pass base are 32-bit
    for (int64_t i0 = 0; i0 < n; i0 += (int64_t)kSumThreads * kSumLoads) {
        const double* base = part + i0;
        const unsigned m = (unsigned)(n - i0 < (int64_t)kSumThreads * kSumLoads ? n - i0 : kSumThreads * kSumLoads);
        // Every load is unconditional (index clamped) and the zero for past
        // the end is selected afterwards: a load under a lane condition is
        // merged with its default right after it, which waits for it, and
        // the loads would run one after another.
        double v[kSumLoads];
#pragma unroll
        for (int k = 0; k < kSumLoads; ++k) {
            const unsigned i = (unsigned)k * kSumThreads + threadIdx.x;
            // a 32-bit byte offset from a uniform base: one address register
            const unsigned off = (i < m ? i : m - 1) * (unsigned)sizeof(double);
            v[k] = *(const double*)((const char*)base + off);
        }
#pragma unroll
        for (int k = 0; k < kSumLoads; ++k) {
            const unsigned i = (unsigned)k * kSumThreads + threadIdx.x;
            acc[k % kSumLanes] += i < m ? v[k] : 0.0;
        }
    }
    double p = 0.0;
#pragma unroll
    for (int k = 0; k < kSumLanes; ++k) p += acc[k];
    p = wave_sum_f64(p);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s_part[w] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kSumThreads / 64; ++i) t += s_part[i];
        *out = t;
    }
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

// First stage of the p.q reduction, spread over kSumGroups workgroups (one
// workgroup summing every partial spends its time issuing loads through one
// CU): workgroup w sums the contiguous range [w c, (w+1) c) of the n
// partials, c = ceil(n / kSumGroups), in a fixed order (thread t adds
// elements t, t + kVecThreads, ... of the range, 4 loads in flight; then the
// wave and workgroup sums) into out[w].  k_cg_r adds the kSumGroups sums.
__global__ __launch_bounds__(kVecThreads) void k_reduce_stage(const double* __restrict__ part, int64_t n,
                                                              double* __restrict__ out) {
    __shared__ double s_part[kVecThreads / 64];
    const int64_t c = (n + kSumGroups - 1) / kSumGroups;
    const int64_t b = (int64_t)blockIdx.x * c;
    const int64_t e = b + c < n ? b + c : n;
    double acc = 0.0;
    for (int64_t i0 = b; i0 < e; i0 += 4 * kVecThreads) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = i0 + threadIdx.x + (int64_t)u * kVecThreads;
            v[u] = i < e ? part[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += v[u];
    }
    acc = block_sum(acc, s_part);
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// K2: p.q = sum of the kSumGroups first-stage sums (workgroup 0 stores it in
// sc->pq for K3); r -= alpha q; partials of r.r (par: parity of the
// iteration)
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_r(X* __restrict__ r, const X* __restrict__ q, int64_t n,
                                                      CgScalars* __restrict__ sc, int par, double* __restrict__ part) {
    static_assert(kSumGroups == 64, "one first-stage sum per lane");
    __shared__ double s_part[kVecThreads / 64];
    // every wave adds the same 64 sums by the same xor butterfly: one value
    // in every lane of every workgroup
    const double pq = wave_sum_f64(sc->pq_stage[threadIdx.x & 63]);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->pq = pq;
    const X alpha = (X)(sc->rr[par] / pq);
    double p = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads) {
        const X ri = r[i] - alpha * q[i];
        r[i] = ri;
        p += (double)ri * (double)ri;
    }
    p = block_sum(p, s_part);
    if (threadIdx.x == 0) part[blockIdx.x] = p;
}

// sum of part[0..m) (m <= kMaxVecGroups) in a fixed order, the same value in
// every thread of every workgroup that calls it: thread t adds part[t],
// part[t + kVecThreads], ...; the xor butterfly leaves identical sums in all
// lanes (each step adds the same two values in either order), and every
// thread adds the wave sums in wave order.  Replaces a separate reduction
// launch: the partials stay in L2 and each workgroup reads them itself.
__device__ __forceinline__ double sum_partials_all(const double* __restrict__ part, unsigned m,
                                                   double* __restrict__ s_part) {
    constexpr int kLoads = kMaxVecGroups / kVecThreads;
    static_assert(kMaxVecGroups % kVecThreads == 0 && kMaxVecGroups >= kVecThreads,
                  "SPC5_VEC_GROUPS must be a positive multiple of the vector workgroup size");
    double v[kLoads];
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
        const unsigned i = (unsigned)k * kVecThreads + threadIdx.x;
        const unsigned off = (i < m ? i : m - 1) * (unsigned)sizeof(double);
        v[k] = *(const double*)((const char*)part + off);
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kLoads; ++k) acc += (unsigned)k * kVecThreads + threadIdx.x < m ? v[k] : 0.0;
    acc = wave_sum_f64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kVecThreads / 64; ++w) t += s_part[w];
    return t;
}

// K3: rr' = sum of the m partials of r.r left by K2 (workgroup 0 stores it in
// sc->rr[par ^ 1]); x += alpha p; p = r + beta p
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_xp(X* __restrict__ x, X* __restrict__ p, const X* __restrict__ r,
                                                       int64_t n, CgScalars* __restrict__ sc, int par,
                                                       const double* __restrict__ part, unsigned m) {
    __shared__ double s_part[kVecThreads / 64];
    const double rr = sc->rr[par];
    const double rr_new = sum_partials_all(part, m, s_part);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->rr[par ^ 1] = rr_new;
    const X alpha = (X)(rr / sc->pq);
    const X beta = (X)(rr_new / rr);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads) {
        const X pi = p[i];
        x[i] += alpha * pi;
        p[i] = r[i] + beta * pi;
    }
}

// K3 split in two around a reliable update: rr' as in K3; x += alpha p ...
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_x(X* __restrict__ x, const X* __restrict__ p, int64_t n,
                                                      CgScalars* __restrict__ sc, int par,
                                                      const double* __restrict__ part, unsigned m) {
    __shared__ double s_part[kVecThreads / 64];
    const double rr_new = sum_partials_all(part, m, s_part);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->rr[par ^ 1] = rr_new;
    const X alpha = (X)(sc->rr[par] / sc->pq);
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        x[i] += alpha * p[i];
}

// ... and p = r + beta p
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_cg_p(X* __restrict__ p, const X* __restrict__ r, int64_t n,
                                                      const CgScalars* __restrict__ sc, int par) {
    const X beta = (X)(sc->rr[par ^ 1] / sc->rr[par]);
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

// dst = (double) src
template <class X>
__global__ __launch_bounds__(kVecThreads) void k_to_f64(double* __restrict__ dst, const X* __restrict__ src,
                                                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        dst[i] = (double)src[i];
}

// dst = (float) src
__global__ __launch_bounds__(kVecThreads) void k_to_f32(float* __restrict__ dst, const double* __restrict__ src,
                                                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kVecThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kVecThreads)
        dst[i] = (float)src[i];
}

__global__ void k_cg_init_scalars(CgScalars* sc, const double* rr) {
    sc->rr[0] = *rr;
    sc->rr[1] = 0.0;
    sc->pq = 0.0;
    sc->pad = 0.0;
}

// A fixed sequence of launches recorded once into a hipGraph on a private
// stream and replayed on the caller's stream (fewer launch gaps per CG
// iteration).  SPC5_CG_GRAPH=0 in the environment launches directly.
class ChunkGraph {
  public:
    ChunkGraph() = default;
    ChunkGraph(const ChunkGraph&) = delete;
    ChunkGraph& operator=(const ChunkGraph&) = delete;
    ~ChunkGraph() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (cs_) (void)hipStreamDestroy(cs_);
    }
    static bool enabled() {
        const char* e = std::getenv("SPC5_CG_GRAPH");
        return !(e && e[0] == '0');
    }
    // Runs fn(stream): recorded on the first call, replayed afterwards;
    // without graphs, fn runs on s every time.
    template <class F>
    int run(hipStream_t s, F&& fn) {
        if (!enabled()) return fn(s);
        if (!exec_) {
            SPC5_HIP_TRY(hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking));
            SPC5_HIP_TRY(hipStreamBeginCapture(cs_, hipStreamCaptureModeThreadLocal));
            const int rc = fn(cs_);
            hipGraph_t g = nullptr;
            const hipError_t err = hipStreamEndCapture(cs_, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            SPC5_HIP_TRY(err);
            graph_ = g;
            SPC5_HIP_TRY(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        }
        SPC5_HIP_TRY(hipGraphLaunch(exec_, s));
        return SPC5_HIP_OK;
    }

  private:
    hipStream_t cs_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

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
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kSumThreads), 0, s, part, (int64_t)g, d_out);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

int read_scalar(const double* d, double* h, hipStream_t s) {
    SPC5_HIP_TRY(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return SPC5_HIP_OK;
}

// CG iterations on A from a given state: x, r, p (type X), r.r in
// sc->rr[it & 1]; stops when r.r <= stop2 (read every `check` iterations)
// or after maxit iterations in total.  it and hrr are updated.  Full chunks
// start at an even iteration and are replayed from a hipGraph.
template <class X>
int cg_run(const spc5_hip_mat& A, X* x, X* r, X* p, X* q, CgScalars* sc, double* part, int64_t n, int check,
           int maxit, double stop2, int& it, double& hrr, hipStream_t s) {
    const unsigned g = vec_grid(n);
    Epilogue ek1;
    ek1.alpha = 1.0;
    ek1.beta = 0.0;
    ek1.dv = p;
    ek1.dot = part;
    auto iteration = [&](hipStream_t t, int par) -> int {
        int e;
        if ((e = launch_spmv(A, p, q, t, ek1))) return e;
        hipLaunchKernelGGL(k_reduce_stage, dim3(kSumGroups), dim3(kVecThreads), 0, t, part, A.plan.ntiles,
                           sc->pq_stage);
        hipLaunchKernelGGL(k_cg_r<X>, dim3(g), dim3(kVecThreads), 0, t, r, q, n, sc, par, part);
        hipLaunchKernelGGL(k_cg_xp<X>, dim3(g), dim3(kVecThreads), 0, t, x, p, r, n, sc, par, part, g);
        SPC5_HIP_TRY(hipGetLastError());
        return SPC5_HIP_OK;
    };
    ChunkGraph chunk_graph;  // `check` iterations from an even one (parity returns to 0)
    int rc;
    while (hrr > stop2 && it < maxit) {
        const int chunk = (maxit - it) < check ? (maxit - it) : check;
        if (chunk == check && check % 2 == 0 && it % 2 == 0) {
            rc = chunk_graph.run(s, [&](hipStream_t t) -> int {
                for (int c = 0; c < check; ++c) {
                    const int e = iteration(t, c & 1);
                    if (e) return e;
                }
                return SPC5_HIP_OK;
            });
            if (rc) return rc;
        } else {
            for (int c = 0; c < chunk; ++c)
                if ((rc = iteration(s, (it + c) & 1))) return rc;
        }
        it += chunk;
        if ((rc = read_scalar(&sc->rr[it & 1], &hrr, s))) return rc;
        if (!std::isfinite(hrr)) break;
    }
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
    if ((rc = cg_run<X>(A, x, r, p, q, sc, part, n, check, maxit, stop2, it, hrr, s))) return rc;
    // true residual ||b - A x|| / ||b||
    hipLaunchKernelGGL(k_copy<X>, dim3(g), dim3(kVecThreads), 0, s, r, b, n);
    if ((rc = launch_spmv(A, x, r, s, ep))) return rc;
    if ((rc = dot<X>(r, r, n, part, d_rr, s))) return rc;
    if ((rc = read_scalar(d_rr, &hrr, s))) return rc;
    if (st) {
        st->iterations = it;
        st->outer = 0;
        st->relres = std::sqrt(hrr / (hbb > 0.0 ? hbb : 1.0));
        st->tail_iterations = 0;
        st->reserved = 0;
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

// Mixed CG with reliable updates.  XV is the type of the iteration's
// vectors: float (r, p, q and a float correction folded into x at every
// update) or double (r, p, q and x itself; only the matrix is float).
template <class XV>
int mixed_cg_impl(const spc5_hip_mat& A64, const spc5_hip_mat& Alo, const double* d_b, double* d_x, double rtol,
                  double delta, int maxit, double tail_rtol, spc5_hip_solve_stats* st, hipStream_t s) {
    constexpr bool kF64 = sizeof(XV) == 8;
    const int64_t n = A64.nrows;
    const int check = 8;
    // Partials: one per workgroup of a vector kernel, one per tile of the
    // float product, and one per tile of the double product of the tail
    // (cg_run on A64), whose tile list can be longer when A64 has another r.
    const int64_t npart = std::max<int64_t>(kMaxVecGroups, std::max(Alo.plan.ntiles, A64.plan.ntiles));
    DeviceBuffer vr64, vr, vp, vq, vd, vsc, vtmp, vpart;
    if (!kF64) {
        SPC5_HIP_TRY(vr64.alloc(sizeof(double) * (size_t)n));
        SPC5_HIP_TRY(vd.alloc(sizeof(XV) * (size_t)n));
    }
    SPC5_HIP_TRY(vr.alloc(sizeof(XV) * (size_t)n));
    SPC5_HIP_TRY(vp.alloc(sizeof(XV) * (size_t)n));
    SPC5_HIP_TRY(vq.alloc(sizeof(XV) * (size_t)n));
    SPC5_HIP_TRY(vsc.alloc(sizeof(CgScalars)));
    SPC5_HIP_TRY(vtmp.alloc(2 * sizeof(double)));
    SPC5_HIP_TRY(vpart.alloc(sizeof(double) * (size_t)npart));
    XV* r = vr.as<XV>();
    XV* p = vp.as<XV>();
    XV* q = vq.as<XV>();
    // the residual in double and the running correction
    double* r64 = kF64 ? reinterpret_cast<double*>(r) : vr64.as<double>();
    XV* d = kF64 ? reinterpret_cast<XV*>(d_x) : vd.as<XV>();
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
    auto spmv_lo = [&](hipStream_t t) -> int {
        if constexpr (kF64)
            return launch_spmv_f64vec(Alo, reinterpret_cast<const double*>(p), reinterpret_cast<double*>(q), t, ek1);
        else
            return launch_spmv(Alo, p, q, t, ek1);
    };
    // r64 = b - A64 x; r = r64 (rounded to XV); *dst = r64.r64
    auto true_residual = [&](double* dst) -> int {
        hipLaunchKernelGGL(k_copy<double>, dim3(g), dim3(kVecThreads), 0, s, r64, d_b, n);
        int e;
        if ((e = launch_spmv(A64, d_x, r64, s, eres))) return e;
        if constexpr (!kF64)
            hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kVecThreads), 0, s, reinterpret_cast<float*>(r), r64, n);
        return dot<double>(r64, r64, n, part, dst, s);
    };

    if ((rc = dot<double>(d_b, d_b, n, part, d_bb, s))) return rc;
    if ((rc = true_residual(d_rr))) return rc;
    hipLaunchKernelGGL(k_copy<XV>, dim3(g), dim3(kVecThreads), 0, s, p, r, n);
    if (!kF64) hipLaunchKernelGGL(k_fill<XV>, dim3(g), dim3(kVecThreads), 0, s, d, n, (XV)0);
    hipLaunchKernelGGL(k_cg_init_scalars, dim3(1), dim3(1), 0, s, sc, d_rr);
    SPC5_HIP_TRY(hipGetLastError());
    double hbb = 0.0, hrr = 0.0;
    if ((rc = read_scalar(d_bb, &hbb, s))) return rc;
    if ((rc = read_scalar(d_rr, &hrr, s))) return rc;
    if (hbb == 0.0) hbb = 1.0;
    const double stop2 = rtol * rtol * hbb;
    double rr_true = hrr;  // ||b - A64 x64||^2 at the last update
    double rr_peak = hrr;  // largest iterated ||r||^2 seen since the last update
    const double tail2 = tail_rtol * tail_rtol * hbb;  // below: finish in double
    int it = 0, updates = 0, tail_par = -1, tail_iterations = 0;
    // One CG iteration of parity par on stream t; with `last`, the p update
    // is left out (done after the host has looked at r.r).
    auto iteration = [&](hipStream_t t, int par, bool last) -> int {
        int e;
        if ((e = spmv_lo(t))) return e;
        hipLaunchKernelGGL(k_reduce_stage, dim3(kSumGroups), dim3(kVecThreads), 0, t, part, Alo.plan.ntiles,
                           sc->pq_stage);
        hipLaunchKernelGGL(k_cg_r<XV>, dim3(g), dim3(kVecThreads), 0, t, r, q, n, sc, par, part);
        if (!last)
            hipLaunchKernelGGL(k_cg_xp<XV>, dim3(g), dim3(kVecThreads), 0, t, d, p, r, n, sc, par, part, g);
        else
            hipLaunchKernelGGL(k_cg_x<XV>, dim3(g), dim3(kVecThreads), 0, t, d, p, n, sc, par, part, g);
        SPC5_HIP_TRY(hipGetLastError());
        return SPC5_HIP_OK;
    };
    ChunkGraph chunk_graph;  // `check` iterations, the last without its p update
    while (rr_true > stop2 && it < maxit && std::isfinite(rr_true)) {
        const int chunk = std::min(check, maxit - it);
        if (chunk == check) {
            rc = chunk_graph.run(s, [&](hipStream_t t) -> int {
                for (int c = 0; c < check; ++c) {
                    const int e = iteration(t, c & 1, c + 1 == check);
                    if (e) return e;
                }
                return SPC5_HIP_OK;
            });
            if (rc) return rc;
        } else {
            for (int c = 0; c < chunk; ++c)
                if ((rc = iteration(s, (it + c) & 1, c + 1 == chunk))) return rc;
        }
        const int par = (it + chunk - 1) & 1;  // parity of the chunk's last iteration
        it += chunk;
        double rr_it = 0.0;
        if ((rc = read_scalar(&sc->rr[par ^ 1], &rr_it, s))) return rc;
        if (!std::isfinite(rr_it)) {
            rr_true = rr_it;
            break;
        }
        rr_peak = std::max(rr_peak, rr_it);
        const bool update = rr_it <= delta * delta * rr_peak || rr_it <= stop2 || it >= maxit ||
                            rr_it <= tail2;
        if (update) {
            // reliable update: x64 += d, d = 0 (float vectors); r = b - A64 x64 (double)
            if (!kF64)
                hipLaunchKernelGGL(k_fold_f32, dim3(g), dim3(kVecThreads), 0, s, d_x, reinterpret_cast<float*>(d), n);
            if ((rc = true_residual(&sc->rr[par ^ 1]))) return rc;
            if ((rc = read_scalar(&sc->rr[par ^ 1], &rr_true, s))) return rc;
            ++updates;
            rr_peak = rr_true;
            if (rr_true <= stop2 || it >= maxit) break;
            if (rr_true <= tail2) {
                tail_par = par;
                break;
            }
        }
        hipLaunchKernelGGL(k_cg_p<XV>, dim3(g), dim3(kVecThreads), 0, s, p, r, n, sc, par);
        SPC5_HIP_TRY(hipGetLastError());
    }
    if (tail_par >= 0) {
        // Tail in double: the residual is below what the reduced-precision
        // matrix resolves.  p goes to double, takes its update with the true
        // residual (the Krylov direction is kept) and CG goes on with A64.
        DeviceBuffer vp64, vq64;
        double* p64 = reinterpret_cast<double*>(p);
        double* q64 = reinterpret_cast<double*>(q);
        if (!kF64) {
            SPC5_HIP_TRY(vp64.alloc(sizeof(double) * (size_t)n));
            SPC5_HIP_TRY(vq64.alloc(sizeof(double) * (size_t)n));
            p64 = vp64.as<double>();
            q64 = vq64.as<double>();
            hipLaunchKernelGGL(k_to_f64<XV>, dim3(g), dim3(kVecThreads), 0, s, p64, p, n);
        }
        hipLaunchKernelGGL(k_cg_p<double>, dim3(g), dim3(kVecThreads), 0, s, p64, r64, n, sc, tail_par);
        SPC5_HIP_TRY(hipGetLastError());
        const int it0 = it;
        double hrr = rr_true;
        if ((rc = cg_run<double>(A64, d_x, r64, p64, q64, sc, part, n, check, maxit, stop2, it, hrr, s))) return rc;
        tail_iterations = it - it0;
        if ((rc = true_residual(d_rr))) return rc;
        if ((rc = read_scalar(d_rr, &rr_true, s))) return rc;
    }
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    if (st) {
        st->iterations = it;
        st->outer = updates;
        st->relres = std::sqrt(rr_true / hbb);
        st->tail_iterations = tail_iterations;
        st->reserved = 0;
    }
    return SPC5_HIP_OK;
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
                      double delta, int maxit, int vec64, double tail_rtol, spc5_hip_solve_stats* st,
                      void* stream) {
    SPC5_CHECK_ARG(A64 != nullptr && Alo != nullptr && d_b != nullptr && d_x != nullptr, "null argument");
    SPC5_CHECK_ARG(A64->value_type == SPC5_HIP_F64, "A64 must hold double values");
    SPC5_CHECK_ARG(Alo->value_type != SPC5_HIP_F64, "Alo must hold float or bfloat16 values");
    SPC5_CHECK_ARG(A64->nrows == Alo->nrows && A64->ncols == Alo->ncols && A64->nrows == A64->ncols,
                   "A64 and Alo must be the same square matrix");
    SPC5_CHECK_ARG(rtol > 0.0 && delta > 0.0 && delta < 1.0 && maxit >= 0, "tolerances out of range");
    SPC5_CHECK_ARG(vec64 == 0 || vec64 == 1, "vec64 must be 0 or 1");
    SPC5_CHECK_ARG(tail_rtol >= 0.0, "tail_rtol must be >= 0");
    hipStream_t s = as_stream(stream);
    if (vec64) return mixed_cg_impl<double>(*A64, *Alo, d_b, d_x, rtol, delta, maxit, tail_rtol, st, s);
    return mixed_cg_impl<float>(*A64, *Alo, d_b, d_x, rtol, delta, maxit, tail_rtol, st, s);
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
        st->tail_iterations = 0;
        st->reserved = 0;
    }
    return SPC5_HIP_OK;
}

}  // extern "C"
