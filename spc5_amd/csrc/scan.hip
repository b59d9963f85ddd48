// scan.hip -- in-place exclusive prefix sum of int32 arrays (reduce, scan
// of partials, apply).  Used by the plan builder and the GPU converter.
#include "spc5_common.hpp"

namespace spc5 {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                            // per thread
constexpr int kScanChunk = kScanThreads * kScanItems;    // per workgroup

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ int block_excl_scan(int v, int* s_wave_tot, int& excl) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    if (lane == 63) s_wave_tot[wave] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const int t = s_wave_tot[w];
        if (w < wave) before += t;
        total += t;
    }
    __syncthreads();
    excl = before + incl - v;
    return total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int* __restrict__ in, int64_t n,
                                                               int* __restrict__ partial) {
    __shared__ int s_tot[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        if (i < n) sum += in[i];
    }
    int excl;
    const int total = block_excl_scan(sum, s_tot, excl);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One workgroup scans all partials, chunk after chunk, carrying the total.
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(int* __restrict__ partial, int64_t np,
                                                                 int* __restrict__ total_out) {
    __shared__ int s_tot[kScanThreads / 64];
    int carry = 0;
    for (int64_t c0 = 0; c0 < np; c0 += kScanChunk) {
        int v[kScanItems];
        int sum = 0;
        const int64_t base = c0 + (int64_t)threadIdx.x * kScanItems;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = base + k;
            v[k] = i < np ? partial[i] : 0;
            sum += v[k];
        }
        int excl;
        const int total = block_excl_scan(sum, s_tot, excl);
        int run = carry + excl;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = base + k;
            if (i < np) partial[i] = run;
            run += v[k];
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(int* __restrict__ data, int64_t n,
                                                              const int* __restrict__ partial) {
    __shared__ int s_tot[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        v[k] = i < n ? data[i] : 0;
        sum += v[k];
    }
    int excl;
    (void)block_excl_scan(sum, s_tot, excl);
    int run = partial[blockIdx.x] + excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        if (i < n) data[i] = run;
        run += v[k];
    }
}

}  // namespace

__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int64_t scan_partials_needed(int64_t n) { return n <= 0 ? 0 : (n + kScanChunk - 1) / kScanChunk; }

// In-place exclusive scan of d_data[0..n) with caller-provided partials
// (scan_partials_needed(n) ints); the grand total goes to *d_total (device
// pointer, may be null).  Asynchronous: the partials are in use until the
// stream reaches the end of the launches.  Totals must fit int32 (checked by
// callers).
int scan_exclusive_i32_ws(int* d_data, int64_t n, int* d_total, int* d_partial, hipStream_t s) {
    if (n <= 0) {
        if (d_total) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_total, (int64_t)1, 0);
        SPC5_HIP_TRY(hipGetLastError());
        return SPC5_HIP_OK;
    }
    const int64_t nblk = scan_partials_needed(n);
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nblk), dim3(kScanThreads), 0, s, d_data, n, d_partial);
    SPC5_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, s, d_partial, nblk, d_total);
    SPC5_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nblk), dim3(kScanThreads), 0, s, d_data, n, d_partial);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

// The same with partials the call allocates (and frees after a stream
// synchronisation).
int scan_exclusive_i32(int* d_data, int64_t n, int* d_total, hipStream_t s) {
    if (n <= 0) {
        if (d_total) SPC5_HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(int), s));
        return SPC5_HIP_OK;
    }
    DeviceBuffer partial;
    SPC5_HIP_TRY(partial.alloc(sizeof(int) * (size_t)scan_partials_needed(n)));
    const int rc = scan_exclusive_i32_ws(d_data, n, d_total, partial.as<int>(), s);
    if (rc) return rc;
    // partial is freed on return: wait for the kernels that use it.
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return SPC5_HIP_OK;
}

}  // namespace spc5
