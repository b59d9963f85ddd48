// workload.hip -- device generator of the banded benchmark matrices (the
// uneven batch).  Same integer hashing and value formula as the host
// generator spc5_synth_banded_* (spc5_amd/csrc/synth.cpp), so a matrix made
// here is byte-identical to the host one: entry (i, j), |i-j| <= bw, is kept
// when splitmix64(seed*0x100000001b3 ^ (i<<24 ^ j)) < thr (always on the
// diagonal); its value is U[-1,1] from splitmix64 on (seed+1, CSR position),
// plus bw on the diagonal.  Counterpart of the reference driver's synthetic
// builddense (examples/load_mm_and_compare.cpp:431-449).
#include "spc5_common.hpp"

using namespace spc5;

namespace {

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform_pm1(unsigned long long seed, unsigned long long i) {
    const unsigned long long z = splitmix64(splitmix64(seed ^ 0x9fb21c651e98df25ull) + i);
    return -1.0 + 2.0 * ((double)(z >> 11) * 0x1.0p-53);
}

__device__ __forceinline__ bool kept(unsigned long long seed, long long i, long long j, unsigned long long thr) {
    return i == j ||
           splitmix64(seed * 0x100000001b3ull ^ (((unsigned long long)i << 24) ^ (unsigned long long)j)) < thr;
}

__global__ void k_banded_count(int64_t n, int64_t bw, unsigned long long thr, unsigned long long seed, int* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j0 = i - bw > 0 ? i - bw : 0, j1 = i + bw < n - 1 ? i + bw : n - 1;
    int c = 0;
    for (int64_t j = j0; j <= j1; ++j) c += kept(seed, i, j, thr) ? 1 : 0;
    cnt[i] = c;
}

__global__ void k_banded_fill(int64_t n, int64_t bw, unsigned long long thr, unsigned long long seed,
                              const int* __restrict__ rowptr, int* __restrict__ col, double* __restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j0 = i - bw > 0 ? i - bw : 0, j1 = i + bw < n - 1 ? i + bw : n - 1;
    int64_t p = rowptr[i];
    for (int64_t j = j0; j <= j1; ++j) {
        if (!kept(seed, i, j, thr)) continue;
        col[p] = (int)j;
        double v = uniform_pm1(seed + 1, (unsigned long long)p);
        if (j == i) v += (double)bw;
        val[p] = v;
        ++p;
    }
}


// Read-dominated streaming probe: each wave (one per workgroup) reads one
// contiguous tile of `tile` 8-byte words with non-temporal loads, 3 KiB per
// step (six coalesced loads of one word per lane), the way a product's
// waves stream their tiles; then the waves write dst (m8 words) in a
// grid-stride loop.  With a write share of a few percent this is the HBM mix
// of a product y = A x (about 97 % reads), so its rate is the ceiling such a
// kernel can reach.
__global__ __launch_bounds__(64) void k_stream_probe(const unsigned long long* __restrict__ src, int64_t n8,
                                                     int64_t tile, unsigned long long* __restrict__ dst, int64_t m8) {
    constexpr int kSteps = 6;
    const int lane = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * tile;
    const int64_t t1 = t0 + tile < n8 ? t0 + tile : n8;
    unsigned long long acc = 0;
    for (int64_t i = t0 + lane; i < t1; i += 64 * kSteps) {
        unsigned long long v[kSteps];
#pragma unroll
        for (int u = 0; u < kSteps; ++u) v[u] = i + 64 * u < t1 ? __builtin_nontemporal_load(src + i + 64 * u) : 0ull;
#pragma unroll
        for (int u = 0; u < kSteps; ++u) acc ^= v[u];
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m8; j += stride) dst[j] = acc ^ (unsigned)j;
    // The loads must not be dropped: the sum reaches memory on one
    // impossible-looking value (dst holds at least one word).
    if (m8 == 0 && acc == 0x9e3779b97f4a7c15ull) dst[0] = acc;
}

}  // namespace

extern "C" {

int spc5_hip_synth_banded_count(int64_t n, int64_t bw, uint64_t thr, uint64_t seed, int* d_rowptr, int64_t* nnz,
                                void* stream) {
    SPC5_CHECK_ARG(n > 0 && n < (1ll << 24) && bw >= 0 && d_rowptr != nullptr && nnz != nullptr, "bad argument");
    hipStream_t s = as_stream(stream);
    const unsigned g = (unsigned)((n + 255) / 256);
    SPC5_HIP_TRY(hipMemsetAsync(d_rowptr + n, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_banded_count, dim3(g), dim3(256), 0, s, n, bw, (unsigned long long)thr,
                       (unsigned long long)seed, d_rowptr);
    SPC5_HIP_TRY(hipGetLastError());
    DeviceBuffer total;
    SPC5_HIP_TRY(total.alloc(sizeof(int)));
    const int rc = scan_exclusive_i32(d_rowptr, n + 1, total.as<int>(), s);
    if (rc) return rc;
    int h = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&h, total.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    *nnz = h;
    return SPC5_HIP_OK;
}

int spc5_hip_synth_banded_fill(int64_t n, int64_t bw, uint64_t thr, uint64_t seed, const int* d_rowptr, int* d_col,
                               double* d_val, void* stream) {
    SPC5_CHECK_ARG(n > 0 && n < (1ll << 24) && bw >= 0 && d_rowptr && d_col && d_val, "bad argument");
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_banded_fill, dim3(g), dim3(256), 0, as_stream(stream), n, bw, (unsigned long long)thr,
                       (unsigned long long)seed, d_rowptr, d_col, d_val);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

int spc5_hip_stream_probe(const void* d_src, int64_t read_bytes, void* d_dst, int64_t write_bytes,
                          int64_t tile_bytes, void* stream) {
    SPC5_CHECK_ARG(d_src != nullptr && d_dst != nullptr, "null buffer");
    SPC5_CHECK_ARG(read_bytes >= 0 && write_bytes >= 0 && read_bytes % 16 == 0 && write_bytes % 16 == 0,
                   "sizes must be non-negative multiples of 16 bytes");
    SPC5_CHECK_ARG(tile_bytes >= 1024 && tile_bytes % 1024 == 0, "tile_bytes must be a positive multiple of 1 KiB");
    const int64_t tile = tile_bytes / 8;
    const int64_t waves = (read_bytes / 8 + tile - 1) / tile;
    SPC5_CHECK_ARG(waves <= (1ll << 30), "too many tiles");
    hipLaunchKernelGGL(k_stream_probe, dim3((unsigned)(waves > 0 ? waves : 1)), dim3(64), 0, as_stream(stream),
                       reinterpret_cast<const unsigned long long*>(d_src), read_bytes / 8, tile,
                       reinterpret_cast<unsigned long long*>(d_dst), write_bytes / 8);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

}  // extern "C"
