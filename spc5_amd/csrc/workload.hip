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


// Read-dominated streaming probe: each wave reads one contiguous tile of
// `tile` 16-byte words with non-temporal loads, 4 KiB per step (four
// 16-byte loads per lane in flight), the way a product's waves stream their
// tiles; then the waves write dst (m16 words) in a grid-stride loop.  With
// a write share of a few percent this is the HBM mix of a product y = A x
// (about 97 % reads), so its rate is the ceiling such a kernel can reach.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_probe(const u32x4* __restrict__ src, int64_t n16, int64_t tile,
                                                      u32x4* __restrict__ dst, int64_t m16) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t t0 = wave * tile;
    const int64_t t1 = t0 + tile < n16 ? t0 + tile : n16;
    unsigned acc = 0;
    int64_t i = t0 + lane;
    for (; i + 192 < t1; i += 256) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + 64 * u);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    for (; i < t1; i += 64) {
        const u32x4 v = __builtin_nontemporal_load(src + i);
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m16; j += stride)
        dst[j] = u32x4{acc, (unsigned)j, 0u, 0u};
    // The loads must not be dropped: the sum reaches memory on one
    // impossible-looking value (dst holds at least one word).
    if (m16 == 0 && acc == 0x9e3779b9u) dst[0] = u32x4{acc, 1u, 0u, 0u};
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
    const int64_t tile = tile_bytes / 16;
    const int64_t waves = (read_bytes / 16 + tile - 1) / tile;
    const int64_t groups = waves > 0 ? (waves + 3) / 4 : 1;
    SPC5_CHECK_ARG(groups <= (1ll << 30), "too many tiles");
    hipLaunchKernelGGL(k_stream_probe, dim3((unsigned)groups), dim3(256), 0, as_stream(stream),
                       reinterpret_cast<const u32x4*>(d_src), read_bytes / 16, tile, reinterpret_cast<u32x4*>(d_dst),
                       write_bytes / 16);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

}  // extern "C"
