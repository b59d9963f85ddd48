// spc5_common.hpp -- shared device helpers and host plumbing for the
// MI355X (gfx950) SPC5 library.  Wave64 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/spc5_hip.h"

namespace spc5 {

// --------------------------------------------------------------------------
// Host-side error plumbing.  Every C-ABI entry point returns a status code
// and records a message retrievable with spc5_hip_last_error().
// --------------------------------------------------------------------------
void set_error(const std::string& msg);
std::string hip_msg(hipError_t e, const char* what);

#define SPC5_HIP_TRY(expr)                                                   \
    do {                                                                     \
        hipError_t e__ = (expr);                                             \
        if (e__ != hipSuccess) {                                             \
            ::spc5::set_error(::spc5::hip_msg(e__, #expr));                  \
            return SPC5_HIP_ERR_HIP;                                         \
        }                                                                    \
    } while (0)

#define SPC5_CHECK_ARG(cond, msg)                                            \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ::spc5::set_error(std::string("invalid argument: ") + (msg));    \
            return SPC5_HIP_ERR_INVALID;                                     \
        }                                                                    \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// --------------------------------------------------------------------------
// Device helpers (wave64, gfx950)
// --------------------------------------------------------------------------
constexpr int kWave = 64;

// DPP control codes (GCN/CDNA "row_shr:n").
constexpr int kDppRowShr1 = 0x111;
constexpr int kDppRowShr2 = 0x112;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppRowShr8 = 0x118;

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

// Compiler + memory ordering point for LDS traffic exchanged between the
// lanes of one wave.  LDS instructions of a wave execute in order, so this
// only has to stop the compiler from moving accesses across it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stops the compiler from moving memory accesses across this point (no
// instruction is emitted).  Used to keep a group of independent loads
// together ahead of their first use.
__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }

// Inclusive prefix sum of an int over the 64 lanes: DPP shifts inside each
// 16-lane row, then the three row totals are added with scalar readlanes.
__device__ __forceinline__ int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr1, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr2, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr4, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, kDppRowShr8, 0xf, 0xf, true);
    const int r0 = __builtin_amdgcn_readlane(v, 15);
    const int r1 = __builtin_amdgcn_readlane(v, 31);
    const int r2 = __builtin_amdgcn_readlane(v, 47);
    const int row = lane_id() >> 4;
    v += (row >= 1 ? r0 : 0) + (row >= 2 ? r1 : 0) + (row >= 3 ? r2 : 0);
    return v;
}

// Inclusive prefix maximum of a non-negative int over the 64 lanes (same
// structure as wave_incl_scan; lanes outside a row read 0).
__device__ __forceinline__ int wave_incl_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppRowShr1, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppRowShr2, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppRowShr4, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, kDppRowShr8, 0xf, 0xf, true));
    const int c1 = __builtin_amdgcn_readlane(v, 15);
    const int c2 = max(c1, __builtin_amdgcn_readlane(v, 31));
    const int c3 = max(c2, __builtin_amdgcn_readlane(v, 47));
    const int row = lane_id() >> 4;
    return max(v, row == 0 ? 0 : (row == 1 ? c1 : (row == 2 ? c2 : c3)));
}

// v of lane-1 across the whole wave (GFX9 wave_shr:1); `first` on lane 0.
__device__ __forceinline__ int wave_prev_lane(int v, int first) {
    return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xf, 0xf, false);
}

// Number of set bits of m below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
}

// DPP moves of 32/64-bit values inside 16-lane rows (0 where the source
// lane is outside the row).
constexpr int kDppRowShl1 = 0x101;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_val(float v) {
    return __int_as_float(dpp_i32<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_val(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i32<CTRL>((int)b);
    const int hi = dpp_i32<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Segmented inclusive sum inside each 16-lane row: lanes with equal `key`
// in adjacent positions form a run; afterwards the last lane of every run
// (inside its row) holds the run's sum.  Returns true on those lanes.
template <class T>
__device__ __forceinline__ bool row_segmented_sum(T& s, int key) {
    const int lr = lane_id() & 15;
    const int kprev = dpp_i32<kDppRowShr1>(key);
    const int knext = dpp_i32<kDppRowShl1>(key);
    int f = (lr == 0 || kprev != key) ? 1 : 0;  // run starts here
#define SPC5_SEG_STEP(CTRL)                          \
    {                                                \
        const T su = dpp_val<CTRL>(s);               \
        const int fu = dpp_i32<CTRL>(f);             \
        if (!f) {                                    \
            s += su;                                 \
            f = fu;                                  \
        }                                            \
    }
    SPC5_SEG_STEP(kDppRowShr1)
    SPC5_SEG_STEP(kDppRowShr2)
    SPC5_SEG_STEP(kDppRowShr4)
    SPC5_SEG_STEP(kDppRowShr8)
#undef SPC5_SEG_STEP
    return lr == 15 || knext != key;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_reduce_add(int v) {
    return __builtin_amdgcn_readlane(wave_incl_scan(v), 63);
}

__device__ __forceinline__ unsigned long long ballot64(bool p) { return __ballot(p); }

// Unaligned little-endian loads from the packed header stream.  gfx950
// supports unaligned global accesses; the compiler emits a single dword (or
// short/byte) load for these.
__device__ __forceinline__ int load_i32_unaligned(const unsigned char* p) {
    int v;
    __builtin_memcpy(&v, p, 4);
    return v;
}

template <int NBYTES>
struct MaskWords;  // combined r-row mask, row s at bits [s*VS, (s+1)*VS)

template <>
struct MaskWords<1> {
    static constexpr int kWords = 1;
    __device__ __forceinline__ static void load(const unsigned char* p, unsigned long long* w) {
        unsigned char v;
        __builtin_memcpy(&v, p, 1);
        w[0] = v;
    }
};
template <>
struct MaskWords<2> {
    static constexpr int kWords = 1;
    __device__ __forceinline__ static void load(const unsigned char* p, unsigned long long* w) {
        unsigned short v;
        __builtin_memcpy(&v, p, 2);
        w[0] = v;
    }
};
template <>
struct MaskWords<4> {
    static constexpr int kWords = 1;
    __device__ __forceinline__ static void load(const unsigned char* p, unsigned long long* w) {
        unsigned int v;
        __builtin_memcpy(&v, p, 4);
        w[0] = v;
    }
};
template <>
struct MaskWords<8> {
    static constexpr int kWords = 1;
    __device__ __forceinline__ static void load(const unsigned char* p, unsigned long long* w) {
        unsigned long long v;
        __builtin_memcpy(&v, p, 8);
        w[0] = v;
    }
};
template <>
struct MaskWords<16> {
    static constexpr int kWords = 2;
    __device__ __forceinline__ static void load(const unsigned char* p, unsigned long long* w) {
        unsigned long long v[2];
        __builtin_memcpy(v, p, 16);
        w[0] = v[0];
        w[1] = v[1];
    }
};

// bf16 stored as raw uint16 (upper half of an IEEE binary32).
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Matrix values are read once per product: non-temporal loads keep them
// from evicting the reused x window from L2.
#ifndef SPC5_NT_VALUES
#define SPC5_NT_VALUES 1
#endif
template <class V>
__device__ __forceinline__ V load_stream(const V* p) {
#if SPC5_NT_VALUES
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

// Cache-policy operand of the raw buffer loads of the value stream (the nt
// bit on gfx94x/gfx950), and word 3 of a raw (stride 0) buffer resource on
// gfx9-family GPUs.
constexpr int kNonTemporal = SPC5_NT_VALUES ? 2 : 0;
constexpr int kRawBufferFlags = 0x00020000;
template <class V, class ACC>
struct ValueLoad {
    __device__ __forceinline__ static ACC get(const V* p) { return (ACC)load_stream(p); }
};
template <>
struct ValueLoad<unsigned short, float> {
    __device__ __forceinline__ static float get(const unsigned short* p) { return bf16_to_f32(load_stream(p)); }
};
template <>
struct ValueLoad<unsigned short, double> {
    __device__ __forceinline__ static double get(const unsigned short* p) {
        return (double)bf16_to_f32(load_stream(p));
    }
};

// --------------------------------------------------------------------------
// Host helpers shared by the translation units
// --------------------------------------------------------------------------
int scan_exclusive_i32(int* d_data, int64_t n, int* d_total, hipStream_t s);  // in place, n entries
// The same scan with caller-provided partials (scan_partials_needed(n) ints), no allocation, asynchronous.
int64_t scan_partials_needed(int64_t n);
int scan_exclusive_i32_ws(int* d_data, int64_t n, int* d_total, int* d_partial, hipStream_t s);
// p[0..n) = v (a kernel, not a memset node).
__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v);

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    hipError_t alloc(size_t b) {
        release();
        bytes = b;
        return hipMalloc(&p, b > 0 ? b : 16);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

inline int header_bytes(int r, int mb) { return 4 + r * mb; }

}  // namespace spc5
