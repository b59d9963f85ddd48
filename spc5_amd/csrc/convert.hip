// convert.hip -- CSR -> beta(r,VS) on the GPU, byte-identical to the
// reference converter (src/spc5.hpp:165-228 and 455-488).
//
// The reference walks each block-row with one cursor per row: a block starts
// at the smallest unconsumed column of the block-row and takes, row by row,
// every entry with column < start+VS.  Block-rows are independent, so the
// GPU gives every 64 rows to one wave, one lane per row, the r lanes of a
// block-row walking together (see "Groups" below), in two passes:
//   count : blocks per block-row -> exclusive scan -> block-row pointers
//           (the reference's rowsSize after conversion) and the block count;
//   fill  : the same walk again, writing each block's packed header at
//           (pointer + k) * (4 + r*MB) bytes and its values at the block-row's
//           value offset, which is the CSR row pointer of its first row
//           (conversion only permutes values inside a block-row).
#include "spc5_common.hpp"

#include <mutex>

#include <climits>

namespace spc5 {

namespace {

// Checks a CSR row pointer: starts at 0, non-decreasing.
__global__ void k_check_rowptr(const int* __restrict__ rowptr, int64_t nrows, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nrows) return;
    bool bad = false;
    if (i == 0) bad = rowptr[0] != 0;
    if (i < nrows) bad |= rowptr[i + 1] < rowptr[i];
    if (bad) atomicOr(err, 1);
}

// Groups: one wave converts kGroupRows = 64 consecutive rows, one lane per
// row; the r lanes of a block-row form a lane group.  For every block the
// group agrees on its start (the minimum of the r row heads, a shuffle
// reduction), each lane consumes its own row's columns inside the block's
// window, and a group prefix sum of the counts places the values (row-major
// inside the block).  The group's column indices are staged in LDS with
// coalesced loads when they fit the stage; the fill pass builds the group's
// packed headers and a destination map of its values in LDS and writes both
// out with coalesced stores.  A group too large for the stage reads its
// columns from global memory and writes directly (the same arithmetic).
//
// The stage is dynamic LDS sized at launch from the matrix's mean values and
// header bytes per group (half as much again as the mean, at least 256 more):
// a small stage keeps many waves resident per CU, which hides the latency of
// each wave's few dependent memory round trips.
constexpr int kGroupRows = 64;
constexpr int kPrefetch = 32;             // words per lane loaded at once: 2048 per group
constexpr int kMaxStageValues = 8192;     // 32 KiB of staged columns
constexpr int kMaxStageHdrBytes = 16384;  // staged header bytes
constexpr int kStagePad = 32;             // words after the staged columns read by the walk's windows

template <int R>
constexpr int group_blockrows() { return kGroupRows / R; }

int64_t round_up64(int64_t v) { return (v + 63) / 64 * 64; }

// Staged values per group: the mean plus half of it (at least 256 more).
int stage_values(int64_t nnz, int64_t ngroups) {
    const int64_t avg = ngroups > 0 ? (nnz + ngroups - 1) / ngroups : 0;
    int64_t s = round_up64(avg + (avg / 2 > 256 ? avg / 2 : 256));
    if (s < 512) s = 512;
    return (int)(s < kMaxStageValues ? s : kMaxStageValues);
}

// Staged header bytes per group, by the same rule.
int stage_header_bytes(int64_t nblocks, int h, int64_t ngroups) {
    const int64_t avg = ngroups > 0 ? (nblocks * h + ngroups - 1) / ngroups : 0;
    int64_t s = round_up64(avg + (avg / 2 > 256 ? avg / 2 : 256));
    if (s < 512) s = 512;
    return (int)(s < kMaxStageHdrBytes ? s : kMaxStageHdrBytes);
}

// Lane-group collectives over R = 1, 2, 4 or 8 consecutive lanes, as DPP
// moves inside 16-lane rows (no LDS round trip): quad_perm [1,0,3,2] and
// [2,3,0,1] exchange within pairs and quads, row_half_mirror (lane i of an
// 8-lane half-row reads lane 7 - i) joins the two quads of an 8-lane group
// once each quad holds its own result, and row_shr:n shifts for the scan.
// Every lane of a group is active whenever one of them calls these.
constexpr int kDppQuadXor1 = 0xB1;
constexpr int kDppQuadXor2 = 0x4E;
constexpr int kDppRowHalfMirror = 0x141;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

template <int R>
__device__ __forceinline__ int group_min(int v) {
    if constexpr (R >= 2) v = min(v, dpp_i32<kDppQuadXor1>(v));
    if constexpr (R >= 4) v = min(v, dpp_i32<kDppQuadXor2>(v));
    if constexpr (R >= 8) v = min(v, dpp_i32<kDppRowHalfMirror>(v));
    return v;
}
template <int R>
__device__ __forceinline__ int group_sum(int v) {
    if constexpr (R >= 2) v += dpp_i32<kDppQuadXor1>(v);
    if constexpr (R >= 4) v += dpp_i32<kDppQuadXor2>(v);
    if constexpr (R >= 8) v += dpp_i32<kDppRowHalfMirror>(v);
    return v;
}
// Exclusive prefix sum over the lanes of a group (lane t: sum over t' < t).
template <int R>
__device__ __forceinline__ int group_excl_scan(int v, int t) {
    int incl = v;
    if constexpr (R >= 2) {
        const int o = dpp_i32<kDppRowShr1>(incl);
        if (t >= 1) incl += o;
    }
    if constexpr (R >= 4) {
        const int o = dpp_i32<kDppRowShr2>(incl);
        if (t >= 2) incl += o;
    }
    if constexpr (R >= 8) {
        const int o = dpp_i32<kDppRowShr4>(incl);
        if (t >= 4) incl += o;
    }
    return incl - v;
}

// Walks the block-rows of a group; for each block of a lane's block-row
// calls on_block(k, start, t, first, n, mask, src0) on every lane of the
// group: k is the block's index in its block-row, t the lane's row in the
// block, first the value position (block order) of the lane's first value in
// the block, n its count and mask its VS-bit row mask; the lane's values were
// at source positions [src0, src0 + n), whose columns the walk never reads
// again.  With CHECK, returns whether a lane saw columns that break the
// format's precondition (non-negative, strictly increasing); without, the
// caller has checked them and it returns false.
//
// `col` is indexed by global value position minus `base` (0 for the global
// array, the group's first value for the stage).  Per block a lane reads the
// window of its next VS + 1 columns at once: its values in the block are the
// window's leading run below start + VS (at most VS, the columns being
// strictly increasing), the entry after them is the next head, and adjacent
// window entries check the precondition (CHECK).  PADDED: `col` may be read up to
// VS + 1 entries past the group's last value (the stage's pad), so the window
// is read without clamping; otherwise each index is clamped to the row.
template <int R, int VS, bool PADDED, bool CHECK, class C, class F>
__device__ __forceinline__ bool walk_group(int rb, int re, C col, int base, F on_block) {
    const int lane = threadIdx.x;
    const int t = lane % R;
    const int end = re;
    int cur = rb;
    int head = cur < end ? col[cur - base] : INT_MAX;  // INT_MAX once the row is consumed
    bool bad = CHECK && head < 0;
    // the block-row's first value: the row pointer of its row 0
    int vbase = __shfl(rb, lane - t, 64);
    int k = 0;
    for (;;) {
        const int start = group_min<R>(head);
        if (!ballot64(start != INT_MAX)) break;  // every block-row of the wave is done
        if (start != INT_MAX) {
            int w[VS + 1];
            if constexpr (PADDED) {
                const auto p = col + (cur - base);
#pragma unroll
                for (int j = 0; j <= VS; ++j) w[j] = p[j];
            } else {
                const int last = end - 1 > cur ? end - 1 : cur;
#pragma unroll
                for (int j = 0; j <= VS; ++j) w[j] = cur < end ? col[(cur + j < last ? cur + j : last) - base] : 0;
            }
            unsigned mask = 0;
            int n = 0;
            bool alive = true;
#pragma unroll
            for (int j = 0; j < VS; ++j) {
                const unsigned d = (unsigned)w[j] - (unsigned)start;
                alive = alive && cur + j < end && d < (unsigned)VS;
                mask |= alive ? 1u << (d & 31u) : 0u;
                n += alive ? 1 : 0;
            }
            if constexpr (CHECK) {
#pragma unroll
                for (int j = 1; j <= VS; ++j) bad |= cur + j < end && w[j] <= w[j - 1];
            }
            int next = w[0];
#pragma unroll
            for (int j = 1; j <= VS; ++j) next = n == j ? w[j] : next;
            const int before = group_excl_scan<R>(n, t);
            on_block(k, start, t, vbase + before, n, mask, cur);
            vbase += group_sum<R>(n);
            cur += n;
            head = cur < end ? next : INT_MAX;
            ++k;
        }
    }
    return bad;
}

// Coalesced copy of n words to LDS, 8 loads in flight per lane per step (a
// load-then-store loop would wait for every load in turn).
template <class T>
__device__ __forceinline__ void stage_words(const T* __restrict__ src, T* dst, int n, int lane) {
    constexpr int U = 8;
    for (int i0 = 0; i0 < n; i0 += 64 * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + lane + 64 * u;
            v[u] = i < n ? src[i] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + lane + 64 * u;
            if (i < n) dst[i] = v[u];
        }
    }
}

// Raw buffer loads of one word type (a read past the resource's range
// returns 0 and touches no memory).
template <class T>
struct RawWord;
template <>
struct RawWord<int> {
    __device__ __forceinline__ static int get(__amdgpu_buffer_rsrc_t r, int off) {
        return (int)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    }
};
template <>
struct RawWord<unsigned int> {
    __device__ __forceinline__ static unsigned int get(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    }
};
template <>
struct RawWord<unsigned long long> {
    __device__ __forceinline__ static unsigned long long get(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(unsigned long long, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
};
template <>
struct RawWord<unsigned short> {
    __device__ __forceinline__ static unsigned short get(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    }
};

// v[u] = src[lane + 64 u] for u < kPrefetch (0 past n), through a buffer
// resource of n words: every load unconditional and one lane offset for all,
// so all are in flight at once at little register cost.
template <class T>
__device__ __forceinline__ void prefetch_words(const T* __restrict__ src, int n, int lane, T (&v)[kPrefetch]) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(src), (short)0, n * (int)sizeof(T), kRawBufferFlags);
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) v[u] = RawWord<T>::get(r, (lane + 64 * u) * (int)sizeof(T));
}

// Stages n column indices in LDS: the first 64 kPrefetch from registers
// already loaded, the rest with stage_words.
__device__ __forceinline__ void stage_columns(const int* __restrict__ src, int* dst, int n, int lane,
                                              const int (&v)[kPrefetch]) {
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int i = lane + 64 * u;
        if (i < n) dst[i] = v[u];
    }
    if (n > 64 * kPrefetch) stage_words(src + 64 * kPrefetch, dst + 64 * kPrefetch, n - 64 * kPrefetch, lane);
    wave_sync();
}

// A group's rows, one per lane: [rb, re) are the lane's values (rows past the
// matrix are empty at its end), v0 the group's first value, nv its count.
struct GroupRows {
    int rb, re, v0, nv;
};
__device__ __forceinline__ GroupRows group_rows(const int* __restrict__ rowptr, int64_t row0, int64_t nrows,
                                                int lane) {
    const int64_t row = row0 + lane;
    GroupRows g;
    g.rb = rowptr[row < nrows ? row : nrows];
    g.re = rowptr[row + 1 < nrows ? row + 1 : nrows];
    g.v0 = __shfl(g.rb, 0, 64);
    g.nv = __shfl(g.re, 63, 64) - g.v0;
    return g;
}

// Row-start bitmap words after a stage of sv values (bit i: value i starts
// a row; one spare word).
__host__ __device__ constexpr int row_start_words(int sv) { return sv / 32 + 2; }

// The column precondition over a group's staged columns (non-negative,
// strictly increasing inside each row), one coalesced pass: value i and its
// successor are compared unless the successor starts a row, as the lanes mark
// their rows' first values in the bitmap `rs`.  True when broken.
__device__ __forceinline__ bool staged_columns_bad(const int* scol, unsigned* rs, int nv, int rb, int re, int v0,
                                                   int lane) {
    const int nw = nv / 32 + 2;
    for (int i = lane; i < nw; i += 64) rs[i] = 0u;
    wave_sync();
    if (rb < re) atomicOr(&rs[(rb - v0) >> 5], 1u << ((rb - v0) & 31));
    wave_sync();
    bool bad = false;
#pragma unroll 4
    for (int i = lane; i < nv; i += 64) {
        const int c0 = scol[i];
        const int c1 = scol[i + 1];  // the stage's pad when i + 1 == nv
        const bool next_starts_row = i + 1 >= nv || ((rs[(i + 1) >> 5] >> ((i + 1) & 31)) & 1u);
        bad |= c0 < 0 || (!next_starts_row && c1 <= c0);
    }
    return bad;
}

// Blocks per block-row (the count pass).  A violation of the column
// precondition sets bit 0 of *err; the fill pass is then never run.  Dynamic
// LDS: sv staged column indices, kStagePad words of pad, then
// row_start_words(sv) words of row-start bitmap.
template <int R, int VS>
__global__ __launch_bounds__(64) void k_convert_count(int64_t nrows, int64_t nbr, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col, int* __restrict__ cnt_out,
                                                      int* __restrict__ err, int sv) {
    constexpr int BR = group_blockrows<R>();
    extern __shared__ __align__(16) unsigned char smem[];
    int* scol = reinterpret_cast<int*>(smem);
    unsigned* rstart = reinterpret_cast<unsigned*>(smem + 4 * (size_t)(sv + kStagePad));
    const int lane = threadIdx.x;
    const int64_t br0 = (int64_t)blockIdx.x * BR;
    const GroupRows g = group_rows(rowptr, br0 * R, nrows, lane);
    int nb = 0;
    auto on_block = [&](int, int, int, int, int, unsigned, int) { ++nb; };
    bool bad = false;
    if (g.nv <= sv) {
        if (g.nv > 0) {
            int c[kPrefetch];
            prefetch_words(col + g.v0, g.nv, lane, c);
            stage_columns(col + g.v0, scol, g.nv, lane, c);
            bad = staged_columns_bad(scol, rstart, g.nv, g.rb, g.re, g.v0, lane);
        }
        walk_group<R, VS, true, false>(g.rb, g.re, scol, g.v0, on_block);
    } else {
        bad = walk_group<R, VS, false, true>(g.rb, g.re, col, 0, on_block);
    }
    const int64_t br = br0 + lane / R;
    if (lane % R == 0 && br < nbr) cnt_out[br] = nb;
    if (bad) atomicOr(err, 1);
}

// Headers and values of the groups' blocks.  Dynamic LDS: sv staged column
// indices (and kStagePad words of pad), overwritten in place by the
// destination of each value once the walk has consumed its column, then
// shb + 4 header bytes (the group's headers, offset by the byte alignment of
// their destination).
template <class W, int R, int VS, int MB>
__global__ __launch_bounds__(64) void k_convert_fill(int64_t nrows, int64_t nbr, const int* __restrict__ rowptr,
                                                     const int* __restrict__ col, const W* __restrict__ vin,
                                                     const int* __restrict__ bptr, int64_t nblocks,
                                                     unsigned char* __restrict__ hdr, W* __restrict__ vout, int sv,
                                                     int shb) {
    constexpr int BR = group_blockrows<R>();
    constexpr int H = 4 + R * MB;
    extern __shared__ __align__(16) unsigned char smem[];
    int* scol = reinterpret_cast<int*>(smem);  // columns, then value destinations
    unsigned char* shdr = smem + 4 * (size_t)(sv + kStagePad);
    const int lane = threadIdx.x;
    const int64_t br0 = (int64_t)blockIdx.x * BR;
    const int64_t br = br0 + lane / R;
    // row pointers and block-row pointers, loaded together
    const GroupRows g = group_rows(rowptr, br0 * R, nrows, lane);
    const int b_lo = bptr[br < nbr ? br : nbr];
    const int b_hi = bptr[br + 1 < nbr ? br + 1 : nbr];
    const int64_t blk0 = __shfl(b_lo, 0, 64);
    const int64_t b_last = __shfl(b_hi, 63, 64);  // pointer past the group's last block-row
    const int64_t blkE = b_last < nblocks ? b_last : nblocks;
    int nblk = 0, bo = 0;  // blocks of the lane's block-row, its first block's index in the group
    if (br < nbr) {
        const int64_t b_end = b_hi < nblocks ? b_hi : nblocks;
        nblk = (int)(b_end > b_lo ? b_end - b_lo : 0);
        bo = (int)(b_lo - blk0);
    }
    const int64_t g0 = blk0 * H, g1 = (blkE > blk0 ? blkE : blk0) * H;  // the group's header bytes
    const int align = (int)(g0 & 3);
    const int v0 = g.v0, nv = g.nv;
    const bool staged = nv <= sv && (g1 - g0) + align <= shb;
    if (!staged) {
        // too large for the stage: global columns, direct writes
        walk_group<R, VS, false, false>(g.rb, g.re, col, 0, [&](int k, int start, int tt, int first, int n,
                                                                 unsigned mask, int src0) {
            if (k >= nblk) return;  // inconsistent pointers: never write past the block-row
            unsigned char* h = hdr + (blk0 + bo + k) * H;
            if (tt == 0) __builtin_memcpy(h, &start, 4);
            __builtin_memcpy(h + 4 + tt * MB, &mask, MB);
            for (int j = 0; j < n; ++j) vout[first + j] = vin[src0 + j];
        });
        return;
    }
    if (nv == 0) return;
    // the group's first 64 kPrefetch columns and values in one round trip
    int c[kPrefetch];
    W v[kPrefetch];
    prefetch_words(col + v0, nv, lane, c);
    prefetch_words(vin + v0, nv, lane, v);
    stage_columns(col + v0, scol, nv, lane, c);
    walk_group<R, VS, true, false>(g.rb, g.re, scol, v0, [&](int k, int start, int tt, int first, int n,
                                                             unsigned mask, int src0) {
        if (k >= nblk) return;
        unsigned char* h = shdr + align + (bo + k) * H;
        if (tt == 0) __builtin_memcpy(h, &start, 4);
        __builtin_memcpy(h + 4 + tt * MB, &mask, MB);
#pragma unroll
        for (int j = 0; j < VS; ++j)
            if (j < n) scol[src0 - v0 + j] = first - v0 + j;
    });
    wave_sync();
    // headers: whole dwords of the destination as dword stores, the partial
    // first and last dwords byte by byte (their other bytes are another
    // group's)
    {
        const int64_t w0 = g0 >> 2, w1 = (g1 + 3) >> 2;  // destination dwords touched
        unsigned* hw = reinterpret_cast<unsigned*>(hdr);
        const unsigned* sw = reinterpret_cast<const unsigned*>(shdr);
        for (int64_t w = w0 + lane; w < w1; w += 64) {
            const int li = (int)(w - w0);
            const int64_t bb = w * 4;  // first destination byte of this dword
            if (bb >= g0 && bb + 4 <= g1) {
                hw[w] = sw[li];
            } else {
                for (int q = 0; q < 4; ++q)
                    if (bb + q >= g0 && bb + q < g1) hdr[bb + q] = shdr[li * 4 + q];
            }
        }
    }
    // values: the prefetched ones from registers, the rest read in CSR order
    // (8 loads in flight per lane); each written to its block-order position
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int i = lane + 64 * u;
        if (i < nv) vout[v0 + scol[i]] = v[u];
    }
    {
        constexpr int U = 8;
        for (int i0 = 64 * kPrefetch; i0 < nv; i0 += 64 * U) {
            W w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + lane + 64 * u;
                w[u] = i < nv ? vin[v0 + i] : W(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + lane + 64 * u;
                if (i < nv) vout[v0 + scol[i]] = w[u];
            }
        }
    }
}

template <int R, int VS>
void launch_count(int64_t nrows, int64_t nbr, const int* rowptr, const int* col, int* out, int* err, int64_t nnz,
                  hipStream_t s) {
    const int64_t groups = (nbr + group_blockrows<R>() - 1) / group_blockrows<R>();
    const int sv = stage_values(nnz, groups);
    const size_t shm = (size_t)(sv + kStagePad + row_start_words(sv)) * 4;
    hipLaunchKernelGGL((k_convert_count<R, VS>), dim3((unsigned)groups), dim3(64), shm, s, nrows, nbr,
                       rowptr, col, out, err, sv);
}

template <class W, int VS, int MB>
int launch_fill_r(int r, int64_t nrows, int64_t nbr, const int* rowptr, const int* col, const void* vin,
                  const int* bptr, int64_t nblocks, int64_t nnz, unsigned char* hdr, void* vout, hipStream_t s) {
    const W* vi = reinterpret_cast<const W*>(vin);
    W* vo = reinterpret_cast<W*>(vout);
#define SPC5_FILL_LAUNCH(R)                                                                                    \
    do {                                                                                                       \
        const int64_t groups = (nbr + group_blockrows<R>() - 1) / group_blockrows<R>();                        \
        const int sv = stage_values(nnz, groups);                                                              \
        const int shb = stage_header_bytes(nblocks, 4 + R * MB, groups);                                       \
        const size_t shm = (size_t)(sv + kStagePad) * 4 + (size_t)shb + 16;                                    \
        hipLaunchKernelGGL((k_convert_fill<W, R, VS, MB>), dim3((unsigned)groups), dim3(64), shm, s, nrows, nbr, \
                           rowptr, col, vi, bptr, nblocks, hdr, vo, sv, shb);                                  \
    } while (0)
    switch (r) {
    case 1: SPC5_FILL_LAUNCH(1); break;
    case 2: SPC5_FILL_LAUNCH(2); break;
    case 4: SPC5_FILL_LAUNCH(4); break;
    case 8: SPC5_FILL_LAUNCH(8); break;
    default: set_error("r must be 1, 2, 4 or 8"); return SPC5_HIP_ERR_INVALID;
    }
#undef SPC5_FILL_LAUNCH
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

template <class W>
int launch_fill_w(int vs, int mb, int r, int64_t nrows, int64_t nbr, const int* rowptr, const int* col,
                  const void* vin, const int* bptr, int64_t nblocks, int64_t nnz, unsigned char* hdr, void* vout,
                  hipStream_t s) {
    if (vs == 8 && mb == 1)
        return launch_fill_r<W, 8, 1>(r, nrows, nbr, rowptr, col, vin, bptr, nblocks, nnz, hdr, vout, s);
    if (vs == 16 && mb == 2)
        return launch_fill_r<W, 16, 2>(r, nrows, nbr, rowptr, col, vin, bptr, nblocks, nnz, hdr, vout, s);
    set_error("(VS, mask bytes) must be (8, 1) or (16, 2)");
    return SPC5_HIP_ERR_INVALID;
}

// Scratch of one count pass: words[0] flags a bad row pointer, words[1]
// unsorted or negative columns, words[2] receives the block total; the scan
// of the block counts uses partials[0 .. partial_cap).
struct CountScratch {
    int* words = nullptr;
    int* partials = nullptr;
    int64_t partial_cap = 0;
};

// Fixed device scratch of spc5_hip_convert_count_noalloc, one copy per GPU
// (a file-scope __device__ array): kScratchWords flag/total words, then the
// scan partials.  Host code uses the address hipGetSymbolAddress returns on
// each device, never the array's name.
constexpr int kScratchInts = 16384;  // 64 KiB
constexpr int kScratchWords = 64;    // keeps the partials 256-byte aligned
constexpr int kMaxDevices = 64;
__device__ __align__(256) int scratch[kScratchInts];

std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (looked up once
// per device, with that device current).
int scratch_address(int** out) {
    int dev = -1;
    SPC5_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index out of range for the fixed scratch table");
        return SPC5_HIP_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        SPC5_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) {
            set_error("hipGetSymbolAddress returned no address for the scratch array");
            return SPC5_HIP_ERR_HIP;
        }
        g_scratch_ptr[dev] = reinterpret_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SPC5_HIP_OK;
}

// Reads n flag/total words back (synchronises the stream).
int read_words(const int* d_words, int* h, int n, hipStream_t s) {
    SPC5_HIP_TRY(hipMemcpyAsync(h, d_words, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return SPC5_HIP_OK;
}

// Count pass shared by spc5_hip_convert_count and its no-allocation variant:
// row pointer check, block count per block-row, exclusive scan into
// d_rowsSize_out[0..nbr], total into *nblocks.  Allocates nothing itself.
int convert_count_impl(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx, int* d_rowsSize_out,
                       int64_t* nblocks, const CountScratch& ws, hipStream_t s) {
    *nblocks = 0;
    const int64_t nbr = (nrows + r - 1) / r;
    if (scan_partials_needed(nbr + 1) > ws.partial_cap) {
        set_error("too many block-rows for the scratch space (" + std::to_string(nbr) + " > " +
                  std::to_string(ws.partial_cap * 2048 - 1) + ")");
        return SPC5_HIP_ERR_INVALID;
    }
    // flags and total start at zero: a fill kernel, not a memset node
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ws.words, (int64_t)3, 0);
    SPC5_HIP_TRY(hipGetLastError());
    {
        const unsigned g = (unsigned)((nrows + 1 + 255) / 256);
        hipLaunchKernelGGL(k_check_rowptr, dim3(g), dim3(256), 0, s, d_rowptr, nrows, ws.words);
        SPC5_HIP_TRY(hipGetLastError());
    }
    // the flag and the value count (which sizes the count pass's stage)
    int bad_rowptr = 0, nnz = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&bad_rowptr, ws.words, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipMemcpyAsync(&nnz, d_rowptr + nrows, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    // the count pass reads columns at the row pointers: only once they are valid
    if (bad_rowptr) {
        set_error("CSR row pointer must start at 0 and be non-decreasing");
        return SPC5_HIP_ERR_INVALID;
    }
    // A matrix without entries may come without a column array (an empty
    // tensor's storage); the count pass then reads no column.
    SPC5_CHECK_ARG(nbr == 0 || d_colidx != nullptr || nnz == 0, "colidx is null");
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_rowsSize_out + nbr, (int64_t)1, 0);
    SPC5_HIP_TRY(hipGetLastError());
    if (nbr > 0) {
        const int key = r * 100 + vs;
        switch (key) {
        case 108: launch_count<1, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 208: launch_count<2, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 408: launch_count<4, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 808: launch_count<8, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 116: launch_count<1, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 216: launch_count<2, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 416: launch_count<4, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        case 816: launch_count<8, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, nnz, s); break;
        }
        SPC5_HIP_TRY(hipGetLastError());
    }
    int rc = scan_exclusive_i32_ws(d_rowsSize_out, nbr + 1, ws.words + 2, ws.partials, s);
    if (rc) return rc;
    int h[3] = {0, 0, 0};
    rc = read_words(ws.words, h, 3, s);
    if (rc) return rc;
    if (h[1]) {
        set_error("CSR columns must be non-negative and strictly increasing inside each row");
        return SPC5_HIP_ERR_INVALID;
    }
    *nblocks = h[2];
    return SPC5_HIP_OK;
}

int check_convert_count_args(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_rowsSize_out,
                             const int64_t* nblocks) {
    SPC5_CHECK_ARG(nblocks != nullptr, "nblocks is null");
    SPC5_CHECK_ARG(r == 1 || r == 2 || r == 4 || r == 8, "r must be 1, 2, 4 or 8");
    SPC5_CHECK_ARG(vs == 8 || vs == 16, "VS must be 8 or 16");
    SPC5_CHECK_ARG(nrows >= 0 && nrows <= INT32_MAX, "nrows out of range");
    SPC5_CHECK_ARG(d_rowptr != nullptr && d_rowsSize_out != nullptr, "null pointer");
    SPC5_CHECK_ARG((const void*)d_rowptr != (const void*)d_rowsSize_out, "rowsSize_out must not alias rowptr");
    return SPC5_HIP_OK;
}

}  // namespace

}  // namespace spc5

using namespace spc5;

extern "C" {

int spc5_hip_convert_count(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx,
                           int* d_rowsSize_out, int64_t* nblocks, void* stream) {
    int rc = check_convert_count_args(vs, r, nrows, d_rowptr, d_rowsSize_out, nblocks);
    if (rc) return rc;
    *nblocks = 0;
    hipStream_t s = as_stream(stream);
    const int64_t nbr = (nrows + r - 1) / r;
    CountScratch ws;
    ws.partial_cap = scan_partials_needed(nbr + 1);
    DeviceBuffer buf;
    SPC5_HIP_TRY(buf.alloc(sizeof(int) * (size_t)(kScratchWords + ws.partial_cap)));
    ws.words = buf.as<int>();
    ws.partials = ws.words + kScratchWords;
    rc = convert_count_impl(vs, r, nrows, d_rowptr, d_colidx, d_rowsSize_out, nblocks, ws, s);
    // the buffer is freed on return: every launch that uses it has ended
    // (each path out of the pass ends with a stream synchronisation)
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return rc;
}

int spc5_hip_convert_count_noalloc(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx,
                                   int* d_rowsSize_out, int64_t* nblocks, void* stream) {
    int rc = check_convert_count_args(vs, r, nrows, d_rowptr, d_rowsSize_out, nblocks);
    if (rc) return rc;
    *nblocks = 0;
    CountScratch ws;
    rc = scratch_address(&ws.words);
    if (rc) return rc;
    ws.partials = ws.words + kScratchWords;
    ws.partial_cap = kScratchInts - kScratchWords;
    return convert_count_impl(vs, r, nrows, d_rowptr, d_colidx, d_rowsSize_out, nblocks, ws, as_stream(stream));
}

int spc5_hip_convert_fill(int value_bytes, int vs, int mask_bytes, int r, int64_t nrows, const int* d_rowptr,
                          const int* d_colidx, const void* d_values_in, const int* d_rowsSize,
                          unsigned char* d_headers_out, void* d_values_out, void* stream) {
    SPC5_CHECK_ARG(value_bytes == 8 || value_bytes == 4 || value_bytes == 2, "value_bytes must be 8, 4 or 2");
    SPC5_CHECK_ARG(nrows >= 0 && nrows <= INT32_MAX, "nrows out of range");
    SPC5_CHECK_ARG(d_rowptr != nullptr && d_rowsSize != nullptr, "null pointer");
    hipStream_t s = as_stream(stream);
    const int64_t nbr = (nrows + r - 1) / r;
    if (nbr == 0) return SPC5_HIP_OK;
    int nb = 0, nnz = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&nb, d_rowsSize + nbr, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipMemcpyAsync(&nnz, d_rowptr + nrows, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    if (nb == 0) return SPC5_HIP_OK;
    SPC5_CHECK_ARG(d_headers_out != nullptr && d_values_out != nullptr && d_values_in != nullptr,
                   "null output/value pointer");
    switch (value_bytes) {
    case 8:
        return launch_fill_w<unsigned long long>(vs, mask_bytes, r, nrows, nbr, d_rowptr, d_colidx, d_values_in,
                                                 d_rowsSize, nb, nnz, d_headers_out, d_values_out, s);
    case 4:
        return launch_fill_w<unsigned int>(vs, mask_bytes, r, nrows, nbr, d_rowptr, d_colidx, d_values_in,
                                           d_rowsSize, nb, nnz, d_headers_out, d_values_out, s);
    default:
        return launch_fill_w<unsigned short>(vs, mask_bytes, r, nrows, nbr, d_rowptr, d_colidx, d_values_in,
                                             d_rowsSize, nb, nnz, d_headers_out, d_values_out, s);
    }
}

}  // extern "C"
