// spmv_tile.hpp -- one wave computes y = alpha A x + beta y over one tile (a
// run of whole block-rows) of a beta(r,VS) matrix in the reference layout.
//
// Per batch of up to 64 blocks (one block per lane):
//   1. the lane's block header (int32 first column + r masks) was loaded
//      ahead; popcounts and a DPP wave scan give every block's value offset
//      inside the batch; blocks are taken while the batch holds at most
//      expand_cap<V>() values;
//   2. the batch's values are loaded at once, expand_cap<V>()/64 coalesced
//      loads per lane at most, straight from the value stream (no zero
//      padding), through a buffer resource based at the batch's first value
//      and as long as the batch (constant lane offsets; the range check
//      returns 0 past the batch's end; 64-value chunks wholly past it are
//      not issued);
//   3. the next batch's headers are requested;
//   4. the block-row of every block of the batch comes from one pass over
//      the tile's block-row pointers in LDS (the block-rows that start in
//      the batch mark their first block, a wave prefix maximum spreads the
//      marks); each taken lane then writes, for every set mask bit, the
//      column and the tile-local row into LDS;
//      Optionally (kSort, r > 1) the values are given row-sorted positions
//      at the same time (row_displacements);
//   5. in value order, every lane gathers x for its values (neighbouring
//      lanes read neighbouring columns; 32-bit byte offsets from x when x
//      spans less than 4 GiB) and forms the products;
//   6. the products go to LDS at their (row-sorted) positions; lane l then
//      owns the K consecutive positions l*K ...: it adds them in registers,
//      one LDS add per run of equal rows, the lane's last run first merged
//      with the neighbouring lanes' (row_segmented_sum).
// At the end of the tile each row is written once: y = alpha*sum + beta*y.
#pragma once

#include "spc5_mat.hpp"

namespace spc5 {

// x through 32-bit byte offsets when it spans < 4 GiB (1) or always through
// 64-bit addresses (0).
#ifndef SPC5_X32
#define SPC5_X32 1
#endif

// Row-sorted batches (see row_displacements): on for r > 1 when asked for.
template <int R, bool kSort>
struct RowSort {
    static constexpr bool kOn = kSort && R > 1;
    static constexpr int kWords = (R + 2) / 3;  // packed words of three 10-bit row counts
};

// Waves per SIMD each tile kernel is compiled for (a register budget of
// 512 / waves VGPRs): the most that compile without spills.  0 leaves the
// choice to the compiler.
#ifndef SPC5_WAVES_PER_EU
#define SPC5_WAVES_PER_EU 8
#endif
constexpr int tile_waves_per_eu(int r, bool sorted, bool wide_x) {
    // wide_x: double x with float sums (twice the registers per x value)
    return SPC5_WAVES_PER_EU == 0 ? 1
                                  : (sorted ? (r == 2 ? 6 : 5) : (r == 8 ? 7 : SPC5_WAVES_PER_EU)) - (wide_x ? 1 : 0);
}

constexpr int kNoRow = kMaxRows;  // row byte of the positions past the batch's end

// LDS of one wave's tile: CAP values per batch.
template <class ACC, int R, int CAP, bool kSort = false>
struct TileSmem {
    ACC acc[kMaxRows];  // row sums of the tile
    // Columns in value order (x byte offsets when x spans less than 4 GiB);
    // every column is read before the first product is stored over them.
    union alignas(16) {
        int col[CAP];
        int2 cp[RowSort<R, kSort>::kOn ? CAP : 1];  // row-sorted: column and row-sorted position
        ACC prod[CAP];                              // products at their (row-sorted) positions
    };
    // Tile-local row of batch position q; kNoRow past the batch's end.
    alignas(8) unsigned char row[CAP];
    int rs[kMaxRows / R + 1];
    int mark[64];  // block-row marks of one batch
    int segs[RowSort<R, kSort>::kOn ? 2 * RowSort<R, kSort>::kWords * 64 : 1];  // row counts at segment ends
};

__device__ __forceinline__ int find_blockrow(const int* rs, int nbr, int top, int jr) {
    // Largest i in [0, nbr) with rs[i] <= jr (rs[0] == 0 <= jr).
    int lo = 0;
    for (int step = top; step >= 1; step >>= 1) {
        const int c = lo + step;
        if (c < nbr && rs[c] <= jr) lo = c;
    }
    return lo;
}

// Header of block j (int32 first column + r masks), zero when j >= b1.
template <int R, int MB>
__device__ __forceinline__ void load_header(const unsigned char* __restrict__ hdr, int j, int b1, int& cs,
                                            unsigned long long* mw) {
    using MW = MaskWords<R * MB>;
#pragma unroll
    for (int k = 0; k < MW::kWords; ++k) mw[k] = 0ull;
    cs = 0;
    if (j < b1) {
        const unsigned char* h = hdr + (int64_t)j * (4 + R * MB);
        cs = load_i32_unaligned(h);
        MW::load(h + 4, mw);
    }
}

// Header of block min(j, last) with no branch around the loads, so that a
// prefetch stays in flight until its first use (a load under a lane
// condition is merged with the default right after it, which waits for
// every older load too).  Lanes past the end are masked where the header is
// used.
template <int R, int MB>
__device__ __forceinline__ void load_header_clamped(const unsigned char* __restrict__ hdr, int j, int last,
                                                    int& cs, unsigned long long* mw) {
    const int jj = j < last ? j : last;
    const unsigned char* h = hdr + (int64_t)jj * (4 + R * MB);
    cs = load_i32_unaligned(h);
    if constexpr (R * MB == 1) {
        mw[0] = (unsigned)h[4];  // the byte load zero-extends; nothing waits for it here
    } else {
        MaskWords<R * MB>::load(h + 4, mw);
    }
}

struct BatchShape {
    int cnt;    // values of this lane's block
    int incl;   // inclusive prefix of cnt over the lanes
    int ntake;  // blocks taken in this batch (a prefix of the lanes)
    int total;  // values taken in this batch
};

template <int NW, int CAP>
__device__ __forceinline__ BatchShape batch_shape(const unsigned long long* mw, bool valid) {
    BatchShape bs;
    bs.cnt = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) bs.cnt += __popcll(mw[k]);
    bs.incl = wave_incl_scan(bs.cnt);
    bs.ntake = __popcll(ballot64(valid && bs.incl <= CAP));
    bs.total = bs.ntake > 0 ? __builtin_amdgcn_readlane(bs.incl, bs.ntake - 1) : 0;
    return bs;
}

template <class V, class ACC, int K>
__device__ __forceinline__ void load_values(const V* __restrict__ val, int v, int total, int lane, ACC* a) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = lane + 64 * k;
        a[k] = q < total ? ValueLoad<V, ACC>::get(val + v + q) : (ACC)0;
    }
}

// Every set mask bit of one block -> (column, tile-local row) in LDS,
// starting at batch position p; rbase is the tile-local first row of the
// block's block-row.  The column written is cs + c * kScale for block
// column c (kScale = sizeof(X) with cs a byte offset).
template <int R, int VS, int NW, int kScale = 1>
__device__ __forceinline__ void expand_rows(int* col, unsigned char* row, int rbase, int cs,
                                            const unsigned long long* mw, int p) {
    if constexpr (R * VS <= 16) {
        // Few mask bits: test each one (no data-dependent trip count).
        const unsigned m = (unsigned)mw[0];
#pragma unroll
        for (int bit = 0; bit < R * VS; ++bit) {
            if ((m >> bit) & 1u) {
                col[p] = cs + (bit % VS) * kScale;
                row[p] = (unsigned char)(rbase + bit / VS);
                ++p;
            }
        }
    } else if constexpr (R * VS <= 32) {
        unsigned m = (unsigned)mw[0];
        while (m) {
            const int bit = __builtin_ctz(m);
            col[p] = cs + (bit & (VS - 1)) * kScale;
            row[p] = (unsigned char)(rbase + bit / VS);
            ++p;
            m &= m - 1u;
        }
    } else {
#pragma unroll
        for (int wd = 0; wd < NW; ++wd) {
            unsigned long long m = mw[wd];
            while (m) {
                const int bit = __builtin_ctzll(m) + wd * 64;
                col[p] = cs + (bit & (VS - 1)) * kScale;
                row[p] = (unsigned char)(rbase + bit / VS);
                ++p;
                m &= m - 1ull;
            }
        }
    }
}

template <int R, int VS, int NW>
__device__ __forceinline__ void expand(int* col, unsigned char* row, const int* rs, int nbr, int top, int jr,
                                       int cs, const unsigned long long* mw, int p) {
    expand_rows<R, VS, NW>(col, row, find_blockrow(rs, nbr, top, jr) * R, cs, mw, p);
}

// Values of row t (0 <= t < R) of a block mask.
template <int R, int VS, int NW>
__device__ __forceinline__ int row_count(const unsigned long long* mw, int t) {
    const int bit = t * VS;
    return __popcll((mw[bit / 64] >> (bit % 64)) & ((1ull << VS) - 1ull));
}

// Row-sorted batch order (r > 1).  In value order the blocks of one
// block-row follow each other and each block holds its rows 0..r-1 in turn,
// so one row's values are spread over every block of the block-row and the
// per-run LDS adds of neighbouring lanes hit the same address.  The
// products are therefore stored in row order inside each block-row segment
// of the batch (every value of the segment's row 0, then of row 1, ...;
// blocks in order inside a row).  The value at value-order position p in
// row t of the lane's block goes to position p + d[t]; this returns d for
// the lane's block.
//
// Per-row counts of the taken blocks are packed three to a word (10-bit
// fields; a batch holds fewer than 1024 values, so no field carries) and prefix-summed over the wave; the first and last lane of
// every block-row segment leave the prefix at the segment's start and end
// in LDS, from which every lane gets its segment's row totals.
template <int R, int VS, int NW>
__device__ __forceinline__ void row_displacements(int* segs, const unsigned long long* mw, int br, int ntake, int p0,
                                                  int lane, int* d) {
    constexpr int kW = RowSort<R, true>::kWords;
    const bool taken = lane < ntake;
    int c[R];
    int pk[kW] = {};
#pragma unroll
    for (int t = 0; t < R; ++t) {
        c[t] = taken ? row_count<R, VS, NW>(mw, t) : 0;
        pk[t / 3] += c[t] << (10 * (t % 3));
    }
    int incl[kW];
#pragma unroll
    for (int w = 0; w < kW; ++w) incl[w] = wave_incl_scan(pk[w]);
    const bool start = taken && wave_prev_lane(br, -1) != br;
    const unsigned long long sb = ballot64(start);
    int seg = lanes_below(sb) + (start ? 1 : 0) - 1;  // index of the lane's segment
    seg = seg < 0 ? 0 : seg;
    const bool end = taken && (lane == ntake - 1 || (((sb >> lane) >> 1) & 1ull));
    if (start) {
#pragma unroll
        for (int w = 0; w < kW; ++w) segs[w * 64 + seg] = incl[w] - pk[w];
    }
    if (end) {
#pragma unroll
        for (int w = 0; w < kW; ++w) segs[(kW + w) * 64 + seg] = incl[w];
    }
    wave_sync();
    int at = 0, pc = p0;  // segment position reached; value-order position of row t of the block
    int f[kW], tot[kW], own[kW];
#pragma unroll
    for (int w = 0; w < kW; ++w) {
        f[w] = segs[w * 64 + seg];
        tot[w] = segs[(kW + w) * 64 + seg] - f[w];
        own[w] = incl[w] - pk[w] - f[w];  // values of each row in earlier blocks of the segment
        at += (f[w] & 0x3ff) + ((f[w] >> 10) & 0x3ff) + ((f[w] >> 20) & 0x3ff);
    }
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int sh = 10 * (t % 3);
        d[t] = at + ((own[t / 3] >> sh) & 0x3ff) - pc;
        at += (tot[t / 3] >> sh) & 0x3ff;
        pc += c[t];
    }
}

// expand_rows for a row-sorted batch: value p (value order) gets its
// column and row-sorted position p + d[t] at cp[p], and its row at
// row[p + d[t]].
template <int R, int VS, int NW, int kScale = 1>
__device__ __forceinline__ void expand_rows_sorted(int2* cp, unsigned char* row, int rbase, int cs,
                                                   const unsigned long long* mw, int p, const int* d) {
    // One loop per block row over its set bits only: the row and its
    // displacement are compile-time indices (no dynamically indexed array),
    // and a loop runs as often as the wave's largest count of that row (a
    // loop over every bit position measured 4-5 % slower for float and
    // bfloat16 values on the 27-point stencil, whose rows hold 3 of 8).
    {
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const int bit0 = t * VS;
            unsigned m = (unsigned)((mw[bit0 / 64] >> (bit0 % 64)) & ((1ull << VS) - 1ull));
            const int dt = d[t];
            const unsigned char rt = (unsigned char)(rbase + t);
            while (m) {
                const int c = __builtin_ctz(m);
                cp[p] = make_int2(cs + c * kScale, p + dt);
                row[p + dt] = rt;
                ++p;
                m &= m - 1u;
            }
        }
    }
}

// Block-row of block `bb + lane` (tile-local block indices) for a batch of
// `ntake` blocks, given `prev`, the block-row of block bb-1 (-1 when bb is
// the tile's first block).  rs holds the tile's block-row pointers (nbr+1
// entries, rs[0] = 0), mark is a 64-entry LDS scratch row.  The block-rows
// prev+1 .. prev+128 are looked at (two per lane); those starting inside the
// batch mark their first block with their rank, and a wave prefix maximum
// gives every block the last block-row started at or before it.  When more
// than 127 block-rows start inside the batch (runs of empty block-rows),
// the lanes fall back to a binary search.
__device__ __forceinline__ int batch_blockrows(const int* rs, int* mark, int nbr, int top, int bb, int ntake,
                                               int prev, int lane) {
    mark[lane] = 0;
    const int i0 = prev + 1 + lane, i1 = i0 + 64;
    const int s0 = i0 <= nbr ? rs[i0] : 0x7fffffff;
    const int s1 = i1 <= nbr ? rs[i1] : 0x7fffffff;
    wave_sync();
    const int d0 = s0 - bb, d1 = s1 - bb;  // >= 0: block-row prev+1 starts at bb or later
    if (d0 < ntake) atomicMax(&mark[d0], lane + 1);
    if (d1 < ntake) atomicMax(&mark[d1], lane + 65);
    const bool overflow = __builtin_amdgcn_readlane(d1, 63) < ntake;  // row prev+128 starts inside
    wave_sync();
    int br;
    if (!overflow) {
        br = prev + wave_incl_max(mark[lane]);
    } else {
        br = find_blockrow(rs, nbr, top, bb + lane);
    }
    return br;
}

// Row sums of one batch by lane-owned runs: lane l owns batch positions
// l*K .. l*K+K-1 (products and rows in LDS in position order; rows kNoRow
// past the batch's end).  Consecutive values of one row are added in
// registers; every run of equal rows ends with one LDS add.
template <class ACC, int K>
__device__ __forceinline__ void reduce_slots(ACC* acc, const unsigned char* row, const ACC* prod, int lane) {
    ACC p[K];
    if constexpr (K % 2 == 0) {
        // Pairs: 8 or 16-byte reads, lane stride 3 or 6 words (no bank conflict).
        using P2 = typename std::conditional<sizeof(ACC) == 8, double2, float2>::type;
        const P2* pp = reinterpret_cast<const P2*>(prod + lane * K);
#pragma unroll
        for (int k = 0; k < K / 2; ++k) {
            const P2 v = pp[k];
            p[2 * k] = v.x;
            p[2 * k + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) p[k] = prod[lane * K + k];
    }
    int rw[K];
#pragma unroll
    for (int k = 0; k < K; ++k) rw[k] = row[lane * K + k];
    // Positions past the end form a suffix, so a run is only ever closed on
    // a real row.  Runs that end inside the lane go to LDS at once; the
    // lane's last run is first summed with the last runs of neighbouring
    // lanes on the same row (DPP segmented sum inside 16-lane rows), so that
    // a row spread over many lanes costs one LDS add, not one per lane.
    // Runs past the end (row kNoRow) are dropped.
    ACC s = p[0];
    int rp = rw[0];
#pragma unroll
    for (int k = 1; k < K; ++k) {
        const bool change = rw[k] != rp;
        if (change) atomicAdd(&acc[rp], s);
        s = change ? p[k] : s + p[k];
        rp = rw[k];
    }
    const bool tail = row_segmented_sum(s, rp);
    if (tail && rp != kNoRow) atomicAdd(&acc[rp], s);
}

// Marks the lane's K positions as past the batch's end (before the next
// batch's expansion overwrites the taken ones).
template <int K>
__device__ __forceinline__ void clear_rows(unsigned char* row, int lane) {
    if constexpr (K % 2 == 0) {
        unsigned short* rr = reinterpret_cast<unsigned short*>(row + lane * K);
#pragma unroll
        for (int k = 0; k < K / 2; ++k) rr[k] = (unsigned short)(kNoRow * 0x101);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) row[lane * K + k] = (unsigned char)kNoRow;
    }
}

// Raw header of one block in registers: the int32 first column and the r
// masks as 32-bit words exactly as loaded (a byte or short load already
// zero-extends), widened only where they are first used.
template <int R, int MB>
struct HeaderRegs {
    static constexpr int kBytes = R * MB;
    static constexpr int kWords = kBytes <= 4 ? 1 : kBytes / 4;
    // A byte or short mask is read as the (unaligned) dword of the header
    // that ends with it, and shifted down only in take_batch: an extension
    // written next to the load would be scheduled right after it and wait.
    int cs = 0;
    unsigned m[kWords] = {};
};

template <int NW>
struct BatchState {
    BatchShape bs;
    int cs;
    unsigned long long mw[NW];
};

// Header of block min(j, last) into raw registers, no branch around the
// loads (see load_header_clamped).
template <int R, int MB>
__device__ __forceinline__ void load_header_raw(const unsigned char* __restrict__ hdr, int j, int last,
                                                HeaderRegs<R, MB>& h) {
    constexpr int B = R * MB;
    const int jj = j < last ? j : last;
    const unsigned char* p = hdr + (int64_t)jj * (4 + B);
    h.cs = load_i32_unaligned(p);
    if constexpr (B <= 2) {
        __builtin_memcpy(&h.m[0], p + B, 4);  // header bytes B .. B+3; the mask is the top B bytes
    } else {
#pragma unroll
        for (int k = 0; k < B / 4; ++k) __builtin_memcpy(&h.m[k], p + 4 + 4 * k, 4);
    }
}

// Shape of the batch starting at block b from its (prefetched) headers;
// lanes past the tile's last block count as empty.  This is the first use
// of the header registers, so their wait happens here.
template <int R, int MB, int NW, int CAP>
__device__ __forceinline__ BatchState<NW> take_batch(const HeaderRegs<R, MB>& h, int b, int b1, int lane) {
    using H = HeaderRegs<R, MB>;
    BatchState<NW> st;
    // The empty asm reads the header registers: it keeps their first use
    // (and so the wait for them) at this point, after earlier requests.
    int cs = h.cs;
    unsigned m[H::kWords];
    if constexpr (H::kBytes == 1) {
        asm volatile("v_lshrrev_b32 %0, 24, %1" : "=v"(m[0]) : "v"(h.m[0]));
    } else if constexpr (H::kBytes == 2) {
        asm volatile("v_lshrrev_b32 %0, 16, %1" : "=v"(m[0]) : "v"(h.m[0]));
    } else {
#pragma unroll
        for (int k = 0; k < H::kWords; ++k) {
            m[k] = h.m[k];
            asm volatile("" : "+v"(m[k]));
        }
    }
    asm volatile("" : "+v"(cs));
    const bool valid = b + lane < b1;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        unsigned long long w = 0;
        if constexpr (H::kWords == 1) {
            w = m[0];
        } else {
            w = (unsigned long long)m[2 * k] | ((unsigned long long)m[2 * k + 1] << 32);
        }
        st.mw[k] = valid ? w : 0ull;
    }
    st.cs = valid ? cs : 0;
    st.bs = batch_shape<NW, CAP>(st.mw, valid);
    return st;
}

// Values of one batch through a buffer resource based at its first value
// (v, wave-uniform) and `total` values long: lane offsets are constants,
// and the hardware range check returns 0 for the positions past the
// batch's end without reading them.
// Raw value bits as loaded (Raw) and their conversion to the accumulation
// type, done where the product is formed: a conversion next to the load
// (inside the branch that skips chunks past the batch's end) would wait for
// the load there.
template <class V>
struct ValueBuf;
template <>
struct ValueBuf<double> {
    using Raw = double;
    __device__ __forceinline__ static Raw get(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kNonTemporal));
    }
    template <class ACC>
    __device__ __forceinline__ static ACC acc(Raw v) { return (ACC)v; }
};
template <>
struct ValueBuf<float> {
    using Raw = float;
    __device__ __forceinline__ static Raw get(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, kNonTemporal));
    }
    template <class ACC>
    __device__ __forceinline__ static ACC acc(Raw v) { return (ACC)v; }
};
template <>
struct ValueBuf<unsigned short> {
    // One 32-bit register per value (the load zero-extends): 16-bit
    // registers would be packed in pairs, and the packing waits for the load.
    using Raw = unsigned int;
    __device__ __forceinline__ static Raw get(__amdgpu_buffer_rsrc_t r, int off) {
        return (unsigned int)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, kNonTemporal);
    }
    template <class ACC>
    __device__ __forceinline__ static ACC acc(Raw v) { return (ACC)__uint_as_float(v << 16); }
};

template <class V, int K>
__device__ __forceinline__ void load_values_buf(const V* __restrict__ val, int v, int total, int lane,
                                                typename ValueBuf<V>::Raw* a) {
    const int vs = __builtin_amdgcn_readfirstlane(v);
    const int nrec = __builtin_amdgcn_readfirstlane(total) * (int)sizeof(V);  // total < 1024
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<V*>(val + vs), (short)0, nrec, kRawBufferFlags);
    const int tot = __builtin_amdgcn_readfirstlane(total);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        // Chunks wholly past the batch's end are not issued (uniform branch);
        // their registers are never read (reduce_batch skips the same chunks).
        if (64 * k < tot) a[k] = ValueBuf<V>::get(r, (lane + 64 * k) * (int)sizeof(V));
    }
}

// x for every value of one expanded batch: columns of every chunk first
// (past the batch's end: column 0, a valid x entry whose product is never
// added), then all K requests.  No branch, so that nothing waits between
// the requests.  kWideX false: the LDS entries are byte offsets from x
// (x spans less than 4 GiB), read as 32-bit offsets from the x base.
// slot[k] is the position of the product of value lane + 64k (row-sorted
// batches: its row-sorted position; past the batch's end: its own, a
// position no row sum reads).
template <class X, int K, bool kSorted, bool kWideX, class SM>
__device__ __forceinline__ void request_x(SM& sm, const X* __restrict__ x, int total, int lane, X* xv, int* slot) {
    int cl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = lane + 64 * k;
        if constexpr (kSorted) {
            const int2 e = sm.cp[q];
            cl[k] = q < total ? e.x : 0;
            slot[k] = q < total ? e.y : q;
        } else {
            const int c = sm.col[q];
            cl[k] = q < total ? c : 0;
            slot[k] = q;
        }
    }
    wave_sync();  // every column is read before products overwrite them
    const int tot = __builtin_amdgcn_readfirstlane(total);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (64 * k >= tot) continue;  // chunk wholly past the batch's end (uniform; never read)
        if constexpr (kWideX) {
            xv[k] = x[cl[k]];
        } else {
            xv[k] = *reinterpret_cast<const X*>(reinterpret_cast<const char*>(x) + (unsigned)cl[k]);
        }
    }
}

// Products of one batch (values a, x xv) at their positions, then the row
// sums into the tile's accumulators.
template <class V, class X, class ACC, int K, class SM>
__device__ __forceinline__ void reduce_batch(SM& sm, const typename ValueBuf<V>::Raw* a, const X* xv,
                                             const int* slot, int total, int lane) {
    const int tot = __builtin_amdgcn_readfirstlane(total);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        // chunks past the end are never read as a run
        if (64 * k < tot) sm.prod[slot[k]] = ValueBuf<V>::template acc<ACC>(a[k]) * (ACC)xv[k];
    }
    wave_sync();
    reduce_slots<ACC, K>(sm.acc, sm.row, sm.prod, lane);
    clear_rows<K>(sm.row, lane);
}

template <class V, class X, class ACC, int R, int VS, int MB, bool kSort, bool kWideX>
__device__ __forceinline__ void spmv_tile_body(TileSmem<ACC, R, expand_cap<V>(), kSort>& sm,
                                               const int* __restrict__ bptr,
                                               const unsigned char* __restrict__ hdr, const V* __restrict__ val,
                                               const X* __restrict__ x, X* __restrict__ y, int br0, int br1,
                                               int vstart, int vend, int64_t nrows, const Epilogue& ep) {
    constexpr int CAP = expand_cap<V>();
    constexpr int K = CAP / 64;  // batch positions owned by one lane in the row sums
    constexpr int NW = MaskWords<R * MB>::kWords;
    constexpr bool kSorted = RowSort<R, kSort>::kOn;
    constexpr int kScale = kWideX ? 1 : (int)sizeof(X);  // LDS column entries: indices or byte offsets
    const int lane = lane_id();
    // The tile bounds are the same on every lane (the tile index comes from
    // the wave index, which the compiler cannot tell is uniform).
    br0 = __builtin_amdgcn_readfirstlane(br0);
    br1 = __builtin_amdgcn_readfirstlane(br1);
    vstart = __builtin_amdgcn_readfirstlane(vstart);
    vend = __builtin_amdgcn_readfirstlane(vend);

    const int nbr = br1 - br0;
    const int b0 = bptr[br0];
    const int b1 = bptr[br1];
    int v = vstart;
    const int64_t row0 = (int64_t)br0 * R;
    const int64_t rows_left = nrows - row0;
    const int nrow_t = (int)((int64_t)nbr * R < rows_left ? (int64_t)nbr * R : rows_left);

    for (int i = lane; i <= nbr; i += 64) sm.rs[i] = bptr[br0 + i] - b0;
    for (int i = lane; i < nbr * R; i += 64) sm.acc[i] = (ACC)0;
    clear_rows<K>(sm.row, lane);
    int top = 1;
    while (top < nbr) top <<= 1;

    int b = b0;
    int prev = -1;  // block-row of block b-1
    HeaderRegs<R, MB> h;
    if (vstart == vend) b = b1;  // no values: x is never read
    if (b < b1) load_header_raw<R, MB>(hdr, b + lane, b1 - 1, h);
    wave_sync();
    while (b < b1) {
        const BatchState<NW> cur = take_batch<R, MB, NW, CAP>(h, b, b1, lane);
        typename ValueBuf<V>::Raw a[K];
        load_values_buf<V, K>(val, v, cur.bs.total, lane, a);
        const int bn = b + cur.bs.ntake;
        load_header_raw<R, MB>(hdr, bn + lane, b1 - 1, h);  // past the end: block b1-1, masked
        const int br = batch_blockrows(sm.rs, sm.mark, nbr, top, b - b0, cur.bs.ntake, prev, lane);
        prev = __builtin_amdgcn_readlane(br, cur.bs.ntake - 1);
        // First column of the block, as an index or as a byte offset from x.
        const int cso = (int)((unsigned)cur.cs * (unsigned)kScale);
        if constexpr (kSorted) {
            int d[R];
            row_displacements<R, VS, NW>(sm.segs, cur.mw, br, cur.bs.ntake, cur.bs.incl - cur.bs.cnt, lane, d);
            if (lane < cur.bs.ntake)
                expand_rows_sorted<R, VS, NW, kScale>(sm.cp, sm.row, br * R, cso, cur.mw,
                                                      cur.bs.incl - cur.bs.cnt, d);
        } else {
            if (lane < cur.bs.ntake)
                expand_rows<R, VS, NW, kScale>(sm.col, sm.row, br * R, cso, cur.mw, cur.bs.incl - cur.bs.cnt);
        }
        wave_sync();
        if (cur.bs.total > 0) {
            X xv[K];  // converted where the products are formed
            int slot[K];
            request_x<X, K, kSorted, kWideX>(sm, x, cur.bs.total, lane, xv, slot);
            reduce_batch<V, X, ACC, K>(sm, a, xv, slot, cur.bs.total, lane);
        }
        v += cur.bs.total;
        b = bn;
        wave_sync();
    }
    wave_sync();
    // y = alpha * (A x) + beta * y; y is not read when beta == 0.  With a dot
    // vector, the tile writes sum(dv[i] * y_new[i]) over its rows to *ep.dot.
    const ACC alpha = (ACC)ep.alpha, beta = (ACC)ep.beta;
    const X* __restrict__ dv = reinterpret_cast<const X*>(ep.dv);
    double part = 0.0;
    for (int i = lane; i < nrow_t; i += 64) {
        const ACC ax = sm.acc[i];
        const X yn = ep.beta == 0.0 ? (X)(alpha * ax) : (X)(alpha * ax + beta * (ACC)y[row0 + i]);
        y[row0 + i] = yn;
        if (dv) part += (double)dv[row0 + i] * (double)yn;
    }
    if (ep.dot) {
        part = wave_sum_f64(part);
        if (lane == 0) ep.dot[0] = part;
    }
}

// One tile; x is addressed by 32-bit byte offsets unless it spans 4 GiB or
// more (ncols * sizeof(X)), a uniform choice per launch.
template <class V, class X, class ACC, int R, int VS, int MB, bool kSort = false>
__device__ __forceinline__ void spmv_tile(TileSmem<ACC, R, expand_cap<V>(), kSort>& sm,
                                          const int* __restrict__ bptr,
                                          const unsigned char* __restrict__ hdr, const V* __restrict__ val,
                                          const X* __restrict__ x, X* __restrict__ y, int br0, int br1, int vstart,
                                          int vend, int64_t nrows, int64_t ncols, const Epilogue& ep) {
    if (SPC5_X32 && (uint64_t)ncols * sizeof(X) <= 0xffffffffull)
        spmv_tile_body<V, X, ACC, R, VS, MB, kSort, false>(sm, bptr, hdr, val, x, y, br0, br1, vstart, vend, nrows,
                                                           ep);
    else
        spmv_tile_body<V, X, ACC, R, VS, MB, kSort, true>(sm, bptr, hdr, val, x, y, br0, br1, vstart, vend, nrows,
                                                          ep);
}

}  // namespace spc5
