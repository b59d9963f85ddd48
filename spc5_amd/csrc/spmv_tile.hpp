// spmv_tile.hpp -- one wave computes y = alpha A x + beta y over one tile (a
// run of whole block-rows) of a beta(r,VS) matrix in the reference layout.
//
// Per batch of up to 64 blocks (one block per lane):
//   1. the lane's block header (int32 first column + r masks) was loaded
//      ahead; popcounts and a DPP wave scan give every block's value offset
//      inside the batch; blocks are taken while the batch holds at most
//      kExpandCap values;
//   2. the batch's values are loaded at once, kExpandCap/64 coalesced loads
//      per lane, straight from the value stream (no zero padding);
//   3. the next batch's headers are requested;
//   4. the block-row of every block of the batch comes from one pass over
//      the tile's block-row pointers in LDS (the block-rows that start in
//      the batch mark their first block, a wave prefix maximum spreads the
//      marks); each taken lane then writes, for every set mask bit, the
//      column and the tile-local row into LDS;
//   5. in value order, every lane gathers x for its values (neighbouring
//      lanes read neighbouring columns) and forms the products;
//   6. row sums, two ways (SPC5_REDUCE):
//      1 (default) -- the products go to LDS slot-major, so that lane l then
//        owns the kSlots consecutive values l*kSlots ...; it adds them in
//        registers and does one LDS add per run of equal rows;
//      0 -- DPP segmented sums over runs of equal rows inside each 16-lane
//        row, one LDS add per run.
// At the end of the tile each row is written once: y = alpha*sum + beta*y.
#pragma once

#include "spc5_mat.hpp"

namespace spc5 {

#ifndef SPC5_REDUCE
#define SPC5_REDUCE 1
#endif


template <class ACC>
struct SlotLayout {
    static constexpr int kSlots = kExpandCap / 64;  // values per lane in the row-run phase
    // Slot-major product layout: value q = l*kSlots + k lives at k*kStride + l.
    // The extra entries put the 64 lanes of one value-order store on
    // distinct banks.
    static constexpr int kStride = sizeof(ACC) == 8 ? 68 : 72;
    __device__ __forceinline__ static int pos(int q) { return (q % kSlots) * kStride + q / kSlots; }
};

template <class ACC, int R>
struct TileSmem {
    ACC acc[kMaxRows];  // row sums of the tile
    // Columns in value order; every column is read before the first
    // product is stored over them.
    union {
        int col[kExpandCap];
#if SPC5_REDUCE
        ACC prod[SlotLayout<ACC>::kSlots * SlotLayout<ACC>::kStride];
#endif
    };
    alignas(8) unsigned char row[kExpandCap];  // tile-local row of batch value q
    int rs[kMaxRows / R + 1];
    int mark[64];  // block-row marks of one batch
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

template <int NW>
__device__ __forceinline__ BatchShape batch_shape(const unsigned long long* mw, bool valid) {
    BatchShape bs;
    bs.cnt = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) bs.cnt += __popcll(mw[k]);
    bs.incl = wave_incl_scan(bs.cnt);
    bs.ntake = __popcll(ballot64(valid && bs.incl <= kExpandCap));
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
// block's block-row.
template <int R, int VS, int NW>
__device__ __forceinline__ void expand_rows(int* col, unsigned char* row, int rbase, int cs,
                                            const unsigned long long* mw, int p) {
    if constexpr (R * VS <= 16) {
        // Few mask bits: test each one (no data-dependent trip count).
        const unsigned m = (unsigned)mw[0];
#pragma unroll
        for (int bit = 0; bit < R * VS; ++bit) {
            if ((m >> bit) & 1u) {
                col[p] = cs + (bit % VS);
                row[p] = (unsigned char)(rbase + bit / VS);
                ++p;
            }
        }
    } else if constexpr (R * VS <= 32) {
        unsigned m = (unsigned)mw[0];
        while (m) {
            const int bit = __builtin_ctz(m);
            col[p] = cs + (bit & (VS - 1));
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
                col[p] = cs + (bit & (VS - 1));
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

// Row sums of one batch by lane-owned runs: lane l owns batch values
// l*K .. l*K+K-1 (products in LDS, slot-major).  Consecutive values of one
// row are added in registers; every run of equal rows ends with one LDS add.
template <class ACC, int K>
__device__ __forceinline__ void reduce_slots(ACC* acc, const unsigned char* row, const ACC* prod, int total,
                                             int lane) {
    using L = SlotLayout<ACC>;
    constexpr int kNone = kMaxRows;  // row of the slots past the batch's end
    ACC p[K];
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = prod[k * L::kStride + lane];  // all reads before the first add
    int rw[K];
    const int nv = total - lane * K;  // valid slots of this lane (<= 0: none, >= K: all)
    if constexpr (K == 8) {
        unsigned long long word;
        __builtin_memcpy(&word, row + lane * 8, 8);
#pragma unroll
        for (int k = 0; k < 8; ++k) rw[k] = k < nv ? (int)((word >> (8 * k)) & 0xffu) : kNone;
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) rw[k] = k < nv ? (int)row[lane * K + k] : kNone;
    }
    // Slots past the end form a suffix, so a run is only ever closed on a
    // real row.  Runs that end inside the lane go to LDS at once; the lane's
    // last run is first summed with the last runs of neighbouring lanes on
    // the same row (DPP segmented sum inside 16-lane rows), so that a row
    // spread over many lanes costs one LDS add, not one per lane.  Runs past
    // the end (row kNone) are dropped.
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
    if (tail && rp != kNone) atomicAdd(&acc[rp], s);
}

template <class ACC, int K>
__device__ __forceinline__ void accumulate_dpp(ACC* acc, int total, const ACC* a, const ACC* xv, const int* rw) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (64 * k >= total) break;  // wave-uniform
        const bool ok = rw[k] >= 0;
        // Runs of one row in adjacent lanes are summed first (DPP), so one
        // LDS atomic per run replaces one per value.
        ACC pr = ok ? a[k] * xv[k] : (ACC)0;
        const bool tail = row_segmented_sum(pr, rw[k]);
        if (ok && tail) atomicAdd(&acc[rw[k]], pr);
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
template <int R, int MB, int NW>
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
    st.bs = batch_shape<NW>(st.mw, valid);
    return st;
}

// x for every value of one expanded batch: columns of every chunk first
// (past the batch's end: column 0, a valid x entry whose product with the
// zero value is unused), then all K requests.  No branch, so that nothing
// waits between the requests.
template <class X, class ACC, int K, class SM>
__device__ __forceinline__ void request_x(SM& sm, const X* __restrict__ x, int total, int lane, ACC* xv) {
    int cl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = lane + 64 * k;
        const int c = sm.col[q];
        cl[k] = q < total ? c : 0;
    }
    wave_sync();  // every column is read before products overwrite them
#pragma unroll
    for (int k = 0; k < K; ++k) xv[k] = (ACC)x[cl[k]];
}

// Products of one batch (values a, x xv) and their row sums into the
// tile's accumulators.  Chunks past the batch's end store unused products.
template <class ACC, int K, class SM>
__device__ __forceinline__ void reduce_batch(SM& sm, const ACC* a, const ACC* xv, int total, int lane) {
#pragma unroll
    for (int k = 0; k < K; ++k) sm.prod[SlotLayout<ACC>::pos(lane + 64 * k)] = a[k] * xv[k];
    wave_sync();
    reduce_slots<ACC, K>(sm.acc, sm.row, sm.prod, total, lane);
}

template <class V, class X, class ACC, int R, int VS, int MB>
__device__ __forceinline__ void spmv_tile(TileSmem<ACC, R>& sm, const int* __restrict__ bptr,
                                          const unsigned char* __restrict__ hdr, const V* __restrict__ val,
                                          const X* __restrict__ x, X* __restrict__ y, int br0, int br1, int vstart,
                                          int vend, int64_t nrows, int64_t ncols, const Epilogue& ep) {
    constexpr int K = kExpandCap / 64;
    constexpr int NW = MaskWords<R * MB>::kWords;
    const int lane = lane_id();

    const int nbr = br1 - br0;
    const int b0 = bptr[br0];
    const int b1 = bptr[br1];
    int v = vstart;
    const int64_t row0 = (int64_t)br0 * R;
    const int64_t rows_left = nrows - row0;
    const int nrow_t = (int)((int64_t)nbr * R < rows_left ? (int64_t)nbr * R : rows_left);

    for (int i = lane; i <= nbr; i += 64) sm.rs[i] = bptr[br0 + i] - b0;
    for (int i = lane; i < nbr * R; i += 64) sm.acc[i] = (ACC)0;
    int top = 1;
    while (top < nbr) top <<= 1;

    int b = b0;
    int prev = -1;  // block-row of block b-1
    HeaderRegs<R, MB> h;
    if (vstart == vend) b = b1;  // no values: x is never read
    if (b < b1) load_header_raw<R, MB>(hdr, b + lane, b1 - 1, h);
    wave_sync();
    while (b < b1) {
        const BatchState<NW> cur = take_batch<R, MB, NW>(h, b, b1, lane);
        ACC a[K];
        load_values<V, ACC, K>(val, v, cur.bs.total, lane, a);
        const int bn = b + cur.bs.ntake;
        load_header_raw<R, MB>(hdr, bn + lane, b1 - 1, h);  // past the end: block b1-1, masked
        const int br = batch_blockrows(sm.rs, sm.mark, nbr, top, b - b0, cur.bs.ntake, prev, lane);
        prev = __builtin_amdgcn_readlane(br, cur.bs.ntake - 1);
        if (lane < cur.bs.ntake)
            expand_rows<R, VS, NW>(sm.col, sm.row, br * R, cur.cs, cur.mw, cur.bs.incl - cur.bs.cnt);
        wave_sync();
        if (cur.bs.total > 0) {
            ACC xv[K];
            request_x<X, ACC, K>(sm, x, cur.bs.total, lane, xv);
#if SPC5_REDUCE
            reduce_batch<ACC, K>(sm, a, xv, cur.bs.total, lane);
#else
            int rw[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int q = lane + 64 * k;
                rw[k] = q < cur.bs.total ? (int)sm.row[q] : -1;
            }
            accumulate_dpp<ACC, K>(sm.acc, cur.bs.total, a, xv, rw);
#endif
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

}  // namespace spc5
