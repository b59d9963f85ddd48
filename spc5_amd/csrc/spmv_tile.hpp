// spmv_tile.hpp -- one wave computes y += A x over one tile (a run of whole
// block-rows) of a beta(r,VS) matrix in the reference layout.
//
// Per batch of up to 64 blocks (one block per lane):
//   1. the lane's block header (int32 first column + r masks) was loaded in
//      the previous iteration (prefetch); popcounts and a DPP wave scan give
//      every block's value offset inside the batch; blocks are taken while
//      the batch holds at most kExpandCap values;
//   2. the batch's values are loaded at once, kExpandCap/64 coalesced loads
//      per lane, straight from the value stream (no zero padding);
//   3. the next batch's headers are requested;
//   4. each taken lane finds its block-row (binary search over the tile's
//      block-row pointers in LDS) and writes, for every set mask bit, the
//      column and the tile-local row into LDS;
//   5. every lane gathers x for its values, multiplies and adds into the
//      tile's per-row LDS accumulators (ds_add);
// and at the end of the tile each row does one y[row] += sum.
#pragma once

#include "spc5_mat.hpp"

namespace spc5 {

#ifndef SPC5_PIPELINE
#define SPC5_PIPELINE 0
#endif

template <class ACC, int R>
struct TileSmem {
    ACC acc[kMaxRows];
    int rs[kMaxRows / R + 1];
    int col[1 + SPC5_PIPELINE][kExpandCap];
    unsigned char row[1 + SPC5_PIPELINE][kExpandCap];
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
    bs.total = __builtin_amdgcn_readlane(bs.incl, bs.ntake - 1);
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

// Every set mask bit of the taken blocks -> (column, tile-local row) in LDS.
template <int R, int VS, int NW>
__device__ __forceinline__ void expand(int* col, unsigned char* row, const int* rs, int nbr, int top, int jr,
                                       int cs, const unsigned long long* mw, int p) {
    const int rbase = find_blockrow(rs, nbr, top, jr) * R;
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

template <class X, class ACC, int K>
__device__ __forceinline__ void gather_x(const int* col, const unsigned char* row, const X* __restrict__ x,
                                         int total, int lane, ACC* xv, int* rw) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = lane + 64 * k;
        xv[k] = (ACC)0;
        rw[k] = -1;
        if (q < total) {
            rw[k] = row[q];
            xv[k] = (ACC)x[col[q]];
        }
    }
}

template <class ACC, int K>
__device__ __forceinline__ void accumulate(ACC* acc, int total, const ACC* a, const ACC* xv, const int* rw) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (64 * k >= total) break;  // wave-uniform
        const bool ok = rw[k] >= 0;
#if SPC5_SEGREDUCE
        // Runs of one row in adjacent lanes are summed first (DPP), so one
        // LDS atomic per run replaces one per value.
        ACC pr = ok ? a[k] * xv[k] : (ACC)0;
        const bool tail = row_segmented_sum(pr, rw[k]);
        if (ok && tail) atomicAdd(&acc[rw[k]], pr);
#else
        if (ok) atomicAdd(&acc[rw[k]], a[k] * xv[k]);
#endif
    }
}

// ---------------------------------------------------------------------------
// Block-stream tile (the default).  Per batch of up to 64 blocks:
//   * values of the batch: coalesced loads from the stream into LDS;
//   * each taken lane loads its block's x window x[start .. start+VS) (the
//     reference's unaligned VS-wide x load) into its own LDS row;
//   * each lane then walks its block's set mask bits in stream order (rows
//     in order, columns ascending): sum += value * x_window[bit % VS] from
//     LDS, and adds the sum to the tile's row accumulator when the row
//     changes.
// No per-value decode, gather or cross-lane reduction: the work per value is
// two LDS reads and an FMA.
// ---------------------------------------------------------------------------
template <class X, class ACC, int R, int VS>
struct StreamSmem {
    static constexpr int kXStride = VS + 1;  // padded per-lane x window (bank spread)
    ACC acc[kMaxRows];
    int rs[kMaxRows / R + 1];
    ACC val[kExpandCap];
    X xw[64 * kXStride];
};

template <class V, class X, class ACC, int R, int VS, int MB>
__device__ __forceinline__ void spmv_tile_stream(StreamSmem<X, ACC, R, VS>& sm, const int* __restrict__ bptr,
                                                 const unsigned char* __restrict__ hdr, const V* __restrict__ val,
                                                 const X* __restrict__ x, X* __restrict__ y, int br0, int br1,
                                                 int vstart, int64_t nrows, int64_t ncols, const Epilogue& ep) {
    constexpr int K = kExpandCap / 64;
    constexpr int NW = MaskWords<R * MB>::kWords;
    constexpr int XS = StreamSmem<X, ACC, R, VS>::kXStride;
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
    int cs;
    unsigned long long mw[NW];
    load_header<R, MB>(hdr, b + lane, b1, cs, mw);
    wave_sync();
    X* xw = &sm.xw[lane * XS];

    while (b < b1) {
        const BatchShape bs = batch_shape<NW>(mw, b + lane < b1);
        // values of the batch -> LDS (coalesced)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (64 * k >= bs.total) break;  // wave-uniform
            const int q = lane + 64 * k;
            if (q < bs.total) sm.val[q] = ValueLoad<V, ACC>::get(val + v + q);
        }
        const bool taken = lane < bs.ntake;
        const int jcur = b + lane;
        const int cs_cur = cs;
        unsigned long long m_cur[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) m_cur[k] = mw[k];
        // the block's x window -> LDS
        if (taken) {
            if ((int64_t)cs_cur + VS <= ncols) {
                X w[VS];
                __builtin_memcpy(w, x + cs_cur, sizeof(w));
#pragma unroll
                for (int j = 0; j < VS; ++j) xw[j] = w[j];
            } else {
#pragma unroll
                for (int j = 0; j < VS; ++j) {
                    const int64_t c = (int64_t)cs_cur + j < ncols ? (int64_t)cs_cur + j : ncols - 1;
                    xw[j] = x[c];
                }
            }
        }
        const int bn = b + bs.ntake;
        load_header<R, MB>(hdr, bn + lane, b1, cs, mw);
        const int rbase = taken ? find_blockrow(sm.rs, nbr, top, jcur - b0) * R : 0;
        wave_sync();
        if (taken) {
            int p = bs.incl - bs.cnt;
            ACC sum = (ACC)0;
            if constexpr (R == 1 && R * MB <= 4) {
                unsigned m = (unsigned)m_cur[0];
                while (m) {
                    const int bit = __builtin_ctz(m);
                    sum = __builtin_fma(sm.val[p], (ACC)xw[bit], sum);
                    ++p;
                    m &= m - 1u;
                }
                if (bs.cnt) atomicAdd(&sm.acc[rbase], sum);
            } else if constexpr (R * MB <= 4) {
                unsigned m = (unsigned)m_cur[0];
                int cur = -1;
                while (m) {
                    const int bit = __builtin_ctz(m);
                    const int s = bit / VS;
                    if (s != cur) {
                        if (cur >= 0) atomicAdd(&sm.acc[rbase + cur], sum);
                        cur = s;
                        sum = (ACC)0;
                    }
                    sum = __builtin_fma(sm.val[p], (ACC)xw[bit & (VS - 1)], sum);
                    ++p;
                    m &= m - 1u;
                }
                if (cur >= 0) atomicAdd(&sm.acc[rbase + cur], sum);
            } else {
                int cur = -1;
#pragma unroll
                for (int wd = 0; wd < NW; ++wd) {
                    unsigned long long m = m_cur[wd];
                    while (m) {
                        const int bit = __builtin_ctzll(m) + wd * 64;
                        const int s = bit / VS;
                        if (s != cur) {
                            if (cur >= 0) atomicAdd(&sm.acc[rbase + cur], sum);
                            cur = s;
                            sum = (ACC)0;
                        }
                        sum = __builtin_fma(sm.val[p], (ACC)xw[bit & (VS - 1)], sum);
                        ++p;
                        m &= m - 1ull;
                    }
                }
                if (cur >= 0) atomicAdd(&sm.acc[rbase + cur], sum);
            }
        }
        v += bs.total;
        b = bn;
        wave_sync();
    }
    wave_sync();
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

template <class V, class X, class ACC, int R, int VS, int MB>
__device__ __forceinline__ void spmv_tile(TileSmem<ACC, R>& sm, const int* __restrict__ bptr,
                                          const unsigned char* __restrict__ hdr, const V* __restrict__ val,
                                          const X* __restrict__ x, X* __restrict__ y, int br0, int br1, int vstart,
                                          int64_t nrows, const Epilogue& ep) {
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
    int cs;
    unsigned long long mw[NW];
    load_header<R, MB>(hdr, b + lane, b1, cs, mw);
    wave_sync();

#if SPC5_PIPELINE
    // Two batches in flight: batch i's x gathers travel while batch i+1 is
    // scanned, its values requested and its blocks expanded into the other
    // LDS buffer.
    ACC a_cur[K], xv_cur[K];
    int rw_cur[K];
    int total_cur = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        a_cur[k] = (ACC)0;
        xv_cur[k] = (ACC)0;
        rw_cur[k] = -1;
    }
    int buf = 0;
    while (b < b1) {
        const BatchShape bs = batch_shape<NW>(mw, b + lane < b1);
        ACC a_nxt[K];
        load_values<V, ACC, K>(val, v, bs.total, lane, a_nxt);
        const int jcur = b + lane;
        const int cs_cur = cs;
        unsigned long long mw_cur[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) mw_cur[k] = mw[k];
        const int bn = b + bs.ntake;
        load_header<R, MB>(hdr, bn + lane, b1, cs, mw);
        if (lane < bs.ntake)
            expand<R, VS, NW>(sm.col[buf], sm.row[buf], sm.rs, nbr, top, jcur - b0, cs_cur, mw_cur,
                              bs.incl - bs.cnt);
        wave_sync();
        accumulate<ACC, K>(sm.acc, total_cur, a_cur, xv_cur, rw_cur);
        gather_x<X, ACC, K>(sm.col[buf], sm.row[buf], x, bs.total, lane, xv_cur, rw_cur);
#pragma unroll
        for (int k = 0; k < K; ++k) a_cur[k] = a_nxt[k];
        total_cur = bs.total;
        v += bs.total;
        b = bn;
        buf ^= 1;
        wave_sync();
    }
    accumulate<ACC, K>(sm.acc, total_cur, a_cur, xv_cur, rw_cur);
#else
    while (b < b1) {
        const BatchShape bs = batch_shape<NW>(mw, b + lane < b1);
        ACC a[K];
        load_values<V, ACC, K>(val, v, bs.total, lane, a);
        const int jcur = b + lane;
        const int cs_cur = cs;
        unsigned long long mw_cur[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) mw_cur[k] = mw[k];
        const int bn = b + bs.ntake;
        load_header<R, MB>(hdr, bn + lane, b1, cs, mw);
        if (lane < bs.ntake)
            expand<R, VS, NW>(sm.col[0], sm.row[0], sm.rs, nbr, top, jcur - b0, cs_cur, mw_cur, bs.incl - bs.cnt);
        wave_sync();
        ACC xv[K];
        int rw[K];
        gather_x<X, ACC, K>(sm.col[0], sm.row[0], x, bs.total, lane, xv, rw);
        accumulate<ACC, K>(sm.acc, bs.total, a, xv, rw);
        v += bs.total;
        b = bn;
        wave_sync();
    }
#endif
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
