// spmm.hip -- multi-right-hand-side product Y = alpha A X + beta Y for a
// beta(r,VS) matrix and row-major dense blocks X (ncols x k, leading
// dimension ldx) and Y (nrows x k, ldy).  The batching form; no counterpart
// in the reference.
//
// One wave per (tile, column chunk of kt <= 64 columns); the grid lists the
// pairs chunk by chunk and each XCD walks one contiguous eighth of the list
// (spmm_item).  Per batch of up to 64 blocks the tile's headers are decoded
// as in the SpMV tile (headers prefetched, values read once, coalesced, from
// the value stream) and every value becomes a 16-byte LDS record (value,
// column, tile-local row), or with column records every set column of a
// block one record (k_spmm_cols).  The wave then splits into groups of
// KC / CPL lanes, each lane taking CPL adjacent columns of the chunk (one
// element, or 16 bytes of the X row with SPC5_HIP_OPT_SPMM_X16); each group
// walks a contiguous run of the batch's records, requests the X rows of U
// records at a time (one contiguous kt-wide row segment per record; with
// 64-column chunks of float or bfloat16 X and one column per lane the whole
// wave is one group and the records are broadcast to scalar registers) and
// keeps the running row sums in registers, adding them to the tile's LDS
// row x column sums only when the row changes.  With kt = k the matrix is
// fetched once for all k columns; smaller kt re-reads the matrix k/kt times
// with a smaller X working set (the autotuner picks kt).
#include "spc5_mat.hpp"
#include "spmv_tile.hpp"

#include <map>
#include <mutex>

using namespace spc5;

namespace spc5 {
namespace {

// Float and bfloat16 values accumulate their row x column sums in double:
// the LDS float atomic add (ds_add_f32) of one 64-lane row per record run
// left the float kernel waiting on LDS for half of its cycles (k = 64 on the
// 27-point stencil: 13.5 ms with float sums, 4.7 ms with double sums, whose
// ds_add_f64 is fast).
#ifndef SPC5_SPMM_FLOAT_ACC_DOUBLE
#define SPC5_SPMM_FLOAT_ACC_DOUBLE 1
#endif
using SpmmFloatAcc = std::conditional<SPC5_SPMM_FLOAT_ACC_DOUBLE != 0, double, float>::type;

// X rows requested together by one record group before the first use
// waits: 16 with 64-column chunks (one group per wave, LDS-bound occupancy;
// k = 64 on the 27-point stencil 5.17 -> 4.87 ms in double), 8 with narrower
// chunks, where 16 measured slower (more registers, and groups with fewer
// than 16 records per batch fall back to one request at a time).
#ifndef SPC5_SPMM_U
#define SPC5_SPMM_U 8
#endif
#ifndef SPC5_SPMM_U64
#define SPC5_SPMM_U64 16
#endif

#ifndef SPC5_SPMM_CAP
#define SPC5_SPMM_CAP 256
#endif
#ifndef SPC5_SPMM_CAP64_NARROW
#define SPC5_SPMM_CAP64_NARROW 128
#endif
// Values expanded per batch.  64-column chunks of 4- or 2-byte X use the
// smaller cap: their records are scalars, LDS then limits occupancy, and
// k = 64 float on the 27-point stencil went 2.99 -> 2.59 ms (narrower
// chunks measured slower with it).  Double X keeps vector records: the
// scalar form measured 2 % slower there (its 512-byte X rows, not
// occupancy, set the pace).
template <int KC, class X, int CPL = 1>
constexpr int spmm_cap() {
    return KC == 64 && CPL == 1 && sizeof(X) < 8 ? SPC5_SPMM_CAP64_NARROW : SPC5_SPMM_CAP;
}
// a batch must hold the largest block (r = 8, VS = 16: 128 values)
static_assert(SPC5_SPMM_CAP >= 128 && SPC5_SPMM_CAP % 64 == 0, "SPC5_SPMM_CAP: a multiple of 64, at least 128");
static_assert(SPC5_SPMM_CAP64_NARROW >= 128 && SPC5_SPMM_CAP64_NARROW % 64 == 0,
              "SPC5_SPMM_CAP64_NARROW: a multiple of 64, at least 128");
// LDS for a tile's row x column sums; it sets the tile height (8 to 128
// rows).  The loop waits on X rows, so occupancy matters most: on the
// 27-point stencil 8192 bytes took 4.88 / 1.83 / 0.83 ms in double for
// k = 64 / 16 / 4, 2048 bytes take 4.39 / 1.38 / 0.74 ms (1024 measured the
// same as 2048, a 128-value batch cap slower in double).
#ifndef SPC5_SPMM_ACC_BYTES
#define SPC5_SPMM_ACC_BYTES 2048
#endif
constexpr int kSpmmAccBytes = SPC5_SPMM_ACC_BYTES;
constexpr int kSpmmMaxKt = 64;       // columns per pass (one lane each)


// One value of a batch: the value, where its X row is (a byte offset from
// X when X spans less than 4 GiB, else the column) and its tile-local row.
template <class ACC>
struct alignas(16) SpmmRec {
    ACC v;
    int col;
    int row;
};
static_assert(sizeof(SpmmRec<float>) == 16 && sizeof(SpmmRec<double>) == 16, "16-byte records");

template <class ACC>
int spmm_max_rows(int kt) {
    int rows = kSpmmAccBytes / (kt * (int)sizeof(ACC));
    if (rows > 128) rows = 128;
    if (rows < 8) rows = 8;
    return rows;
}

template <class ACC, int KC, class X, int CPL>
size_t spmm_smem_bytes(int maxrows, int kt, int r) {
    return sizeof(SpmmRec<ACC>) * spmm_cap<KC, X, CPL>() + sizeof(ACC) * (size_t)maxrows * kt +
           sizeof(int) * (maxrows / r + 1);
}

template <bool WIDE, class X>
__device__ __forceinline__ X spmm_x(const X* __restrict__ xcol, const X* __restrict__ x, int col, int64_t ldx,
                                    unsigned lane_off) {
    if constexpr (WIDE) {
        return xcol[(int64_t)col * ldx];
    } else {
        return *reinterpret_cast<const X*>(reinterpret_cast<const char*>(x) + ((unsigned)col + lane_off));
    }
}

// CPL adjacent X columns of one record: one X value, or (CPL = 16 /
// sizeof(X)) one 16-byte load of two doubles or four floats; the caller has
// checked that every X row start and every lane's first column are then
// 16-byte aligned.
template <int CPL, bool WIDE, class X, class ACC>
__device__ __forceinline__ void spmm_xs(const X* __restrict__ xcol, const X* __restrict__ x, int col, int64_t ldx,
                                        unsigned lane_off, ACC* out) {
    if constexpr (CPL == 1) {
        out[0] = (ACC)spmm_x<WIDE>(xcol, x, col, ldx, lane_off);
    } else {
        static_assert(CPL * sizeof(X) == 16, "one 16-byte load per lane");
        const X* p = WIDE ? xcol + (int64_t)col * ldx
                          : reinterpret_cast<const X*>(reinterpret_cast<const char*>(x) + ((unsigned)col + lane_off));
        const uint4 w = *reinterpret_cast<const uint4*>(p);
        X v[CPL];
        __builtin_memcpy(v, &w, 16);
#pragma unroll
        for (int e = 0; e < CPL; ++e) out[e] = (ACC)v[e];
    }
}

// The value of lane l (a compile-time or uniform lane index) on every lane.
__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The (tile, column chunk) pair of this wave.  The 1-D grid lists the pairs
// chunk by chunk, tiles in order, padded to a multiple of 8.  Workgroups are
// dealt to the 8 XCDs round-robin (group g runs on XCD g % 8), so groups are
// renumbered so that each XCD walks one contiguous eighth of that list: the
// tiles an XCD holds at once are neighbours, and the X rows they gather (the
// column footprint of one contiguous run of rows) stay in its own L2 instead
// of every XCD's L2 holding the footprint of the whole front.  False for the
// padding.
__device__ __forceinline__ bool spmm_item(int64_t ntiles, int k, int kt, int64_t& t, int& c0) {
    const unsigned g = blockIdx.x;
#if SPC5_XCD_REMAP
    const unsigned per_xcd = gridDim.x >> 3;
    const unsigned gx = (g & 7u) * per_xcd + (g >> 3);
#else
    const unsigned gx = g;
#endif
    const unsigned nt = (unsigned)ntiles;  // ntiles * chunks < 2^32 (checked at launch)
    const unsigned ch = gx / nt;
    t = (int64_t)(gx - ch * nt);
    c0 = (int)ch * kt;
    return (int64_t)ch * kt < k;  // false for the padding after the last chunk
}

// WIDE: X spans 4 GiB or more (or rows of 16 MiB or more), so records hold
// columns and X is addressed with 64-bit arithmetic; otherwise they hold
// 32-bit byte offsets of the X rows, computed with 24-bit multiplies.
// CPL: adjacent X columns per lane (1, or 16 / sizeof(X): one 16-byte X
// request), so KC / CPL lanes share a record.
template <class V, class X, class ACC, int R, int VS, int MB, int KC, bool WIDE, int CPL>
__global__ __launch_bounds__(64) void k_spmm_tiles(const int* __restrict__ bptr, const unsigned char* __restrict__ hdr,
                                                    const V* __restrict__ val, const X* __restrict__ x, int64_t ldx,
                                                    X* __restrict__ y, int64_t ldy, int k, int kt,
                                                    const int* __restrict__ tile_br, const int* __restrict__ tile_val,
                                                    int64_t ntiles, int64_t nrows, int maxrows, ACC alpha, ACC beta) {
    constexpr int CAP = spmm_cap<KC, X, CPL>();
    constexpr int K = CAP / 64;
    constexpr int NW = MaskWords<R * MB>::kWords;
    constexpr int LPR = KC / CPL;  // lanes per record
    constexpr int G = 64 / LPR;    // record groups per wave
    // records whose X rows are requested together (half as many with
    // several columns per lane and more than two groups per wave)
    constexpr int U = (LPR == 64 ? SPC5_SPMM_U64 : SPC5_SPMM_U) / (CPL > 1 && G > 2 ? 2 : 1);
    using Rec = SpmmRec<ACC>;
    extern __shared__ __align__(16) unsigned char smem[];
    Rec* rec = reinterpret_cast<Rec*>(smem);                            // [CAP]
    ACC* acc = reinterpret_cast<ACC*>(rec + CAP);                       // [maxrows * kt]
    int* rs = reinterpret_cast<int*>(acc + (size_t)maxrows * kt);       // [maxrows / R + 1]

    int64_t t;
    int c0;
    if (!spmm_item(ntiles, k, kt, t, c0)) return;
    const int ktc = (k - c0) < kt ? (k - c0) : kt;
    const int lane = lane_id();
    const int grp = lane / LPR;
    const int jl = (lane % LPR) * CPL;  // the lane's first column in the chunk
    const bool colok = jl < ktc;        // ktc is a multiple of CPL
    const X* __restrict__ xcol = x + c0 + (colok ? jl : 0);
    const unsigned lane_off = (unsigned)(c0 + (colok ? jl : 0)) * (unsigned)sizeof(X);  // byte offset in an X row
    const unsigned row_bytes = (unsigned)ldx * (unsigned)sizeof(X);
    const int br0 = tile_br[t], br1 = tile_br[t + 1];
    const int nbr = br1 - br0;
    const int b0 = bptr[br0], b1 = bptr[br1];
    int v = tile_val[t];
    const int64_t row0 = (int64_t)br0 * R;
    const int64_t rows_left = nrows - row0;
    const int nrow_t = (int)((int64_t)nbr * R < rows_left ? (int64_t)nbr * R : rows_left);

    for (int i = lane; i <= nbr; i += 64) rs[i] = bptr[br0 + i] - b0;
    for (int i = lane; i < nbr * R * kt; i += 64) acc[i] = (ACC)0;
    int top = 1;
    while (top < nbr) top <<= 1;

    int b = b0;
    int cs = 0;
    unsigned long long mw[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) mw[w] = 0ull;
    if (b0 < b1) load_header_clamped<R, MB>(hdr, b + lane, b1 - 1, cs, mw);
    wave_sync();
    while (b < b1) {
        const bool valid = b + lane < b1;
        unsigned long long mw_cur[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) mw_cur[w] = valid ? mw[w] : 0ull;
        int cs_cur = valid ? cs : 0;
        asm volatile("" : "+v"(cs_cur));
        int cnt = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) cnt += __popcll(mw_cur[w]);
        const int incl = wave_incl_scan(cnt);
        const int ntake = __popcll(ballot64(valid && incl <= CAP));
        const int total = __builtin_amdgcn_readlane(incl, ntake - 1);
        ACC a[K];
        load_values<V, ACC, K>(val, v, total, lane, a);
        const int jcur = b + lane;
        const int bn = b + ntake;
        load_header_clamped<R, MB>(hdr, bn + lane, b1 - 1, cs, mw);
        if (lane < ntake) {
            // records (X row, tile-local row) of this block's set bits
            const int rbase = find_blockrow(rs, nbr, top, jcur - b0) * R;
            const unsigned cbase = WIDE ? (unsigned)cs_cur : (unsigned)cs_cur * row_bytes;
            int p = incl - cnt;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                unsigned long long m = mw_cur[w];
                while (m) {
                    const int bit = __builtin_ctzll(m) + w * 64;
                    const unsigned c = (unsigned)(bit & (VS - 1));
                    rec[p].col = (int)(WIDE ? cbase + c : cbase + __umul24(c, row_bytes));
                    rec[p].row = rbase + bit / VS;
                    ++p;
                    m &= m - 1ull;
                }
            }
        }
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const int q = lane + 64 * kk;
            if (q < total) rec[q].v = a[kk];
        }
        wave_sync();
        // group grp walks records [qa, qb)
        const int per = (total + G - 1) / G;
        int q = grp * per;
        const int qb = q + per < total ? q + per : total;
        ACC run[CPL];
#pragma unroll
        for (int e = 0; e < CPL; ++e) run[e] = (ACC)0;
        int prow = -1;
        // the running sums of row prow go to the tile's LDS sums
        auto close = [&]() {
            if (prow >= 0 && colok) {
#pragma unroll
                for (int e = 0; e < CPL; ++e) atomicAdd(&acc[prow * kt + jl + e], run[e]);
            }
#pragma unroll
            for (int e = 0; e < CPL; ++e) run[e] = (ACC)0;
        };
        if constexpr (KC == 64 && CPL == 1 && !WIDE && sizeof(X) < 8) {
            // One group: every record is the same on all lanes.  Lane u reads
            // record q + u (one LDS read for U records) and the fields are
            // broadcast to scalars, so the records take no vector registers
            // and each X row request is a scalar base plus the lane's offset.
            static_assert(U <= 64 && (U & (U - 1)) == 0, "U: a power of two up to 64");
            for (; q + U <= qb; q += U) {
                const Rec mine = rec[q + (lane & (U - 1))];
                unsigned cu[U];
                int ru[U];
                ACC vu[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    cu[u] = (unsigned)__builtin_amdgcn_readlane(mine.col, u);
                    ru[u] = __builtin_amdgcn_readlane(mine.row, u);
                    vu[u] = read_lane(mine.v, u);
                }
                ACC xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    xv[u] = (ACC) * reinterpret_cast<const X*>(reinterpret_cast<const char*>(x) + cu[u] + lane_off);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (ru[u] != prow) {
                        close();
                        prow = ru[u];
                    }
                    run[0] += vu[u] * xv[u];
                }
            }
        } else {
            for (; q + U <= qb; q += U) {
                Rec rr[U];
#pragma unroll
                for (int u = 0; u < U; ++u) rr[u] = rec[q + u];
                // All U record reads are issued before the first use waits.
                __builtin_amdgcn_sched_barrier(0);
                ACC xv[U][CPL];
#pragma unroll
                for (int u = 0; u < U; ++u) spmm_xs<CPL, WIDE>(xcol, x, rr[u].col, ldx, lane_off, xv[u]);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (rr[u].row != prow) {
                        close();
                        prow = rr[u].row;
                    }
#pragma unroll
                    for (int e = 0; e < CPL; ++e) run[e] += rr[u].v * xv[u][e];
                }
            }
        }
        for (; q < qb; ++q) {
            const Rec r1 = rec[q];
            ACC xv[CPL];
            spmm_xs<CPL, WIDE>(xcol, x, r1.col, ldx, lane_off, xv);
            if (r1.row != prow) {
                close();
                prow = r1.row;
            }
#pragma unroll
            for (int e = 0; e < CPL; ++e) run[e] += r1.v * xv[e];
        }
        close();
        v += total;
        b = bn;
        wave_sync();
    }
    wave_sync();
    for (int i = lane; i < nrow_t * ktc; i += 64) {
        const int rr = i / ktc, j = i % ktc;
        X* yp = &y[(row0 + rr) * ldy + c0 + j];
        const ACC s = acc[rr * kt + j];
        *yp = beta == (ACC)0 ? (X)(alpha * s) : (X)(alpha * s + beta * (ACC)*yp);
    }
}

// ---- column records ------------------------------------------------------
//
// In a beta(r,VS) block the r rows share one window of VS columns, and one
// column is usually set in several rows (r = 4 on the 27-point stencil: 12
// values over 6 columns).  This kernel makes one record per (block, set
// column): the X row offset, the block-row and the r values of that column
// (0 where a row has none).  Each record requests its X row once and adds
// v[t] * x into r running sums, one per row of the block-row, written to
// the tile's LDS sums when the block-row changes.  X is then requested
// once per distinct column of a block instead of once per value.

template <class ACC, int R>
struct alignas(16) SpmmColRec {
    int col;   // X row: byte offset from X, or the column (WIDE)
    int brow;  // tile-local block-row
    ACC v[R];  // the column's value in each row of the block (0 where none)
};

// Records and values per batch (LDS about 4-5 KiB of records per wave).
template <int R>
constexpr int spmm_col_cap_rec() { return R == 2 ? 128 : (R == 4 ? 96 : 64); }
template <int R>
constexpr int spmm_col_cap_val() { return R == 2 ? 256 : (R == 4 ? 384 : 512); }

// With 64-column chunks of one column per lane, one wave walks the whole
// record sequence of its tile in block-row order, so it writes each finished
// block-row of Y directly (no LDS row x column sums, tiles of up to kMaxRows
// rows).
template <int KC, int CPL>
constexpr bool spmm_col_direct() { return KC == 64 && CPL == 1; }


template <class ACC, int R>
size_t spmm_col_smem_bytes(int maxrows, int kt, bool direct) {
    return sizeof(SpmmColRec<ACC, R>) * spmm_col_cap_rec<R>() + (direct ? 0 : sizeof(ACC) * (size_t)maxrows * kt) +
           sizeof(int) * (maxrows / R + 1) + sizeof(unsigned short) * spmm_col_cap_val<R>();
}

// Columns set in any row of a block (the OR of its r row masks).
template <int R, int VS, int NW>
__device__ __forceinline__ unsigned block_columns(const unsigned long long* mw) {
    unsigned cm = 0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const int bit0 = t * VS;
        cm |= (unsigned)((mw[bit0 / 64] >> (bit0 % 64)) & ((1ull << VS) - 1ull));
    }
    return cm;
}

// KC: the chunk's columns rounded up to a power of two; CPL: adjacent
// columns per lane (1, or 16 / sizeof(X)), so KC / CPL lanes share a record.
template <class V, class X, class ACC, int R, int VS, int MB, int KC, bool WIDE, int CPL>
__global__ __launch_bounds__(64) void k_spmm_cols(const int* __restrict__ bptr, const unsigned char* __restrict__ hdr,
                                                   const V* __restrict__ val, const X* __restrict__ x, int64_t ldx,
                                                   X* __restrict__ y, int64_t ldy, int k, int kt,
                                                   const int* __restrict__ tile_br, const int* __restrict__ tile_val,
                                                   int64_t ntiles, int64_t nrows, int maxrows, ACC alpha, ACC beta) {
    constexpr int CAPR = spmm_col_cap_rec<R>();
    constexpr int CAP = spmm_col_cap_val<R>();
    constexpr int K = CAP / 64;
    constexpr int NW = MaskWords<R * MB>::kWords;
    constexpr int LPR = KC / CPL;        // lanes per record
    constexpr int G = 64 / LPR;          // record groups per wave
    // X rows requested together.  With several columns per lane and more
    // than two groups per wave, half as many: k = 64 in double on the
    // 27-point stencil at beta(4,8) took 3.16 ms with 32-column chunks and 4
    // requests (3.53 ms with 8), and 2.95 ms with 64-column chunks (two
    // groups) and 8 requests (3.12 ms with 4).
    constexpr int U = (R <= 2 ? 16 : 8) / (CPL > 1 && G > 2 ? 2 : 1);
    constexpr bool DIRECT = spmm_col_direct<KC, CPL>();
    using Rec = SpmmColRec<ACC, R>;
    static_assert(sizeof(Rec) % 16 == 0, "16-byte aligned records");
    extern __shared__ __align__(16) unsigned char smem[];
    Rec* rec = reinterpret_cast<Rec*>(smem);                                          // [CAPR]
    ACC* acc = reinterpret_cast<ACC*>(rec + CAPR);                                    // [maxrows * kt] unless DIRECT
    int* rs = reinterpret_cast<int*>(acc + (DIRECT ? 0 : (size_t)maxrows * kt));      // [maxrows / R + 1]
    unsigned short* slot = reinterpret_cast<unsigned short*>(rs + (maxrows / R + 1));  // [CAP]

    int64_t t;
    int c0;
    if (!spmm_item(ntiles, k, kt, t, c0)) return;
    const int ktc = (k - c0) < kt ? (k - c0) : kt;
    const int lane = lane_id();
    const int grp = lane / LPR;
    const int jl = (lane % LPR) * CPL;  // the lane's first column in the chunk
    const bool colok = jl < ktc;        // ktc is a multiple of CPL
    const X* __restrict__ xcol = x + c0 + (colok ? jl : 0);
    const unsigned lane_off = (unsigned)(c0 + (colok ? jl : 0)) * (unsigned)sizeof(X);
    const unsigned row_bytes = (unsigned)ldx * (unsigned)sizeof(X);
    const int br0 = tile_br[t], br1 = tile_br[t + 1];
    const int nbr = br1 - br0;
    const int b0 = bptr[br0], b1 = bptr[br1];
    int v = tile_val[t];
    const int64_t row0 = (int64_t)br0 * R;
    const int64_t rows_left = nrows - row0;
    const int nrow_t = (int)((int64_t)nbr * R < rows_left ? (int64_t)nbr * R : rows_left);

    for (int i = lane; i <= nbr; i += 64) rs[i] = bptr[br0 + i] - b0;
    if constexpr (!DIRECT) {
        for (int i = lane; i < nbr * R * kt; i += 64) acc[i] = (ACC)0;
    }
    int top = 1;
    while (top < nbr) top <<= 1;

    // Running sums of the current block-row, one per row of it; with DIRECT
    // they persist across batches (the wave walks the tile's records in
    // block-row order) and finished block-rows go straight to Y.
    ACC run[R][CPL];
#pragma unroll
    for (int tr = 0; tr < R; ++tr)
#pragma unroll
        for (int e = 0; e < CPL; ++e) run[tr][e] = (ACC)0;
    int pb = -1;    // block-row the running sums belong to
    int done = 0;   // DIRECT: block-rows [0, done) of the tile are written
    // DIRECT: Y of block-row pb, requested when the block-row opens so that
    // the read is in flight while its records are processed
    X yv[DIRECT ? R : 1];
    auto request_y = [&](int brow) {
        if constexpr (DIRECT) {
#pragma unroll
            for (int tr = 0; tr < R; ++tr) {
                const int rr = brow * R + tr;
                yv[tr] = (beta != (ACC)0 && brow >= 0 && rr < nrow_t && colok) ? y[(row0 + rr) * ldy + c0 + jl]
                                                                                 : (X)0;
            }
        }
    };
    auto write_blockrow = [&](int brow, bool have) {
#pragma unroll
        for (int tr = 0; tr < R; ++tr) {
            const int rr = brow * R + tr;
            if (rr < nrow_t && colok) {
                X* yp = &y[(row0 + rr) * ldy + c0 + jl];
                const ACC s = have ? run[tr][0] : (ACC)0;
                const ACC yo = beta == (ACC)0 ? (ACC)0 : (have ? (ACC)yv[tr] : (ACC)*yp);
                *yp = (X)(alpha * s + beta * yo);
            }
        }
    };
    // block-rows [done, upto) hold no value: y = beta * y (nothing when beta == 1)
    auto skip_to = [&](int upto) {
        if (beta != (ACC)1) {
            for (int br = done; br < upto; ++br) write_blockrow(br, false);
        }
        done = upto > done ? upto : done;
    };
    auto flush = [&](int brow) {
        if (pb >= 0) {
            if constexpr (DIRECT) {
                skip_to(pb);
                write_blockrow(pb, true);
                done = pb + 1;
            } else if (colok) {
#pragma unroll
                for (int tr = 0; tr < R; ++tr)
#pragma unroll
                    for (int e = 0; e < CPL; ++e) atomicAdd(&acc[(pb * R + tr) * kt + jl + e], run[tr][e]);
            }
        }
#pragma unroll
        for (int tr = 0; tr < R; ++tr)
#pragma unroll
            for (int e = 0; e < CPL; ++e) run[tr][e] = (ACC)0;
        pb = brow;
        request_y(brow);
    };

    int b = b0;
    int cs = 0;
    unsigned long long mw[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) mw[w] = 0ull;
    if (b0 < b1) load_header_clamped<R, MB>(hdr, b + lane, b1 - 1, cs, mw);
    wave_sync();
    while (b < b1) {
        const bool valid = b + lane < b1;
        unsigned long long mw_cur[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) mw_cur[w] = valid ? mw[w] : 0ull;
        int cs_cur = valid ? cs : 0;
        asm volatile("" : "+v"(cs_cur));
        int cnt = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) cnt += __popcll(mw_cur[w]);
        const unsigned cm = block_columns<R, VS, NW>(mw_cur);
        const int crec = __popc(cm);
        const int incl = wave_incl_scan(cnt);
        const int inclr = wave_incl_scan(crec);
        const int ntake = __popcll(ballot64(valid && incl <= CAP && inclr <= CAPR));
        const int total = __builtin_amdgcn_readlane(incl, ntake - 1);
        const int nrec = __builtin_amdgcn_readlane(inclr, ntake - 1);
        ACC a[K];
        load_values<V, ACC, K>(val, v, total, lane, a);
        const int jcur = b + lane;
        const int bn = b + ntake;
        load_header_clamped<R, MB>(hdr, bn + lane, b1 - 1, cs, mw);
        // records of this batch start at zero (every lane clears 16-byte words)
        {
            uint4* words = reinterpret_cast<uint4*>(rec);
            const int nw = nrec * (int)(sizeof(Rec) / 16);
            for (int i = lane; i < nw; i += 64) words[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        wave_sync();
        if (lane < ntake) {
            // record header per set column, and the record slot of every value
            const int brow = find_blockrow(rs, nbr, top, jcur - b0);
            const unsigned cbase = WIDE ? (unsigned)cs_cur : (unsigned)cs_cur * row_bytes;
            const int rb = inclr - crec;
            unsigned m = cm;
            int i = rb;
            while (m) {
                const unsigned c = (unsigned)__builtin_ctz(m);
                rec[i].col = (int)(WIDE ? cbase + c : cbase + __umul24(c, row_bytes));
                rec[i].brow = brow;
                ++i;
                m &= m - 1u;
            }
            int p = incl - cnt;
#pragma unroll
            for (int tr = 0; tr < R; ++tr) {
                const int bit0 = tr * VS;
                unsigned mr = (unsigned)((mw_cur[bit0 / 64] >> (bit0 % 64)) & ((1ull << VS) - 1ull));
                while (mr) {
                    const unsigned c = (unsigned)__builtin_ctz(mr);
                    const int ri = rb + __popc(cm & ((1u << c) - 1u));  // record of column c
                    slot[p] = (unsigned short)(ri * (int)sizeof(Rec) + 8 + tr * (int)sizeof(ACC));
                    ++p;
                    mr &= mr - 1u;
                }
            }
        }
        wave_sync();
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const int q = lane + 64 * kk;
            if (q < total) *reinterpret_cast<ACC*>(reinterpret_cast<unsigned char*>(rec) + slot[q]) = a[kk];
        }
        wave_sync();
        // group grp walks records [q, qb)
        const int per = (nrec + G - 1) / G;
        int q = grp * per;
        const int qb = q + per < nrec ? q + per : nrec;
        for (; q + U <= qb; q += U) {
            int cu[U], bu[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int2 hb = *reinterpret_cast<const int2*>(&rec[q + u]);
                cu[u] = hb.x;
                bu[u] = hb.y;
            }
            // all U X rows are requested before the first use waits
            __builtin_amdgcn_sched_barrier(0);
            ACC xv[U][CPL];
#pragma unroll
            for (int u = 0; u < U; ++u) spmm_xs<CPL, WIDE>(xcol, x, cu[u], ldx, lane_off, xv[u]);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (bu[u] != pb) flush(bu[u]);
#pragma unroll
                for (int tr = 0; tr < R; ++tr)
#pragma unroll
                    for (int e = 0; e < CPL; ++e) run[tr][e] += rec[q + u].v[tr] * xv[u][e];
            }
        }
        for (; q < qb; ++q) {
            const int2 hb = *reinterpret_cast<const int2*>(&rec[q]);
            ACC xv[CPL];
            spmm_xs<CPL, WIDE>(xcol, x, hb.x, ldx, lane_off, xv);
            if (hb.y != pb) flush(hb.y);
#pragma unroll
            for (int tr = 0; tr < R; ++tr)
#pragma unroll
                for (int e = 0; e < CPL; ++e) run[tr][e] += rec[q].v[tr] * xv[e];
        }
        // groups end their runs with the batch (each batch splits anew);
        // one group keeps its block-row open across batches
        if constexpr (!DIRECT) flush(-1);
        v += total;
        b = bn;
        wave_sync();
    }
    if constexpr (DIRECT) {
        flush(-1);
        skip_to(nbr);
    } else {
        wave_sync();
        for (int i = lane; i < nrow_t * ktc; i += 64) {
            const int rr = i / ktc, j = i % ktc;
            X* yp = &y[(row0 + rr) * ldy + c0 + j];
            const ACC s = acc[rr * kt + j];
            *yp = beta == (ACC)0 ? (X)(alpha * s) : (X)(alpha * s + beta * (ACC)*yp);
        }
    }
}

struct SpmmPlans {
    std::mutex mu;
    std::map<const spc5_hip_mat*, std::map<int, Plan>> plans;
};
SpmmPlans& spmm_plans() {
    static SpmmPlans p;
    return p;
}

// Groups for ntiles tiles times the column chunks, a multiple of 8 (the
// caller has checked that it fits 32 bits).
unsigned spmm_grid(int64_t ntiles, int k, int kt) {
    const uint64_t items = (uint64_t)ntiles * (uint64_t)((k + kt - 1) / kt);
    return (unsigned)((items + 7) / 8 * 8);
}

template <class V, class X, class ACC, int VS, int MB, int KC, bool WIDE, int CPL>
int launch_spmm_cols(const spc5_hip_mat& m, const Plan& p, int maxrows, const void* x, int64_t ldx, void* y,
                     int64_t ldy, int k, int kt, double alpha, double beta, hipStream_t s) {
    const dim3 grid(spmm_grid(p.ntiles, k, kt));
    const int* tb = p.tile_br->as<int>();
    const int* tv = p.tile_val->as<int>();
    const V* val = reinterpret_cast<const V*>(m.values);
    const X* xx = reinterpret_cast<const X*>(x);
    X* yy = reinterpret_cast<X*>(y);
#define SPC5_SPMM_COLS(R)                                                                                          \
    {                                                                                                              \
        const size_t shm = spmm_col_smem_bytes<ACC, R>(maxrows, kt, spmm_col_direct<KC, CPL>());                  \
        hipLaunchKernelGGL((k_spmm_cols<V, X, ACC, R, VS, MB, KC, WIDE, CPL>), grid, dim3(64), shm, s, m.bptr, m.hdr, \
                           val,                                                                                    \
                           xx, ldx, yy, ldy, k, kt, tb, tv, p.ntiles, m.nrows, maxrows, (ACC)alpha, (ACC)beta);       \
    }
    switch (m.r) {
    case 2: SPC5_SPMM_COLS(2); break;
    case 4: SPC5_SPMM_COLS(4); break;
    case 8: SPC5_SPMM_COLS(8); break;
    default: set_error("column records need r = 2, 4 or 8"); return SPC5_HIP_ERR_INVALID;
    }
#undef SPC5_SPMM_COLS
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

template <class V, class X, class ACC, int VS, int MB, int KC, bool WIDE, int CPL>
int launch_spmm_vals(const spc5_hip_mat& m, const Plan& p, int maxrows, const void* x, int64_t ldx, void* y,
                     int64_t ldy, int k, int kt, double alpha, double beta, hipStream_t s) {
    const dim3 grid(spmm_grid(p.ntiles, k, kt));
    const size_t shm = spmm_smem_bytes<ACC, KC, X, CPL>(maxrows, kt, m.r);
    const int* tb = p.tile_br->as<int>();
    const int* tv = p.tile_val->as<int>();
    const V* val = reinterpret_cast<const V*>(m.values);
    const X* xx = reinterpret_cast<const X*>(x);
    X* yy = reinterpret_cast<X*>(y);
#define SPC5_SPMM_LAUNCH(R)                                                                                   \
    hipLaunchKernelGGL((k_spmm_tiles<V, X, ACC, R, VS, MB, KC, WIDE, CPL>), grid, dim3(64), shm, s, m.bptr, m.hdr, \
                       val, xx, ldx, yy, ldy, k, kt, tb, tv, p.ntiles, m.nrows, maxrows, (ACC)alpha, (ACC)beta)
    switch (m.r) {
    case 1: SPC5_SPMM_LAUNCH(1); break;
    case 2: SPC5_SPMM_LAUNCH(2); break;
    case 4: SPC5_SPMM_LAUNCH(4); break;
    default: SPC5_SPMM_LAUNCH(8); break;
    }
#undef SPC5_SPMM_LAUNCH
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

template <class V, class X, class ACC, int VS, int MB, int KC, bool WIDE>
int launch_spmm_kc(const spc5_hip_mat& m, const Plan& p, int maxrows, const void* x, int64_t ldx, void* y,
                   int64_t ldy, int k, int kt, double alpha, double beta, bool vec, hipStream_t s) {
    constexpr int CPL = 16 / (int)sizeof(X);
    const bool cols = m.spmm_cols && m.r > 1;
    if constexpr (KC >= CPL) {
        if (vec)
            return cols ? launch_spmm_cols<V, X, ACC, VS, MB, KC, WIDE, CPL>(m, p, maxrows, x, ldx, y, ldy, k, kt,
                                                                            alpha, beta, s)
                        : launch_spmm_vals<V, X, ACC, VS, MB, KC, WIDE, CPL>(m, p, maxrows, x, ldx, y, ldy, k, kt,
                                                                            alpha, beta, s);
    }
    return cols ? launch_spmm_cols<V, X, ACC, VS, MB, KC, WIDE, 1>(m, p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s)
                : launch_spmm_vals<V, X, ACC, VS, MB, KC, WIDE, 1>(m, p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
}

template <class V, class X, class ACC, int VS, int MB>
int launch_spmm_t(const spc5_hip_mat& m, const void* x, int64_t ldx, void* y, int64_t ldy, int k, int kt,
                  double alpha, double beta, hipStream_t s) {
    int kc = 1;  // lanes per record: the smallest power of two holding the chunk's kt columns
    while (kc < kt) kc <<= 1;
    // With SPMM_X16, records take 16 / sizeof(X) adjacent columns per lane
    // (one 16-byte X load) when every chunk's width is a multiple of that
    // and every X row and chunk start is 16-byte aligned.  A record's X
    // request then takes a half (double) or a quarter (float) of the lanes,
    // and the address path and the LDS record reads cost that much less per
    // column: k = 64 in double on the 27-point stencil (160^3, beta(4,8),
    // column records) went from 3.65 to 2.95 ms.  For float X it wins or
    // loses by r and chunk (column records with 64-column chunks: 2.37 ->
    // 2.74 ms at beta(4,8), 2.28 ms at beta(2,8)), so it is off by default
    // there.
    const int cpl = 16 / (int)sizeof(X);
    const bool x16 = m.spmm_x16 < 0 ? sizeof(X) == 8 : m.spmm_x16 == 1;
    const bool vec = x16 && kc >= cpl && k % cpl == 0 && kt % cpl == 0 && ldx % cpl == 0 &&
                     ((uintptr_t)x & 15u) == 0;
    // column records with 64-column chunks of one column per lane write Y
    // directly: whole SpMV-sized tiles
    const bool direct = m.spmm_cols && m.r > 1 && kc == 64 && !vec;
    const int maxrows = direct ? kMaxRows : spmm_max_rows<ACC>(kt);
    Plan* p = nullptr;
    {
        SpmmPlans& sp = spmm_plans();
        std::lock_guard<std::mutex> lk(sp.mu);
        auto& per = sp.plans[&m];
        auto it = per.find(maxrows);
        if (it == per.end()) {
            Plan np;
            int64_t target = m.nnz / 16384;
            if (target < 256) target = 256;
            if (target > 4096) target = 4096;
            const int rc = build_tiles(m, np, maxrows, (int)target, s);
            if (rc) return rc;
            it = per.emplace(maxrows, np).first;
        }
        p = &it->second;
    }
    if (p->ntiles == 0) return SPC5_HIP_OK;
    if ((uint64_t)p->ntiles * (uint64_t)((k + kt - 1) / kt) > 0xfffffff8ull) {
        set_error("spmm: tiles times column chunks exceed 2^32 - 8 workgroups");
        return SPC5_HIP_ERR_INVALID;
    }
    // 32-bit X offsets when every X row offset plus a column fits, and a
    // row is under 16 MiB (the 24-bit multiply of the expansion)
    const uint64_t row_bytes = (uint64_t)ldx * sizeof(X);
    const bool wide = row_bytes >= (1ull << 24) || (uint64_t)m.ncols * row_bytes > 0xffffffffull;
#define SPC5_SPMM_KC(KC)                                                                                         \
    return wide ? launch_spmm_kc<V, X, ACC, VS, MB, KC, true>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, \
                                                               vec, s)                                            \
                : launch_spmm_kc<V, X, ACC, VS, MB, KC, false>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta,  \
                                                                vec, s)
    switch (kc) {
    case 64: SPC5_SPMM_KC(64);
    case 32: SPC5_SPMM_KC(32);
    case 16: SPC5_SPMM_KC(16);
    case 8: SPC5_SPMM_KC(8);
    case 4: SPC5_SPMM_KC(4);
    case 2: SPC5_SPMM_KC(2);
    default: SPC5_SPMM_KC(1);
    }
#undef SPC5_SPMM_KC
}

}  // namespace

void forget_spmm_plans(const spc5_hip_mat* m) {
    SpmmPlans& sp = spmm_plans();
    std::lock_guard<std::mutex> lk(sp.mu);
    sp.plans.erase(m);
}

}  // namespace spc5

extern "C" {

int spc5_hip_spmm(spc5_hip_mat_t h, int k, int kt, double alpha, const void* d_x, int64_t ldx, double beta, void* d_y,
                  int64_t ldy, void* stream) {
    SPC5_CHECK_ARG(h != nullptr, "null handle");
    SPC5_CHECK_ARG(k >= 1 && ldx >= k && ldy >= k, "need k >= 1, ldx >= k, ldy >= k");
    SPC5_CHECK_ARG(d_x != nullptr && d_y != nullptr, "x or y is null");
    if (kt <= 0 || kt > k) kt = k;
    if (kt > kSpmmMaxKt) kt = kSpmmMaxKt;
    hipStream_t s = as_stream(stream);
    if (k == 1 && ldx == 1 && ldy == 1) {
        // One contiguous column: the single-vector product.
        Epilogue ep;
        ep.alpha = alpha;
        ep.beta = beta;
        return launch_spmv(*h, d_x, d_y, s, ep);
    }
    if (h->value_type == SPC5_HIP_F64)
        return launch_spmm_t<double, double, double, 8, 1>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
    if (h->value_type == SPC5_HIP_F32 && h->vs == 16)
        return launch_spmm_t<float, float, SpmmFloatAcc, 16, 2>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
    if (h->value_type == SPC5_HIP_F32)
        return launch_spmm_t<float, float, SpmmFloatAcc, 8, 1>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
    return launch_spmm_t<unsigned short, float, SpmmFloatAcc, 8, 1>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
}

}  // extern "C"
