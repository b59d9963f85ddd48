// spmm.hip -- multi-right-hand-side product Y = alpha A X + beta Y for a
// beta(r,VS) matrix and row-major dense blocks X (ncols x k, leading
// dimension ldx) and Y (nrows x k, ldy).  The batching form; no counterpart
// in the reference.
//
// One wave per (tile, column chunk).  The tile's blocks are decoded and
// expanded exactly as in the SpMV tile (headers prefetched, values read
// once, coalesced, from the value stream) and every value is then applied
// to the kt columns of its X row: lanes are spread over (value, column)
// pairs, KC lanes per value, so a chunk of kt <= 64 columns of one X row is
// read contiguously.  With kt = k the matrix is fetched once for all k
// columns; smaller kt re-reads the matrix k/kt times but keeps the x window
// of each pass small enough to stay in L2 (the autotuner picks kt).
// Per-row, per-column sums live in LDS (tiles are cut to fit them).
#include "spc5_mat.hpp"
#include "spmv_tile.hpp"

#include <map>
#include <mutex>

using namespace spc5;

namespace spc5 {
namespace {

constexpr int kSpmmCap = 256;        // values expanded per batch
constexpr int kSpmmAccBytes = 16384;  // LDS for the row x column sums of a tile

template <class ACC>
int spmm_max_rows(int kt) {
    int rows = kSpmmAccBytes / (kt * (int)sizeof(ACC));
    if (rows > 256) rows = 256;
    if (rows < 8) rows = 8;
    return rows;
}

template <class V, class X, class ACC, int R, int VS, int MB, int KC>
__global__ __launch_bounds__(64) void k_spmm_tiles(const int* __restrict__ bptr, const unsigned char* __restrict__ hdr,
                                                    const V* __restrict__ val, const X* __restrict__ x, int64_t ldx,
                                                    X* __restrict__ y, int64_t ldy, int k, int kt,
                                                    const int* __restrict__ tile_br, const int* __restrict__ tile_val,
                                                    int64_t ntiles, int64_t nrows, int maxrows, ACC alpha, ACC beta) {
    constexpr int K = kSpmmCap / 64;
    constexpr int NW = MaskWords<R * MB>::kWords;
    constexpr int VPL = 64 / KC;  // values per step
    extern __shared__ __align__(16) unsigned char smem[];
    ACC* acc = reinterpret_cast<ACC*>(smem);                          // [maxrows * kt]
    ACC* sval = acc + (size_t)maxrows * kt;                           // [kSpmmCap]
    int* scol = reinterpret_cast<int*>(sval + kSpmmCap);              // [kSpmmCap]
    int* rs = scol + kSpmmCap;                                        // [maxrows / R + 1]
    unsigned char* srow = reinterpret_cast<unsigned char*>(rs + maxrows / R + 1);  // [kSpmmCap]

    const int64_t t = blockIdx.x;
    if (t >= ntiles) return;
    const int c0 = blockIdx.y * kt;
    const int ktc = (k - c0) < kt ? (k - c0) : kt;
    const int lane = threadIdx.x;
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
    int cs;
    unsigned long long mw[NW];
    load_header<R, MB>(hdr, b + lane, b1, cs, mw);
    wave_sync();
    const int sub = lane / KC;  // value slot of this lane inside a step
    const int jl = lane % KC;   // column of this lane inside a pass
    while (b < b1) {
        // shape of the batch (at most kSpmmCap values)
        int cnt = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) cnt += __popcll(mw[w]);
        const int incl = wave_incl_scan(cnt);
        const int ntake = __popcll(ballot64(b + lane < b1 && incl <= kSpmmCap));
        const int total = __builtin_amdgcn_readlane(incl, ntake - 1);
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const int q = lane + 64 * kk;
            if (q < total) sval[q] = ValueLoad<V, ACC>::get(val + v + q);
        }
        const int jcur = b + lane;
        const int cs_cur = cs;
        unsigned long long mw_cur[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) mw_cur[w] = mw[w];
        const int bn = b + ntake;
        load_header<R, MB>(hdr, bn + lane, b1, cs, mw);
        if (lane < ntake) expand<R, VS, NW>(scol, srow, rs, nbr, top, jcur - b0, cs_cur, mw_cur, incl - cnt);
        wave_sync();
        // apply every value of the batch to its X row chunk
        for (int q0 = 0; q0 < total; q0 += VPL * 4) {
            int qs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) qs[u] = q0 + u * VPL + sub;
            for (int j = jl; j < ktc; j += KC) {
                ACC xx[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    xx[u] = qs[u] < total ? (ACC)x[(int64_t)scol[qs[u]] * ldx + c0 + j] : (ACC)0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (qs[u] < total) {
                        ACC* a = &acc[(int)srow[qs[u]] * kt + j];
                        if (VPL == 1)
                            *a += sval[qs[u]] * xx[u];
                        else
                            atomicAdd(a, sval[qs[u]] * xx[u]);
                    }
                }
            }
        }
        v += total;
        b = bn;
        wave_sync();
    }
    for (int i = lane; i < nrow_t * ktc; i += 64) {
        const int rr = i / ktc, j = i % ktc;
        X* yp = &y[(row0 + rr) * ldy + c0 + j];
        const ACC s = acc[rr * kt + j];
        *yp = beta == (ACC)0 ? (X)(alpha * s) : (X)(alpha * s + beta * (ACC)*yp);
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

template <class V, class X, class ACC, int VS, int MB, int KC>
int launch_spmm_kc(const spc5_hip_mat& m, const Plan& p, int maxrows, const void* x, int64_t ldx, void* y,
                   int64_t ldy, int k, int kt, double alpha, double beta, hipStream_t s) {
    const dim3 grid((unsigned)p.ntiles, (unsigned)((k + kt - 1) / kt));
    const size_t shm = sizeof(ACC) * ((size_t)maxrows * kt + kSpmmCap) + sizeof(int) * (kSpmmCap + maxrows + 1) +
                       kSpmmCap + 16;
    const int* tb = p.tile_br->as<int>();
    const int* tv = p.tile_val->as<int>();
    const V* val = reinterpret_cast<const V*>(m.values);
    const X* xx = reinterpret_cast<const X*>(x);
    X* yy = reinterpret_cast<X*>(y);
#define SPC5_SPMM_LAUNCH(R)                                                                                   \
    hipLaunchKernelGGL((k_spmm_tiles<V, X, ACC, R, VS, MB, KC>), grid, dim3(64), shm, s, m.bptr, m.hdr, val, \
                       xx, ldx, yy, ldy, k, kt, tb, tv, p.ntiles, m.nrows, maxrows, (ACC)alpha, (ACC)beta)
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

template <class V, class X, class ACC, int VS, int MB>
int launch_spmm_t(const spc5_hip_mat& m, const void* x, int64_t ldx, void* y, int64_t ldy, int k, int kt,
                  double alpha, double beta, hipStream_t s) {
    const int maxrows = spmm_max_rows<ACC>(kt);
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
    const int kc = kt >= 64 ? 64 : (kt >= 32 ? 32 : (kt >= 16 ? 16 : (kt >= 8 ? 8 : (kt >= 4 ? 4 : (kt >= 2 ? 2 : 1)))));
    switch (kc) {
    case 64: return launch_spmm_kc<V, X, ACC, VS, MB, 64>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    case 32: return launch_spmm_kc<V, X, ACC, VS, MB, 32>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    case 16: return launch_spmm_kc<V, X, ACC, VS, MB, 16>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    case 8: return launch_spmm_kc<V, X, ACC, VS, MB, 8>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    case 4: return launch_spmm_kc<V, X, ACC, VS, MB, 4>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    case 2: return launch_spmm_kc<V, X, ACC, VS, MB, 2>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    default: return launch_spmm_kc<V, X, ACC, VS, MB, 1>(m, *p, maxrows, x, ldx, y, ldy, k, kt, alpha, beta, s);
    }
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
    hipStream_t s = as_stream(stream);
    if (h->value_type == SPC5_HIP_F64)
        return launch_spmm_t<double, double, double, 8, 1>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
    if (h->value_type == SPC5_HIP_F32 && h->vs == 16)
        return launch_spmm_t<float, float, float, 16, 2>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
    if (h->value_type == SPC5_HIP_F32)
        return launch_spmm_t<float, float, float, 8, 1>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
    return launch_spmm_t<unsigned short, float, float, 8, 1>(*h, d_x, ldx, d_y, ldy, k, kt, alpha, beta, s);
}

}  // extern "C"
