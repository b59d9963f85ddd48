// spmv.hip -- beta(r,VS) SpMV on gfx950: y += A x with A in the reference
// block layout, plus the plan builder and the handle API.
//
// Kernel design (one wave = one tile of whole block-rows):
//   * the tile's block-row pointers go to LDS; a running value offset walks
//     the value stream exactly like the reference's `values += popcount`
//     (src/spc5.hpp:2571-2617), starting from the tile's planned offset;
//   * batch step: lane i decodes block b+i (unaligned int32 column + r
//     masks), counts its values (popcount), and a DPP wave scan gives each
//     block's position in the batch; blocks are taken while the batch holds
//     at most expand_cap<V>() values;
//   * each taken lane finds its block-row by binary search in LDS and
//     expands its set bits into (column, local row) pairs in LDS -- values
//     are never zero-padded;
//   * the batch's values are then read with fully coalesced loads, one per
//     lane, multiplied by x[column] and added into per-row LDS accumulators;
//   * at the end of the tile every row does one y[row] += sum (rows of a
//     tile belong to it alone, so no global atomics and no second pass).
#include "spc5_mat.hpp"
#include "spmv_tile.hpp"

#include <mutex>
#include <vector>

namespace spc5 {

namespace {

std::mutex g_err_mu;
std::string g_err;

// Workgroups are dealt to the 8 XCDs round-robin (group g runs on XCD g % 8).
// Renumbering groups so that each XCD walks one contiguous eighth of the
// tile list keeps the x window each XCD gathers from inside its own L2.
__device__ __forceinline__ int64_t xcd_tile(int wave) {
    const int64_t g = blockIdx.x;
#if SPC5_XCD_REMAP
    const int64_t per_xcd = gridDim.x >> 3;  // the grid is a multiple of 8
    const int64_t gx = (g & 7) * per_xcd + (g >> 3);
#else
    const int64_t gx = g;
#endif
    return gx * kWavesPerGroup + wave;
}


template <class V, class X, class ACC, int R, int VS, int MB, bool SORT>
__global__ __launch_bounds__(kWavesPerGroup * 64, tile_waves_per_eu(R, SORT, sizeof(X) > sizeof(ACC))) void
k_spmv_tiles(
    const int* __restrict__ bptr, const unsigned char* __restrict__ hdr, const V* __restrict__ val,
    const X* __restrict__ x, X* __restrict__ y, const int* __restrict__ tile_br,
    const int* __restrict__ tile_val, int64_t ntiles, int64_t nrows, int64_t ncols, Epilogue ep) {
    __shared__ TileSmem<ACC, R, expand_cap<V>(), SORT> sm[kWavesPerGroup];
    const int w = threadIdx.x >> 6;
    const int64_t t = xcd_tile(w);
    if (t >= ntiles) return;  // whole wave leaves; no workgroup barrier is used
    Epilogue e = ep;
    if (e.dot) e.dot += t;  // one partial per tile
    spmv_tile<V, X, ACC, R, VS, MB, SORT>(sm[w], bptr, hdr, val, x, y, tile_br[t], tile_br[t + 1], tile_val[t],
                                          tile_val[t + 1], nrows, ncols, e);
}

// ---- access probe ----------------------------------------------------------

// Memory-access probe over a double beta(1,8) matrix (a measurement aid, not
// a product): the product's tiles on the product's grid, one wave per tile,
// batches of 64 blocks (one per lane, at most 512 values).  Every block's
// header (fetched one batch ahead) and the batch's values (coalesced,
// non-temporal) are read, and then
//   MODE 0: nothing more -- the matrix stream alone;
//   MODE 1: x at every value's column, expanded through LDS in value order as
//           the product gathers it;
//   MODE 2: as 1 with each column folded into a 1024-entry window of x at the
//           tile's first column: the same gathers, all L1 hits.
// There are no row sums and no y; the sum of the products reaches memory only
// on an impossible value, which keeps every load live.
template <int MODE>
__global__ __launch_bounds__(kWavesPerGroup * 64) void k_access_probe(
    const int* __restrict__ bptr, const unsigned char* __restrict__ hdr, const double* __restrict__ val,
    const double* __restrict__ x, const int* __restrict__ tile_br, const int* __restrict__ tile_val, int64_t ntiles,
    int64_t ncols, double* __restrict__ out) {
    constexpr int H = 5;  // int32 first column + one mask byte
    __shared__ int scol[kWavesPerGroup][512];
    const int w = threadIdx.x >> 6;
    const int64_t t = xcd_tile(w);
    if (t >= ntiles) return;
    const int lane = threadIdx.x & 63;
    int* sc = scol[w];
    const int b0 = bptr[tile_br[t]], b1 = bptr[tile_br[t + 1]];
    int64_t v = tile_val[t];
    double acc = 0.0;
    int cs = 0;
    unsigned m = 0;
    if (b0 + lane < b1) {
        cs = load_i32_unaligned(hdr + (int64_t)(b0 + lane) * H);
        m = hdr[(int64_t)(b0 + lane) * H + 4];
    }
    // the folded window (MODE 2): 1024 columns from the tile's first column
    const int64_t wbase = ncols >= 1024 ? (__shfl(cs, 0, 64) < ncols - 1024 ? __shfl(cs, 0, 64) : ncols - 1024) : 0;
    for (int b = b0; b < b1; b += 64) {
        const int cnt = __popc(m);
        const int incl = wave_incl_scan(cnt);
        const int total = __builtin_amdgcn_readlane(incl, 63);
        double a[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = lane + 64 * k;
            a[k] = q < total ? __builtin_nontemporal_load(val + v + q) : 0.0;
        }
        // next batch's header
        int ncs = 0;
        unsigned nm = 0;
        if (b + 64 + lane < b1) {
            ncs = load_i32_unaligned(hdr + (int64_t)(b + 64 + lane) * H);
            nm = hdr[(int64_t)(b + 64 + lane) * H + 4];
        }
        if constexpr (MODE == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) acc += a[k];
            acc += (double)cs;
        } else {
            int p = incl - cnt;
#pragma unroll
            for (int bit = 0; bit < 8; ++bit)
                if ((m >> bit) & 1u) sc[p++] = MODE == 1 ? cs + bit : (int)(wbase + ((cs + bit) & 1023));
            wave_sync();
            double xv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int q = lane + 64 * k;
                xv[k] = q < total ? x[sc[q]] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) acc += a[k] * xv[k];
            wave_sync();
        }
        v += total;
        cs = ncs;
        m = nm;
    }
    if (acc == 1.2345e300) out[0] = acc;
}

// ---- plan kernels --------------------------------------------------------

// Checks the block-row pointers: 0-based, non-decreasing, ending at nblocks.
__global__ void k_check_bptr(const int* __restrict__ bptr, int64_t nbr, int64_t nblocks, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nbr) return;
    const int v = bptr[i];
    bool bad = v < 0 || (int64_t)v > nblocks;
    if (i == 0) bad |= v != 0;
    if (i == nbr) bad |= (int64_t)v != nblocks;
    if (i < nbr) bad |= bptr[i + 1] < v;
    if (bad) atomicOr(err, 1);
}

// Values per block-row (popcount of every mask) and a column range check.
// One wave per 64 block-rows (the pointers already checked): the lanes read
// the group's headers in block order, coalesced, find each block's block-row
// by a binary search of the 65 group pointers in LDS and add its popcount to
// that block-row's LDS counter.
template <int R, int MB, int VS>
__global__ __launch_bounds__(256) void k_blockrow_values(const int* __restrict__ bptr,
                                                         const unsigned char* __restrict__ hdr, int64_t nbr,
                                                         int64_t ncols, int* __restrict__ cnt_out, int* err) {
    constexpr int H = 4 + R * MB;
    using MW = MaskWords<R * MB>;
    __shared__ int sptr[4][65];
    __shared__ int scnt[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t br0 = ((int64_t)blockIdx.x * 4 + w) * 64;
    if (br0 >= nbr) return;  // whole wave: nothing below synchronises across waves
    int* sp = sptr[w];
    int* sc = scnt[w];
    {
        const int64_t i = br0 + lane;
        sp[lane] = bptr[i < nbr ? i : nbr];
        if (lane == 0) sp[64] = bptr[br0 + 64 < nbr ? br0 + 64 : nbr];
        sc[lane] = 0;
    }
    wave_sync();
    const int b0 = sp[0], bE = sp[64];
    bool bad = false;
    for (int b = b0 + lane; b < bE; b += 64) {
        const unsigned char* h = hdr + (int64_t)b * H;
        const int cs = load_i32_unaligned(h);
        unsigned long long mw[MW::kWords];
        MW::load(h + 4, mw);
        int c = 0;
        unsigned rowor = 0;
        for (int k = 0; k < MW::kWords; ++k) {
            c += __popcll(mw[k]);
            for (int q = 0; q < 64 / VS; ++q) rowor |= (unsigned)((mw[k] >> (q * VS)) & ((1ull << VS) - 1ull));
        }
        if (rowor != 0) {
            const int hi = 31 - __builtin_clz(rowor);
            bad |= cs < 0 || (int64_t)cs + hi >= ncols;
        }
        // the block's block-row: the largest i with sp[i] <= b (sp[0] <= b < sp[64])
        int lo = 0, up = 64;
        while (up - lo > 1) {
            const int mid = (lo + up) >> 1;
            if (sp[mid] <= b) lo = mid;
            else up = mid;
        }
        atomicAdd(&sc[lo], c);
    }
    wave_sync();
    if (br0 + lane < nbr) cnt_out[br0 + lane] = sc[lane];
    if (bad) atomicOr(err, 2);
}

__device__ __forceinline__ bool tile_starts(const int* valptr, int64_t br, int target, int maxbr) {
    if (br % maxbr == 0) return true;
    return valptr[br] / target != valptr[br - 1] / target;
}

__global__ void k_tile_flags(const int* __restrict__ valptr, int64_t nbr, int target, int maxbr,
                             int* __restrict__ flags) {
    const int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= nbr) return;
    flags[br] = tile_starts(valptr, br, target, maxbr) ? 1 : 0;
}

__global__ void k_tile_scatter(const int* __restrict__ valptr, int64_t nbr, int target, int maxbr,
                               const int* __restrict__ idx, int64_t ntiles, int* __restrict__ tile_br,
                               int* __restrict__ tile_val) {
    const int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (br > nbr) return;
    if (br == nbr) {
        tile_br[ntiles] = (int)nbr;
        tile_val[ntiles] = valptr[nbr];
        return;
    }
    if (tile_starts(valptr, br, target, maxbr)) {
        tile_br[idx[br]] = (int)br;
        tile_val[idx[br]] = valptr[br];
    }
}

template <int R, int MB, int VS>
void launch_blockrow_values(const spc5_hip_mat& m, int* cnt, int* err, hipStream_t s) {
    const unsigned g = (unsigned)((m.nbr + 255) / 256);  // 4 waves of 64 block-rows per workgroup
    hipLaunchKernelGGL((k_blockrow_values<R, MB, VS>), dim3(g), dim3(256), 0, s, m.bptr, m.hdr, m.nbr, m.ncols, cnt,
                       err);
}

template <class V, class X, class ACC, int R, int VS, int MB, bool SORT>
void launch_tiles(const spc5_hip_mat& m, const void* x, void* y, hipStream_t s, const Epilogue& ep) {
    const int64_t groups = (m.plan.ntiles + kWavesPerGroup - 1) / kWavesPerGroup;
    const unsigned grid = (unsigned)((groups + 7) / 8 * 8);
    hipLaunchKernelGGL((k_spmv_tiles<V, X, ACC, R, VS, MB, SORT>), dim3(grid), dim3(kWavesPerGroup * 64), 0, s,
                       m.bptr, m.hdr, reinterpret_cast<const V*>(m.values), reinterpret_cast<const X*>(x),
                       reinterpret_cast<X*>(y), m.plan.tile_br->as<int>(), m.plan.tile_val->as<int>(),
                       m.plan.ntiles, m.nrows, m.ncols, ep);
}

template <class V, class X, class ACC, int VS, int MB>
int launch_spmv_vs(const spc5_hip_mat& m, const void* x, void* y, hipStream_t s, const Epilogue& ep) {
    const bool sort = m.row_sort != 0;
    switch (m.r) {
    case 1: launch_tiles<V, X, ACC, 1, VS, MB, false>(m, x, y, s, ep); break;
    case 2:
        if (sort) launch_tiles<V, X, ACC, 2, VS, MB, true>(m, x, y, s, ep);
        else launch_tiles<V, X, ACC, 2, VS, MB, false>(m, x, y, s, ep);
        break;
    case 4:
        if (sort) launch_tiles<V, X, ACC, 4, VS, MB, true>(m, x, y, s, ep);
        else launch_tiles<V, X, ACC, 4, VS, MB, false>(m, x, y, s, ep);
        break;
    case 8: launch_tiles<V, X, ACC, 8, VS, MB, false>(m, x, y, s, ep); break;  // row sort: r = 2, 4
    default:
        set_error("invalid r");
        return SPC5_HIP_ERR_INVALID;
    }
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

}  // namespace

void set_error(const std::string& msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
}

std::string hip_msg(hipError_t e, const char* what) {
    return std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
}

const char* last_error_cstr() {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_err_mu);
    copy = g_err;
    return copy.c_str();
}

int launch_spmv(const spc5_hip_mat& m, const void* x, void* y, hipStream_t s, const Epilogue& ep) {
    if (m.plan.ntiles == 0) return SPC5_HIP_OK;
    if (m.value_type == SPC5_HIP_F64) return launch_spmv_vs<double, double, double, 8, 1>(m, x, y, s, ep);
    if (m.value_type == SPC5_HIP_F32 && m.vs == 16) return launch_spmv_vs<float, float, FloatAcc, 16, 2>(m, x, y, s, ep);
    if (m.value_type == SPC5_HIP_F32 && m.vs == 8) return launch_spmv_vs<float, float, FloatAcc, 8, 1>(m, x, y, s, ep);
    if (m.value_type == SPC5_HIP_BF16) return launch_spmv_vs<unsigned short, float, FloatAcc, 8, 1>(m, x, y, s, ep);
    set_error("unsupported value type / VS");
    return SPC5_HIP_ERR_INVALID;
}

int launch_spmv_f64vec(const spc5_hip_mat& m, const double* x, double* y, hipStream_t s, const Epilogue& ep) {
    if (m.plan.ntiles == 0) return SPC5_HIP_OK;
    if (m.value_type == SPC5_HIP_F64) return launch_spmv_vs<double, double, double, 8, 1>(m, x, y, s, ep);
    if (m.value_type == SPC5_HIP_F32 && m.vs == 16) return launch_spmv_vs<float, double, FloatAcc, 16, 2>(m, x, y, s, ep);
    if (m.value_type == SPC5_HIP_F32 && m.vs == 8) return launch_spmv_vs<float, double, FloatAcc, 8, 1>(m, x, y, s, ep);
    if (m.value_type == SPC5_HIP_BF16) return launch_spmv_vs<unsigned short, double, FloatAcc, 8, 1>(m, x, y, s, ep);
    set_error("unsupported value type / VS");
    return SPC5_HIP_ERR_INVALID;
}

int build_plan(spc5_hip_mat& m, hipStream_t s) {
    Plan& p = m.plan;
    p.ntiles = 0;
    if (m.nbr == 0) return SPC5_HIP_OK;
    DeviceBuffer err;
    SPC5_HIP_TRY(err.alloc(sizeof(int)));
    SPC5_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), s));
    {
        const unsigned g = (unsigned)((m.nbr + 1 + 255) / 256);
        hipLaunchKernelGGL(k_check_bptr, dim3(g), dim3(256), 0, s, m.bptr, m.nbr, m.nblocks, err.as<int>());
        SPC5_HIP_TRY(hipGetLastError());
        int herr = 0;
        SPC5_HIP_TRY(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
        SPC5_HIP_TRY(hipStreamSynchronize(s));
        if (herr) {
            set_error("rowsSize is not a 0-based non-decreasing block pointer array ending at nblocks");
            return SPC5_HIP_ERR_INVALID;
        }
    }
    p.valptr = std::make_shared<DeviceBuffer>();
    SPC5_HIP_TRY(p.valptr->alloc(sizeof(int) * (size_t)(m.nbr + 1)));
    int* valptr = p.valptr->as<int>();
    SPC5_HIP_TRY(hipMemsetAsync(valptr + m.nbr, 0, sizeof(int), s));
    const int key = m.r * 100 + m.mb * 10 + (m.vs == 16 ? 1 : 0);
    switch (key) {
    case 110: launch_blockrow_values<1, 1, 8>(m, valptr, err.as<int>(), s); break;
    case 210: launch_blockrow_values<2, 1, 8>(m, valptr, err.as<int>(), s); break;
    case 410: launch_blockrow_values<4, 1, 8>(m, valptr, err.as<int>(), s); break;
    case 810: launch_blockrow_values<8, 1, 8>(m, valptr, err.as<int>(), s); break;
    case 121: launch_blockrow_values<1, 2, 16>(m, valptr, err.as<int>(), s); break;
    case 221: launch_blockrow_values<2, 2, 16>(m, valptr, err.as<int>(), s); break;
    case 421: launch_blockrow_values<4, 2, 16>(m, valptr, err.as<int>(), s); break;
    case 821: launch_blockrow_values<8, 2, 16>(m, valptr, err.as<int>(), s); break;
    default: set_error("unsupported (r, VS, mask bytes)"); return SPC5_HIP_ERR_INVALID;
    }
    SPC5_HIP_TRY(hipGetLastError());
    DeviceBuffer dtotal;
    SPC5_HIP_TRY(dtotal.alloc(sizeof(int)));
    // Counts are < 2^31 each; the scan is checked against nnz below.
    int rc = scan_exclusive_i32(valptr, m.nbr + 1, dtotal.as<int>(), s);
    if (rc) return rc;
    int herr = 0, hnnz = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipMemcpyAsync(&hnnz, valptr + m.nbr, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    if (herr) {
        set_error("a block touches a column outside [0, ncols)");
        return SPC5_HIP_ERR_INVALID;
    }
    if ((int64_t)hnnz != m.nnz) {
        set_error("nnz (" + std::to_string(m.nnz) + ") differs from the mask popcount (" + std::to_string(hnnz) + ")");
        return SPC5_HIP_ERR_INVALID;
    }
    return build_tiles(m, p, kMaxRows, default_tile_values(m.nnz), s);
}

// Values per tile aimed at when no tile size is set: about 8192 tiles, at
// least 512 and at most 8192 values each (a tile also ends every
// kMaxRows rows).
int default_tile_values(int64_t nnz) {
    int64_t target = nnz / 8192;
    if (target < 512) target = 512;
    if (target > 8192) target = 8192;
    return (int)target;
}

// Tile cut from the block-row value offsets: a tile starts where the running
// value count crosses a multiple of `target`, and at least every
// maxrows / r block-rows.
int build_tiles(const spc5_hip_mat& m, Plan& p, int maxrows, int target, hipStream_t s) {
    p.ntiles = 0;
    p.target = target;
    if (m.nbr == 0) return SPC5_HIP_OK;
    if (!p.valptr) p.valptr = m.plan.valptr;
    const int* valptr = p.valptr->as<int>();
    const int maxbr = maxrows / m.r > 0 ? maxrows / m.r : 1;
    int rc;
    DeviceBuffer dtotal;
    SPC5_HIP_TRY(dtotal.alloc(sizeof(int)));
    DeviceBuffer flags;
    SPC5_HIP_TRY(flags.alloc(sizeof(int) * (size_t)(m.nbr + 1)));
    const unsigned g = (unsigned)((m.nbr + 1 + 255) / 256);
    hipLaunchKernelGGL(k_tile_flags, dim3(g), dim3(256), 0, s, valptr, m.nbr, p.target, maxbr, flags.as<int>());
    SPC5_HIP_TRY(hipGetLastError());
    rc = scan_exclusive_i32(flags.as<int>(), m.nbr, dtotal.as<int>(), s);
    if (rc) return rc;
    int hn = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&hn, dtotal.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    p.ntiles = hn;
    p.tile_br = std::make_shared<DeviceBuffer>();
    p.tile_val = std::make_shared<DeviceBuffer>();
    SPC5_HIP_TRY(p.tile_br->alloc(sizeof(int) * (size_t)(hn + 1)));
    SPC5_HIP_TRY(p.tile_val->alloc(sizeof(int) * (size_t)(hn + 1)));
    hipLaunchKernelGGL(k_tile_scatter, dim3(g), dim3(256), 0, s, valptr, m.nbr, p.target, maxbr, flags.as<int>(),
                       (int64_t)hn, p.tile_br->as<int>(), p.tile_val->as<int>());
    SPC5_HIP_TRY(hipGetLastError());
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return SPC5_HIP_OK;
}

}  // namespace spc5

using namespace spc5;

extern "C" {

const char* spc5_hip_last_error(void) { return spc5::last_error_cstr(); }

const char* spc5_hip_version(void) { return "spc5_hip 0.1 (gfx950)"; }

int spc5_hip_mat_create(int value_type, int vs, int r, int64_t nrows, int64_t ncols, int64_t nnz, int64_t nblocks,
                        const int* d_rowsSize, const unsigned char* d_headers, const void* d_values, void* stream,
                        spc5_hip_mat_t* out) {
    SPC5_CHECK_ARG(out != nullptr, "out is null");
    *out = nullptr;
    const int mb = (vs == 16) ? 2 : 1;
    SPC5_CHECK_ARG(r == 1 || r == 2 || r == 4 || r == 8, "r must be 1, 2, 4 or 8");
    SPC5_CHECK_ARG(valid_format(value_type, vs, mb), "unsupported (value type, VS) combination");
    SPC5_CHECK_ARG(nrows >= 0 && ncols >= 0 && nnz >= 0 && nblocks >= 0, "negative size");
    SPC5_CHECK_ARG(nnz <= INT32_MAX && nblocks <= INT32_MAX && nrows <= INT32_MAX && ncols <= INT32_MAX,
                   "sizes must fit the 32-bit format");
    SPC5_CHECK_ARG(nrows == 0 || d_rowsSize != nullptr, "rowsSize is null");
    SPC5_CHECK_ARG(nblocks == 0 || d_headers != nullptr, "headers are null");
    SPC5_CHECK_ARG(nnz == 0 || d_values != nullptr, "values are null");
    std::unique_ptr<spc5_hip_mat> m(new spc5_hip_mat());
    m->value_type = value_type;
    m->vs = vs;
    m->mb = mb;
    m->r = r;
    m->nrows = nrows;
    m->ncols = ncols;
    m->nnz = nnz;
    m->nblocks = nblocks;
    m->nbr = (nrows + r - 1) / r;
    m->bptr = d_rowsSize;
    m->hdr = d_headers;
    m->values = d_values;
    const int rc = build_plan(*m, as_stream(stream));
    if (rc) return rc;
    *out = m.release();
    return SPC5_HIP_OK;
}

int spc5_hip_mat_create_values_view(spc5_hip_mat_t base, int value_type, const void* d_values, spc5_hip_mat_t* out) {
    SPC5_CHECK_ARG(base != nullptr && out != nullptr, "null handle");
    SPC5_CHECK_ARG(valid_format(value_type, base->vs, base->mb), "value type incompatible with the block format");
    SPC5_CHECK_ARG(base->nnz == 0 || d_values != nullptr, "values are null");
    std::unique_ptr<spc5_hip_mat> m(new spc5_hip_mat(*base));
    m->value_type = value_type;
    m->values = d_values;
    *out = m.release();
    return SPC5_HIP_OK;
}

int spc5_hip_mat_destroy(spc5_hip_mat_t h) {
    if (h) forget_spmm_plans(h);
    delete h;
    return SPC5_HIP_OK;
}

int spc5_hip_mat_info(spc5_hip_mat_t h, int64_t* ntiles, int64_t* nblockrows, int64_t* nblocks, int64_t* nnz) {
    SPC5_CHECK_ARG(h != nullptr, "null handle");
    if (ntiles) *ntiles = h->plan.ntiles;
    if (nblockrows) *nblockrows = h->nbr;
    if (nblocks) *nblocks = h->nblocks;
    if (nnz) *nnz = h->nnz;
    return SPC5_HIP_OK;
}

int spc5_hip_mat_set_option(spc5_hip_mat_t h, int option, int64_t value) {
    SPC5_CHECK_ARG(h != nullptr, "null handle");
    switch (option) {
    case SPC5_HIP_OPT_ROW_SORT:
        SPC5_CHECK_ARG(value == 0 || value == 1, "ROW_SORT must be 0 or 1");
        h->row_sort = (int)value;
        return SPC5_HIP_OK;
    case SPC5_HIP_OPT_SPMM_COLUMNS:
        SPC5_CHECK_ARG(value == 0 || value == 1, "SPMM_COLUMNS must be 0 or 1");
        h->spmm_cols = (int)value;
        return SPC5_HIP_OK;
    case SPC5_HIP_OPT_SPMM_X16:
        SPC5_CHECK_ARG(value == 0 || value == 1, "SPMM_X16 must be 0 or 1");
        h->spmm_x16 = (int)value;
        return SPC5_HIP_OK;
    case SPC5_HIP_OPT_TILE_VALUES: {
        SPC5_CHECK_ARG(value == 0 || (value >= 64 && value <= (1 << 24)), "TILE_VALUES must be 0 or in [64, 2^24]");
        const int target = value == 0 ? default_tile_values(h->nnz) : (int)value;
        if (h->nbr == 0 || target == h->plan.target) return SPC5_HIP_OK;
        Plan np;
        np.valptr = h->plan.valptr;
        const int rc = build_tiles(*h, np, kMaxRows, target, nullptr);
        if (rc) return rc;
        h->plan = np;  // shared buffers of the old plan stay alive while views or batches use them
        return SPC5_HIP_OK;
    }
    default:
        set_error("unknown option");
        return SPC5_HIP_ERR_INVALID;
    }
}

int spc5_hip_access_probe(spc5_hip_mat_t h, int mode, const void* d_x, void* d_out, void* stream) {
    SPC5_CHECK_ARG(h != nullptr && d_x != nullptr && d_out != nullptr, "null argument");
    SPC5_CHECK_ARG(h->value_type == SPC5_HIP_F64 && h->r == 1 && h->vs == 8 && h->mb == 1,
                   "the access probe takes a double beta(1,8) matrix");
    SPC5_CHECK_ARG(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    SPC5_CHECK_ARG(mode != 2 || h->ncols >= 1024, "mode 2 needs at least 1024 columns");
    const spc5_hip_mat& m = *h;
    if (m.plan.ntiles == 0) return SPC5_HIP_OK;
    const int64_t groups = (m.plan.ntiles + kWavesPerGroup - 1) / kWavesPerGroup;
    const unsigned grid = (unsigned)((groups + 7) / 8 * 8);
    const double* xv = reinterpret_cast<const double*>(d_x);
    const double* vv = reinterpret_cast<const double*>(m.values);
    double* o = reinterpret_cast<double*>(d_out);
    const int* tb = m.plan.tile_br->as<int>();
    const int* tv = m.plan.tile_val->as<int>();
    hipStream_t s = as_stream(stream);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL(k_access_probe<0>, dim3(grid), dim3(kWavesPerGroup * 64), 0, s, m.bptr, m.hdr, vv, xv, tb, tv,
                           m.plan.ntiles, m.ncols, o);
        break;
    case 1:
        hipLaunchKernelGGL(k_access_probe<1>, dim3(grid), dim3(kWavesPerGroup * 64), 0, s, m.bptr, m.hdr, vv, xv, tb, tv,
                           m.plan.ntiles, m.ncols, o);
        break;
    default:
        hipLaunchKernelGGL(k_access_probe<2>, dim3(grid), dim3(kWavesPerGroup * 64), 0, s, m.bptr, m.hdr, vv, xv, tb, tv,
                           m.plan.ntiles, m.ncols, o);
        break;
    }
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

int spc5_hip_spmv(spc5_hip_mat_t h, const void* d_x, void* d_y, void* stream) {
    SPC5_CHECK_ARG(h != nullptr, "null handle");
    SPC5_CHECK_ARG(h->nnz == 0 || (d_x != nullptr && d_y != nullptr), "x or y is null");
    return launch_spmv(*h, d_x, d_y, as_stream(stream));
}

}  // extern "C"
