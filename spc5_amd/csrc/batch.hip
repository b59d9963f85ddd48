// batch.hip -- many independent beta(r,VS) matrices in one persistent,
// hipGraph-captured launch (the batching form; no counterpart in the
// reference).
//
// The tiles of every matrix form one list (matrix index + local tile index
// per entry, built once at batch creation).  A persistent grid of
// `waves` waves walks the list with a static stride; each tile runs the
// same tile routine as the single-matrix kernel, with the matrix's own r
// (chosen per matrix, e.g. by the autotuner) selected at run time.  When
// every matrix has the same r, a kernel holding only that r's body runs
// instead (the four bodies together need far more registers than one); with
// r = 2 or 4 and row sorting switched on for every matrix, that kernel uses
// the row-sorted tile (as the single-matrix product does).
// The launch is captured once into a hipGraph; every product replays it.
#include "spc5_mat.hpp"
#include "spmv_tile.hpp"

#include <algorithm>
#include <vector>

using namespace spc5;

struct BatchEntry {
    const int* bptr;
    const unsigned char* hdr;
    const void* val;
    const void* x;
    void* y;
    const int* tile_br;
    const int* tile_val;
    int64_t nrows;
    int64_t ncols;
    int r;
    int pad;
};

struct spc5_hip_batch {
    int value_type = SPC5_HIP_F64;
    int vs = 8;
    int nmats = 0;
    int64_t ntiles = 0;
    int waves = 0;
    int only_r = 0;         // the r shared by every matrix, 0 if they differ
    int sorted = 0;         // only_r is 2 or 4 and every matrix has row sorting on
    DeviceBuffer entries;   // BatchEntry[nmats]
    DeviceBuffer tile_mat;  // int32[ntiles]: matrix of each tile
    DeviceBuffer tile_loc;  // int32[ntiles]: tile index inside its matrix
    std::vector<spc5_hip_mat_t> mats;  // borrowed handles
    // The tile lists the batch was built from: a matrix's plan may be
    // rebuilt later (SPC5_HIP_OPT_TILE_VALUES); the batch keeps its own.
    std::vector<std::shared_ptr<DeviceBuffer>> tiles_kept;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    ~spc5_hip_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
    }
};

namespace spc5 {
namespace {

template <class V, class ACC>
union BatchSmem {
    TileSmem<ACC, 1, expand_cap<V>()> r1;
    TileSmem<ACC, 2, expand_cap<V>()> r2;
    TileSmem<ACC, 4, expand_cap<V>()> r4;
    TileSmem<ACC, 8, expand_cap<V>()> r8;
};
#define SPC5_BATCH_TILE(RR, SMEM) \
    spmv_tile<V, X, ACC, RR, VS, MB>(SMEM, m.bptr, m.hdr, val, x, y, br0, br1, v0, v1, m.nrows, m.ncols, ep)

// Waves per SIMD of the batch kernel: with every r body (ONLY_R = 0) the
// most that compile without spills; with one r, the tile kernel's rule.
constexpr int batch_waves_per_eu(int only_r, bool wide_x, int value_bytes, bool sorted = false) {
    // the persistent loop's own state costs one wave (two with 8-byte values)
    return only_r == 0 ? 5 : tile_waves_per_eu(only_r, sorted, wide_x) - (value_bytes == 8 ? 2 : 1);
}

// The batch loop over tiles: tile t of the list runs body(matrix entry,
// tile bounds) on its wave.
template <class Body>
__device__ __forceinline__ void batch_tiles(const BatchEntry* __restrict__ mats, const int* __restrict__ tile_mat,
                                            const int* __restrict__ tile_loc, int64_t ntiles, Body&& body) {
    const int w = threadIdx.x >> 6;
    const int64_t stride = (int64_t)gridDim.x * kWavesPerGroup;
    // The tile index is the same on every lane (it comes from the wave index,
    // which the compiler cannot tell is uniform): read it as a scalar, so the
    // matrix descriptor is loaded into scalar registers.
    const int64_t t0 = (int64_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerGroup + w));
    for (int64_t t = t0; t < ntiles; t += stride) {
        const BatchEntry& m = mats[tile_mat[t]];
        const int lt = tile_loc[t];
        body(m, m.tile_br[lt], m.tile_br[lt + 1], m.tile_val[lt], m.tile_val[lt + 1]);
        wave_sync();
    }
}

// Every matrix has r = R and row sorting on.
template <class V, class X, class ACC, int VS, int MB, int R>
__global__ __launch_bounds__(kWavesPerGroup * 64, batch_waves_per_eu(R, sizeof(X) > sizeof(ACC), sizeof(V), true)) void
k_spmv_batch_sorted(const BatchEntry* __restrict__ mats, const int* __restrict__ tile_mat,
                    const int* __restrict__ tile_loc, int64_t ntiles) {
    static_assert(R == 2 || R == 4, "row-sorted tiles exist for r = 2 and 4");
    __shared__ TileSmem<ACC, R, expand_cap<V>(), true> sm[kWavesPerGroup];
    const int w = threadIdx.x >> 6;
    const Epilogue ep;
    batch_tiles(mats, tile_mat, tile_loc, ntiles, [&](const BatchEntry& m, int br0, int br1, int v0, int v1) {
        spmv_tile<V, X, ACC, R, VS, MB, true>(sm[w], m.bptr, m.hdr, reinterpret_cast<const V*>(m.val),
                                              reinterpret_cast<const X*>(m.x), reinterpret_cast<X*>(m.y), br0, br1,
                                              v0, v1, m.nrows, m.ncols, ep);
    });
}

template <class V, class X, class ACC, int VS, int MB, int ONLY_R>
__global__ __launch_bounds__(kWavesPerGroup * 64, batch_waves_per_eu(ONLY_R, sizeof(X) > sizeof(ACC), sizeof(V))) void
k_spmv_batch(const BatchEntry* __restrict__ mats,
                                                                     const int* __restrict__ tile_mat,
                                                                     const int* __restrict__ tile_loc,
                                                                     int64_t ntiles) {
    __shared__ BatchSmem<V, ACC> sm[kWavesPerGroup];
    const int w = threadIdx.x >> 6;
    const int64_t stride = (int64_t)gridDim.x * kWavesPerGroup;
    const Epilogue ep;
    // The tile index is the same on every lane (it comes from the wave index,
    // which the compiler cannot tell is uniform): read it as a scalar, so the
    // matrix descriptor is loaded into scalar registers.
    const int64_t t0 = (int64_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWavesPerGroup + w));
    for (int64_t t = t0; t < ntiles; t += stride) {
        const BatchEntry& m = mats[tile_mat[t]];
        const int lt = tile_loc[t];
        const int br0 = m.tile_br[lt], br1 = m.tile_br[lt + 1], v0 = m.tile_val[lt], v1 = m.tile_val[lt + 1];
        const V* val = reinterpret_cast<const V*>(m.val);
        const X* x = reinterpret_cast<const X*>(m.x);
        X* y = reinterpret_cast<X*>(m.y);
        switch (ONLY_R ? ONLY_R : m.r) {
        case 1:
            SPC5_BATCH_TILE(1, sm[w].r1);
            break;
        case 2:
            SPC5_BATCH_TILE(2, sm[w].r2);
            break;
        case 4:
            SPC5_BATCH_TILE(4, sm[w].r4);
            break;
        default:
            SPC5_BATCH_TILE(8, sm[w].r8);
            break;
        }
        wave_sync();
    }
}

template <class V, class X, class ACC, int VS, int MB, int ONLY_R>
void launch_batch_r(const spc5_hip_batch& b, hipStream_t s) {
    const unsigned groups = (unsigned)((b.waves + kWavesPerGroup - 1) / kWavesPerGroup);
    hipLaunchKernelGGL((k_spmv_batch<V, X, ACC, VS, MB, ONLY_R>), dim3(groups), dim3(kWavesPerGroup * 64), 0, s,
                       b.entries.as<const BatchEntry>(), b.tile_mat.as<const int>(), b.tile_loc.as<const int>(),
                       b.ntiles);
}

template <class V, class X, class ACC, int VS, int MB, int R>
void launch_batch_sorted(const spc5_hip_batch& b, hipStream_t s) {
    const unsigned groups = (unsigned)((b.waves + kWavesPerGroup - 1) / kWavesPerGroup);
    hipLaunchKernelGGL((k_spmv_batch_sorted<V, X, ACC, VS, MB, R>), dim3(groups), dim3(kWavesPerGroup * 64), 0, s,
                       b.entries.as<const BatchEntry>(), b.tile_mat.as<const int>(), b.tile_loc.as<const int>(),
                       b.ntiles);
}

template <class V, class X, class ACC, int VS, int MB>
void launch_batch(const spc5_hip_batch& b, hipStream_t s) {
    if (b.sorted) {
        if (b.only_r == 2) launch_batch_sorted<V, X, ACC, VS, MB, 2>(b, s);
        else launch_batch_sorted<V, X, ACC, VS, MB, 4>(b, s);
        return;
    }
    switch (b.only_r) {
    case 1: launch_batch_r<V, X, ACC, VS, MB, 1>(b, s); break;
    case 2: launch_batch_r<V, X, ACC, VS, MB, 2>(b, s); break;
    case 4: launch_batch_r<V, X, ACC, VS, MB, 4>(b, s); break;
    case 8: launch_batch_r<V, X, ACC, VS, MB, 8>(b, s); break;
    default: launch_batch_r<V, X, ACC, VS, MB, 0>(b, s); break;
    }
}

// Waves per SIMD the batch's kernel runs at (sizes the persistent grid).
int batch_kernel_waves(int value_type, int only_r, bool sorted) {
    const bool wide_x = false;  // batches use the matrices' own vector type
    return batch_waves_per_eu(only_r, wide_x, value_type == SPC5_HIP_F64 ? 8 : value_type == SPC5_HIP_F32 ? 4 : 2,
                              sorted);
}

int launch_batch_any(const spc5_hip_batch& b, hipStream_t s) {
    if (b.value_type == SPC5_HIP_F64)
        launch_batch<double, double, double, 8, 1>(b, s);
    else if (b.value_type == SPC5_HIP_F32 && b.vs == 16)
        launch_batch<float, float, FloatAcc, 16, 2>(b, s);
    else if (b.value_type == SPC5_HIP_F32)
        launch_batch<float, float, FloatAcc, 8, 1>(b, s);
    else
        launch_batch<unsigned short, float, FloatAcc, 8, 1>(b, s);
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

}  // namespace
}  // namespace spc5

extern "C" {

int spc5_hip_batch_create(int nmats, const spc5_hip_mat_t* mats, const void* const* d_x, void* const* d_y,
                          int waves, spc5_hip_batch_t* out) {
    SPC5_CHECK_ARG(out != nullptr && nmats > 0 && mats != nullptr && d_x != nullptr && d_y != nullptr,
                   "null argument or empty batch");
    *out = nullptr;
    std::unique_ptr<spc5_hip_batch> b(new spc5_hip_batch());
    b->value_type = mats[0]->value_type;
    b->vs = mats[0]->vs;
    b->nmats = nmats;
    std::vector<BatchEntry> ent((size_t)nmats);
    std::vector<int> tmat, tloc;
    for (int i = 0; i < nmats; ++i) {
        const spc5_hip_mat* m = mats[i];
        SPC5_CHECK_ARG(m != nullptr, "null matrix handle");
        SPC5_CHECK_ARG(m->value_type == b->value_type && m->vs == b->vs, "all matrices need one value type and VS");
        SPC5_CHECK_ARG(m->nnz == 0 || (d_x[i] != nullptr && d_y[i] != nullptr), "null x or y");
        BatchEntry& e = ent[(size_t)i];
        e.bptr = m->bptr;
        e.hdr = m->hdr;
        e.val = m->values;
        e.x = d_x[i];
        e.y = d_y[i];
        e.tile_br = m->plan.ntiles ? m->plan.tile_br->as<const int>() : nullptr;
        e.tile_val = m->plan.ntiles ? m->plan.tile_val->as<const int>() : nullptr;
        if (m->plan.ntiles) {
            b->tiles_kept.push_back(m->plan.tile_br);
            b->tiles_kept.push_back(m->plan.tile_val);
        }
        e.nrows = m->nrows;
        e.ncols = m->ncols;
        e.r = m->r;
        e.pad = 0;
        for (int64_t t = 0; t < m->plan.ntiles; ++t) {
            tmat.push_back(i);
            tloc.push_back((int)t);
        }
        b->mats.push_back(const_cast<spc5_hip_mat*>(m));
    }
    b->ntiles = (int64_t)tmat.size();
    SPC5_CHECK_ARG(b->ntiles <= 0x7ffffff0, "batch: more than 2^31 - 16 tiles");
    int dev = 0, ncu = 256;
    SPC5_HIP_TRY(hipGetDevice(&dev));
    SPC5_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    b->only_r = mats[0]->r;
    bool all_sorted = true;
    for (int i = 0; i < nmats; ++i) {
        if (mats[i]->r != b->only_r) b->only_r = 0;
        all_sorted = all_sorted && mats[i]->row_sort != 0;
    }
    b->sorted = (b->only_r == 2 || b->only_r == 4) && all_sorted ? 1 : 0;
    int64_t wv = waves > 0 ? waves : (int64_t)ncu * 4 * batch_kernel_waves(b->value_type, b->only_r, b->sorted != 0);
    if (wv > b->ntiles) wv = std::max<int64_t>(b->ntiles, 1);
    b->waves = (int)wv;
    SPC5_HIP_TRY(b->entries.alloc(sizeof(BatchEntry) * (size_t)nmats));
    SPC5_HIP_TRY(b->tile_mat.alloc(sizeof(int) * (size_t)std::max<int64_t>(b->ntiles, 1)));
    SPC5_HIP_TRY(b->tile_loc.alloc(sizeof(int) * (size_t)std::max<int64_t>(b->ntiles, 1)));
    SPC5_HIP_TRY(hipMemcpy(b->entries.p, ent.data(), sizeof(BatchEntry) * (size_t)nmats, hipMemcpyHostToDevice));
    if (b->ntiles) {
        SPC5_HIP_TRY(hipMemcpy(b->tile_mat.p, tmat.data(), sizeof(int) * tmat.size(), hipMemcpyHostToDevice));
        SPC5_HIP_TRY(hipMemcpy(b->tile_loc.p, tloc.data(), sizeof(int) * tloc.size(), hipMemcpyHostToDevice));
    }
    // Capture the persistent launch once; products replay the graph.
    SPC5_HIP_TRY(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    if (b->ntiles) {
        SPC5_HIP_TRY(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch_batch_any(*b, b->capture_stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
        if (rc) return rc;
        SPC5_HIP_TRY(e);
        b->graph = g;
        SPC5_HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    *out = b.release();
    return SPC5_HIP_OK;
}

int spc5_hip_batch_spmv(spc5_hip_batch_t b, void* stream) {
    SPC5_CHECK_ARG(b != nullptr, "null batch");
    if (!b->exec) return SPC5_HIP_OK;
    SPC5_HIP_TRY(hipGraphLaunch(b->exec, as_stream(stream)));
    return SPC5_HIP_OK;
}

int spc5_hip_batch_spmv_direct(spc5_hip_batch_t b, void* stream) {
    SPC5_CHECK_ARG(b != nullptr, "null batch");
    if (!b->ntiles) return SPC5_HIP_OK;
    return launch_batch_any(*b, as_stream(stream));
}

int spc5_hip_batch_info(spc5_hip_batch_t b, int64_t* ntiles, int* waves) {
    SPC5_CHECK_ARG(b != nullptr, "null batch");
    if (ntiles) *ntiles = b->ntiles;
    if (waves) *waves = b->waves;
    return SPC5_HIP_OK;
}

int spc5_hip_batch_destroy(spc5_hip_batch_t b) {
    delete b;
    return SPC5_HIP_OK;
}

}  // extern "C"
