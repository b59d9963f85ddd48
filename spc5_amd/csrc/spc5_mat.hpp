// spc5_mat.hpp -- the device-resident matrix handle and the SpMV tile plan.
#pragma once

#include "spc5_common.hpp"

#include <memory>
#include <vector>

namespace spc5 {

// Work plan of one matrix: a tile is a run of whole block-rows processed by
// one wave.  Tiles are cut where the running value count crosses a multiple
// of the value target, and at least every kMaxRows/r block-rows, so every
// tile's rows fit the wave's LDS accumulators and every tile carries about
// the same number of values.
#ifndef SPC5_MAX_ROWS
#define SPC5_MAX_ROWS 128
#endif
#ifndef SPC5_WAVES_PER_GROUP
#define SPC5_WAVES_PER_GROUP 1
#endif
#ifndef SPC5_EXPAND_CAP
#define SPC5_EXPAND_CAP 384  // values expanded into LDS per batch, 8-byte values
#endif
#ifndef SPC5_EXPAND_CAP_NARROW
#define SPC5_EXPAND_CAP_NARROW 512  // same, 4- and 2-byte values (fixed costs weigh more per byte)
#endif
#ifndef SPC5_SEGREDUCE
#define SPC5_SEGREDUCE 1
#endif
#ifndef SPC5_XCD_REMAP
#define SPC5_XCD_REMAP 1
#endif
// Accumulator type of the float and bfloat16 kernels' LDS row sums (the
// products are formed in it too): SPC5_FLOAT_ACC_DOUBLE = 1 selects double.
#ifndef SPC5_FLOAT_ACC_DOUBLE
#define SPC5_FLOAT_ACC_DOUBLE 0
#endif
#if SPC5_FLOAT_ACC_DOUBLE
using FloatAcc = double;
#else
using FloatAcc = float;
#endif
constexpr int kMaxRows = SPC5_MAX_ROWS;              // scalar rows per tile (LDS accumulators), <= 256
constexpr int kWavesPerGroup = SPC5_WAVES_PER_GROUP; // independent waves per workgroup
static_assert(kMaxRows <= 255, "tile-local rows are stored as bytes, kMaxRows marks none");
// Values expanded into LDS per batch for value type V.
template <class V>
constexpr int expand_cap() {
    return sizeof(V) == 8 ? SPC5_EXPAND_CAP : SPC5_EXPAND_CAP_NARROW;
}
static_assert(SPC5_EXPAND_CAP % 64 == 0 && SPC5_EXPAND_CAP >= 128 && SPC5_EXPAND_CAP < 1024,
              "expansion cap: whole chunks, a full block, 10-bit counts");
static_assert(SPC5_EXPAND_CAP_NARROW % 64 == 0 && SPC5_EXPAND_CAP_NARROW >= 128 && SPC5_EXPAND_CAP_NARROW < 1024,
              "expansion cap: whole chunks, a full block, 10-bit counts");

// Output stage of a product: y = alpha * (A x) + beta * y (y is not read
// when beta == 0); with dv != null, tile t writes sum(dv[i] * y_new[i]) over
// its rows to dot[t] (one partial per tile, summed by the caller).
struct Epilogue {
    double alpha = 1.0;
    double beta = 1.0;
    const void* dv = nullptr;
    double* dot = nullptr;
};

struct Plan {
    int64_t ntiles = 0;
    int target = 0;                   // values per tile aimed at
    std::shared_ptr<DeviceBuffer> valptr;    // int32[nbr+1]: first value of each block-row
    std::shared_ptr<DeviceBuffer> tile_br;   // int32[ntiles+1]: first block-row of each tile
    std::shared_ptr<DeviceBuffer> tile_val;  // int32[ntiles+1]: first value of each tile
};

}  // namespace spc5

struct spc5_hip_mat {
    int value_type = SPC5_HIP_F64;
    int vs = 8;
    int mb = 1;          // mask bytes per row
    int r = 1;
    int64_t nrows = 0, ncols = 0, nnz = 0, nblocks = 0, nbr = 0;
    const int* bptr = nullptr;            // device, [nbr+1]
    const unsigned char* hdr = nullptr;   // device, [nblocks*(4+r*mb)]
    const void* values = nullptr;         // device, [nnz]
    int row_sort = 0;                     // SPC5_HIP_OPT_ROW_SORT
    int spmm_cols = 0;                    // SPC5_HIP_OPT_SPMM_COLUMNS
    int spmm_x16 = -1;                    // SPC5_HIP_OPT_SPMM_X16; -1: 1 for double vectors, else 0
    spc5::Plan plan;
    std::vector<std::shared_ptr<spc5::DeviceBuffer>> owned;  // buffers the handle allocated itself
};

namespace spc5 {

// Launch y = alpha A x + beta y for a handle (defined in spmv.hip).
int launch_spmv(const spc5_hip_mat& m, const void* x, void* y, hipStream_t s, const Epilogue& ep = Epilogue());
// The same product with double x and y whatever the value type (float and
// bfloat16 values are multiplied and summed in float; defined in spmv.hip).
int launch_spmv_f64vec(const spc5_hip_mat& m, const double* x, double* y, hipStream_t s,
                       const Epilogue& ep = Epilogue());
// Build the tile plan of a handle (defined in spmv.hip).
int build_plan(spc5_hip_mat& m, hipStream_t s);
// Values per tile of the default plan (defined in spmv.hip).
int default_tile_values(int64_t nnz);
// Another tile list of the same matrix with at most `maxrows` rows per tile
// (needs the value offsets of m.plan; defined in spmv.hip).
int build_tiles(const spc5_hip_mat& m, Plan& p, int maxrows, int target, hipStream_t s);
// Drops the multi-right-hand-side tile lists cached for a handle (spmm.hip).
void forget_spmm_plans(const spc5_hip_mat* m);

inline bool valid_format(int vt, int vs, int mb) {
    return (vt == SPC5_HIP_F64 && vs == 8 && mb == 1) || (vt == SPC5_HIP_F32 && vs == 16 && mb == 2) ||
           (vt == SPC5_HIP_F32 && vs == 8 && mb == 1) || (vt == SPC5_HIP_BF16 && vs == 8 && mb == 1);
}

inline size_t value_bytes(int vt) { return vt == SPC5_HIP_F64 ? 8 : (vt == SPC5_HIP_F32 ? 4 : 2); }

}  // namespace spc5
