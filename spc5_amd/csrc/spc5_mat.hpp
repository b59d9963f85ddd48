// spc5_mat.hpp -- the device-resident matrix handle and the SpMV tile plan.
#pragma once

#include "spc5_common.hpp"

#include <memory>

namespace spc5 {

// Work plan of one matrix: a tile is a run of whole block-rows processed by
// one wave.  Tiles are cut where the running value count crosses a multiple
// of the value target, and at least every kMaxRows/r block-rows, so every
// tile's rows fit the wave's LDS accumulators and every tile carries about
// the same number of values.
constexpr int kMaxRows = 256;        // scalar rows per tile (LDS accumulators)
constexpr int kWavesPerGroup = 4;    // independent waves per workgroup
constexpr int kExpandCap = 512;      // values expanded into LDS per batch

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
    spc5::Plan plan;
};

namespace spc5 {

// Launch y += A x for a handle (defined in spmv.hip).
int launch_spmv(const spc5_hip_mat& m, const void* x, void* y, hipStream_t s);
// Build the tile plan of a handle (defined in spmv.hip).
int build_plan(spc5_hip_mat& m, hipStream_t s);

inline bool valid_format(int vt, int vs, int mb) {
    return (vt == SPC5_HIP_F64 && vs == 8 && mb == 1) || (vt == SPC5_HIP_F32 && vs == 16 && mb == 2) ||
           (vt == SPC5_HIP_F32 && vs == 8 && mb == 1) || (vt == SPC5_HIP_BF16 && vs == 8 && mb == 1);
}

inline size_t value_bytes(int vt) { return vt == SPC5_HIP_F64 ? 8 : (vt == SPC5_HIP_F32 ? 4 : 2); }

}  // namespace spc5
