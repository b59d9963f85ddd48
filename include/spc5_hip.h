/*
 * spc5_hip.h -- C ABI of the MI355X (gfx950) SPC5 library, libspc5_hip.so.
 *
 * Two groups of entry points:
 *
 *  1. The reference's own C ABI, same names, same arguments, same
 *     semantics (y += A*x on host pointers).  A program built against the
 *     reference header links against this library instead and its SpMV
 *     calls run on the GPU.  These copy their inputs to the GPU, run the
 *     HIP kernel and copy y back; they are the drop-in, not the fast path.
 *
 *  2. spc5_hip_* -- the device-resident API: caller-owned device buffers,
 *     explicit sizes, an explicit stream (a hipStream_t passed as void*;
 *     NULL = the default stream), an int status code (SPC5_HIP_OK = 0), and
 *     a message from spc5_hip_last_error() on failure.  Nothing here throws.
 *
 * Matrix layout (identical to the reference, src/spc5.hpp:44-62, 165-228):
 *   rowsSize[ceil(N/r)+1]   int32 block-row pointers (first block of each
 *                           block-row; the last entry is the block count)
 *   headers[nb*(4+r*MB)]    per block: int32 first column (little endian)
 *                           then r masks of MB bytes; bit k of mask s set
 *                           <=> entry (row b*r+s, column start+k) is stored
 *   values[nnz]             stored entries, block after block, and inside a
 *                           block row after row, columns ascending
 *   double: VS=8 columns, MB=1; float: VS=16, MB=2 (the reference formats).
 *   Extension formats of this library: float or bfloat16 values with VS=8,
 *   MB=1 (the double-precision block structure with cheaper values).
 */
#ifndef SPC5_HIP_H
#define SPC5_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes                                                           */
/* ---------------------------------------------------------------------- */
#define SPC5_HIP_OK 0
#define SPC5_HIP_ERR_INVALID 1     /* bad argument or unsupported combination */
#define SPC5_HIP_ERR_HIP 2         /* a HIP runtime call failed */
#define SPC5_HIP_ERR_ALLOC 3       /* device allocation failed */
#define SPC5_HIP_ERR_OVERFLOW 4    /* a count does not fit the 32-bit format */

/* Value types */
#define SPC5_HIP_F64 0   /* values, x, y and sums in double */
#define SPC5_HIP_F32 1   /* values, x, y and sums in float */
#define SPC5_HIP_BF16 2  /* bfloat16 values (raw uint16), float x, y and sums */

/* ---------------------------------------------------------------------- */
/* 1. Reference C ABI (drop-in).                                          */
/* ---------------------------------------------------------------------- */

/* y[i] += (A x)[i] for the beta(r,VS) matrix given by the reference arrays.
 * nbRows is the number of scalar rows; rowsSizes is indexed by block-row and
 * only differences are used; headers and values are read from their first
 * element.  x must hold every column any block touches (the reference pads
 * it by VS-1); rows past nbRows are not written.
 * Replaces src/spc5.hpp:516-523 (1rVc), 641-648 (2rVc), 769-776 (4rVc),
 * 895-902 (8rVc).  Aborts with a message on a GPU error (the reference
 * interface has no error channel). */
void core_SPC5_1rVc_Spmv_double(const long int nbRows, const int* rowsSizes,
                                const unsigned char* blocksColumnIndexesWithMasks,
                                const double* values, const double* x, double* y);
void core_SPC5_1rVc_Spmv_float(const long int nbRows, const int* rowsSizes,
                               const unsigned char* blocksColumnIndexesWithMasks,
                               const float* values, const float* x, float* y);
void core_SPC5_2rVc_Spmv_double(const long int nbRows, const int* rowsSizes,
                                const unsigned char* blocksColumnIndexesWithMasks,
                                const double* values, const double* x, double* y);
void core_SPC5_2rVc_Spmv_float(const long int nbRows, const int* rowsSizes,
                               const unsigned char* blocksColumnIndexesWithMasks,
                               const float* values, const float* x, float* y);
void core_SPC5_4rVc_Spmv_double(const long int nbRows, const int* rowsSizes,
                                const unsigned char* blocksColumnIndexesWithMasks,
                                const double* values, const double* x, double* y);
void core_SPC5_4rVc_Spmv_float(const long int nbRows, const int* rowsSizes,
                               const unsigned char* blocksColumnIndexesWithMasks,
                               const float* values, const float* x, float* y);
void core_SPC5_8rVc_Spmv_double(const long int nbRows, const int* rowsSizes,
                                const unsigned char* blocksColumnIndexesWithMasks,
                                const double* values, const double* x, double* y);
void core_SPC5_8rVc_Spmv_float(const long int nbRows, const int* rowsSizes,
                               const unsigned char* blocksColumnIndexesWithMasks,
                               const float* values, const float* x, float* y);

/* dest[i] += src[i] for i < nbValues (host pointers).
 * Replaces src/spc5.hpp:434-435 (declared), 2157-2175 (AVX-512 body). */
void SPC5_opti_merge_double(double dest[], const double src[], const int nbValues);
void SPC5_opti_merge_float(float dest[], const float src[], const int nbValues);

/* ---------------------------------------------------------------------- */
/* 2. Device-resident API                                                 */
/* ---------------------------------------------------------------------- */

/* Message of the last failing call in this process ("" if none). */
const char* spc5_hip_last_error(void);

/* Library version string. */
const char* spc5_hip_version(void);

typedef struct spc5_hip_mat* spc5_hip_mat_t;

/* Wrap device arrays in the reference layout (above) into a matrix handle
 * and build its work plan on the GPU (block-row value offsets and the
 * nnz-balanced tile list).  The arrays stay caller-owned and must outlive
 * the handle.  (value_type, vs, mask bytes) must be one of
 *   (F64, 8, 1), (F32, 16, 2), (F32, 8, 1), (BF16, 8, 1).
 * Counterpart of holding an SPC5Mat after CSR_to_SPC5 (src/spc5.hpp:44-62,
 * 989-1018) plus the work split of core_SPC5_rVc_threadsplit
 * (src/spc5.hpp:346-427), which the tile plan replaces. */
int spc5_hip_mat_create(int value_type, int vs, int r, int64_t nrows, int64_t ncols,
                        int64_t nnz, int64_t nblocks, const int* d_rowsSize,
                        const unsigned char* d_headers, const void* d_values,
                        void* stream, spc5_hip_mat_t* out);

/* Same, for a matrix whose block structure is shared with another handle
 * (same r, VS, rowsSize and headers) but with other values: used for the
 * float/bfloat16 copies of a double matrix.  Reuses the plan of `base`. */
int spc5_hip_mat_create_values_view(spc5_hip_mat_t base, int value_type, const void* d_values,
                                    spc5_hip_mat_t* out);

int spc5_hip_mat_destroy(spc5_hip_mat_t h);

/* Plan and shape information: number of tiles, block-rows, blocks, nnz. */
int spc5_hip_mat_info(spc5_hip_mat_t h, int64_t* ntiles, int64_t* nblockrows, int64_t* nblocks,
                      int64_t* nnz);

/* Serialisation of a matrix in the beta(r,VS) layout (no counterpart in
 * the reference; amortises conversion across runs).  save copies the
 * handle's arrays to a file; load reads one into a handle that owns its
 * device copies (freed by spc5_hip_mat_destroy) and builds its plan.  File:
 * "SPC5BRVC", uint32 version 1, int32 value type / VS / mask bytes / r / 0,
 * int64 nrows / ncols / nnz / nblocks, then rowsSize[ceil(nrows/r)+1]
 * (int32), the headers and the values, little-endian. */
int spc5_hip_mat_save(spc5_hip_mat_t h, const char* path);
int spc5_hip_mat_load(const char* path, spc5_hip_mat_t* out);

/* Per-handle kernel options (spc5_hip_mat_set_option).
 * ROW_SORT (0 or 1, default 0; used for r = 2 and 4): the products of every batch of
 * blocks are summed in row order instead of block order.  Pays off where
 * every block of a block-row covers the same rows (stencils: all blocks of
 * a block-row add to the same r rows at the same step), costs extra work
 * for irregular masks.  Results agree with the default order to rounding.
 * No counterpart in the reference (a property of the GPU kernel). */
#define SPC5_HIP_OPT_ROW_SORT 1
/* SPMM_COLUMNS (0 or 1, default 0; used for r = 2, 4 and 8): spc5_hip_spmm
 * requests each X row once per set column of a block and applies it to the
 * block's r rows, instead of once per stored value.  Pays off where the
 * rows of a block share columns (stencils); results agree with the default
 * to rounding.  No counterpart in the reference. */
#define SPC5_HIP_OPT_SPMM_COLUMNS 2
/* TILE_VALUES (0, or 64 .. 2^24): values one wave's tile aims at; a tile
 * also ends every 128 rows.  0 restores the default (nnz / 8192, clamped
 * to [512, 8192]).  Rebuilds the tile plan (synchronises the device); a
 * batch created afterwards uses the new tiles.  No counterpart in the
 * reference (the work split of the GPU kernel). */
#define SPC5_HIP_OPT_TILE_VALUES 3
/* SPMM_X16 (0 or 1; default 1 for double vectors, 0 for float): in
 * spc5_hip_spmm each lane requests 16 bytes of an X row (two doubles or four
 * floats, adjacent columns) instead of one element, where X allows it
 * (16-byte aligned X, a leading dimension, k and column chunk that are
 * multiples of the element count); otherwise one element per lane.  Fewer
 * lanes then share a record, which halves (double) or quarters (float) the
 * address work per column.  Results agree with one element per lane to
 * rounding.  No counterpart in the reference. */
#define SPC5_HIP_OPT_SPMM_X16 4
int spc5_hip_mat_set_option(spc5_hip_mat_t h, int option, int64_t value);

/* y += A x on device pointers (x and y in the matrix's vector type).
 * Replaces core_SPC5_<r>rVc_Spmv_<T> (src/spc5.hpp:2309-2715) and the
 * OpenMP driver SPC5_<r>rVc_Spmv_omp (src/spc5.hpp:557-599 etc.). */
int spc5_hip_spmv(spc5_hip_mat_t h, const void* d_x, void* d_y, void* stream);

/* COO (device) -> CSR (device): the reference's COO_unsorted_to_CSR
 * (src/spc5.hpp:106-114; sorted = 0) and COO_sorted_to_CSR (77-104;
 * sorted = 1, order checked).  Entries are ordered by (row, column) with a
 * stable radix sort; duplicates are kept in input order.  value_bytes is 8
 * (double) or 4 (float).  d_rowptr gets nrows+1 entries, d_col / d_val nnz.
 * Indices out of range or an unsorted input with sorted = 1 are errors. */
int spc5_hip_coo_to_csr(int value_bytes, int64_t nrows, int64_t ncols, int64_t nnz, int sorted, const int* d_i,
                        const int* d_j, const void* d_v, int* d_rowptr, int* d_col, void* d_val, void* stream);

/* MatrixMarket coordinate files (real, integer or pattern; general,
 * symmetric or skew-symmetric), read as examples/load_mm_and_compare.cpp
 * loadMM (287-428) reads them: 0-based indices, pattern values 1, the
 * mirrored entry of every off-diagonal symmetric (negated for
 * skew-symmetric) entry right after it.  info gives the dimensions and the
 * entry count after mirroring; read fills host arrays of `capacity` entries
 * (the sort by (row, column) is spc5_hip_coo_to_csr). */
int spc5_hip_mm_info(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nentries);
int spc5_hip_mm_read(const char* path, int64_t capacity, int* i, int* j, double* v, int64_t* nread);

/* GPU converter, CSR (device) -> beta(r,VS), in two calls.
 * count: writes the block-row pointers into d_rowsSize_out[0..ceil(N/r)]
 *   (count pass + exclusive scan) and returns the block count in
 *   *nblocks (synchronises the stream to read it).
 * fill : writes the packed headers (nblocks*(4+r*MB) bytes) and the values
 *   permuted into block order (nnz * value_bytes).
 * d_rowsSize_out must not alias d_rowptr.  Columns must be sorted and
 * unique inside each row, as the reference requires.
 * Replaces CSR_to_SPC5_<r>rVc / CSR_to_SPC5 (src/spc5.hpp:165-228,
 * 455-488, 609-613, 738-742, 864-868, 989-1018) and the block counters
 * SPC5_<r>rVc_block_count (src/spc5.hpp:276-317, 510-513, 635-638,
 * 764-767, 890-893). */
int spc5_hip_convert_count(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx,
                           int* d_rowsSize_out, int64_t* nblocks, void* stream);
int spc5_hip_convert_fill(int value_bytes, int vs, int mask_bytes, int r, int64_t nrows,
                          const int* d_rowptr, const int* d_colidx, const void* d_values_in,
                          const int* d_rowsSize, unsigned char* d_headers_out, void* d_values_out,
                          void* stream);

/* The count pass without any allocation: same arguments, results and
 * errors as spc5_hip_convert_count.  No counterpart in the reference; it is
 * for callers that cannot allocate (serving loops, harnesses that forbid
 * hipMalloc).  Its flags, total and scan partials live in a fixed 64 KiB
 * device array, one copy per GPU, held until the process ends; that limits
 * the matrix to 16320 * 2048 - 1 (about 33.4 M) block-rows, and larger
 * sizes are refused with SPC5_HIP_ERR_INVALID before any launch.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (two
 * host threads, two streams) race.  With spc5_hip_convert_fill, which
 * allocates nothing either, a whole conversion runs allocation-free. */
int spc5_hip_convert_count_noalloc(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx,
                                   int* d_rowsSize_out, int64_t* nblocks, void* stream);

/* y = alpha * (A x) + beta * y on device vectors (y is not read when
 * beta == 0).  alpha = beta = 1 is spc5_hip_spmv.  No counterpart in the
 * reference (its kernels only accumulate); used by the solvers. */
int spc5_hip_spmv_ex(spc5_hip_mat_t h, double alpha, const void* d_x, double beta, void* d_y, void* stream);

/* ---------------------------------------------------------------------- */
/* Precision form: reduced-precision values and refinement solvers        */
/* (no counterpart in the reference).                                     */
/* ---------------------------------------------------------------------- */

typedef struct {
    int iterations;      /* CG iterations (IR: float CG iterations in total; mixed CG: all iterations) */
    int outer;           /* refinement steps (IR) or reliable updates (mixed CG) */
    double relres;       /* ||b - A x|| / ||b||, recomputed at the end */
    int tail_iterations; /* mixed CG: iterations of the double tail (included in `iterations`) */
    int reserved;        /* 0 */
} spc5_hip_solve_stats;

/* d_out[i] = d_in[i] rounded to float (to_type F32) or to bfloat16 through
 * float (to_type BF16, raw uint16), each step round-to-nearest-even.  Makes
 * the value array of a values_view of a double matrix. */
int spc5_hip_values_convert(int to_type, const double* d_in, void* d_out, int64_t n, void* stream);

/* Conjugate gradient on A x = b (A symmetric positive definite), vectors
 * in the handle's vector type, x holding the initial guess.  Stops when the
 * recurrence residual falls to rtol*||b|| (checked every 8 iterations) or
 * after maxit iterations; st->relres is the true residual. */
int spc5_hip_cg(spc5_hip_mat_t A, const void* d_b, void* d_x, double rtol, int maxit, spc5_hip_solve_stats* st,
                void* stream);

/* Mixed-precision iterative refinement around CG: residuals r = b - A64 x
 * and updates of x in double (A64: double values), correction equations
 * A_lo d = r/||r|| solved by float CG to inner_rtol (A_lo: float or
 * bfloat16 values, same matrix).  Stops at ||b - A64 x|| <= rtol ||b|| or
 * after max_outer refinements; x holds the initial guess. */
int spc5_hip_ir_cg(spc5_hip_mat_t A64, spc5_hip_mat_t Alo, const double* d_b, double* d_x, double rtol,
                   double inner_rtol, int max_outer, int max_inner, spc5_hip_solve_stats* st, void* stream);

/* Mixed CG with reliable updates: CG iterates with A_lo (float or bfloat16
 * values, products and row sums in float).  vec64 = 0: the iteration's
 * vectors are float and a float correction d is accumulated; vec64 = 1: the
 * vectors (and x) stay double, only the matrix is float.  Every 8
 * iterations, when the iterated residual has dropped by `delta` (0 < delta
 * < 1) since the last update, d is folded into x (double), r = b - A64 x is
 * recomputed with the double SpMV and replaces the iterated residual, and
 * CG goes on with its search direction kept.  Once the true residual is
 * below tail_rtol ||b|| (0: never), the rest of the solve is double CG with
 * A64, continuing from the same search direction.  Stops at
 * ||b - A64 x|| <= rtol ||b|| or after maxit iterations in total. */
int spc5_hip_mixed_cg(spc5_hip_mat_t A64, spc5_hip_mat_t Alo, const double* d_b, double* d_x, double rtol,
                      double delta, int maxit, int vec64, double tail_rtol, spc5_hip_solve_stats* st,
                      void* stream);

/* ---------------------------------------------------------------------- */
/* Batching form: many independent matrices in one persistent launch      */
/* (no counterpart in the reference).                                     */
/* ---------------------------------------------------------------------- */

/* Y = alpha * (A X) + beta * Y for k right-hand sides stored as row-major
 * dense blocks: X is ncols x k with leading dimension ldx, Y is nrows x k
 * with leading dimension ldy (device pointers, the matrix's vector type).
 * kt is the column chunk one pass handles (0 or >= k: all k columns in
 * one pass, the matrix read once); smaller chunks re-read the matrix k/kt
 * times with a smaller x working set.  Tile lists for the chunk width are
 * built on first use and kept with the handle. */
int spc5_hip_spmm(spc5_hip_mat_t h, int k, int kt, double alpha, const void* d_x, int64_t ldx, double beta,
                  void* d_y, int64_t ldy, void* stream);

typedef struct spc5_hip_batch* spc5_hip_batch_t;

/* Group nmats matrix handles (one value type and VS; each its own r) with
 * their device vectors x_i, y_i.  The tiles of all matrices form one work
 * list walked by `waves` persistent waves (0: as many as the kernel keeps
 * resident, 12-28 per compute unit; more waves than tiles means one wave per
 * tile, the tiles then dealt out by the hardware's workgroup dispatcher; a
 * kernel with only that r's code runs
 * when all matrices share one r, with row-sorted batches when that r is 2
 * or 4 and every handle has SPC5_HIP_OPT_ROW_SORT set at creation); the
 * launch is captured once into a hipGraph.  Handles and vectors are
 * borrowed and must outlive the batch. */
int spc5_hip_batch_create(int nmats, const spc5_hip_mat_t* mats, const void* const* d_x, void* const* d_y,
                          int waves, spc5_hip_batch_t* out);

/* y_i += A_i x_i for every matrix of the batch: one graph replay. */
int spc5_hip_batch_spmv(spc5_hip_batch_t b, void* stream);

/* Same product through a direct kernel launch (no graph). */
int spc5_hip_batch_spmv_direct(spc5_hip_batch_t b, void* stream);

int spc5_hip_batch_info(spc5_hip_batch_t b, int64_t* ntiles, int* waves);
int spc5_hip_batch_destroy(spc5_hip_batch_t b);

/* ---------------------------------------------------------------------- */
/* Benchmark workloads generated on the device (counterpart of the        */
/* reference driver's synthetic builddense,                               */
/* examples/load_mm_and_compare.cpp:431-449).                             */
/* ---------------------------------------------------------------------- */

/* Banded n x n CSR matrix, byte-identical to spc5_synth_banded_count/fill
 * of libspc5_synth.so: entry (i,j), |i-j| <= bw, kept on the diagonal and
 * where splitmix64(seed*0x100000001b3 ^ (i<<24 ^ j)) < thr; values U[-1,1]
 * (splitmix64 on seed+1 and the CSR position), diagonal + bw.  count writes
 * d_rowptr[0..n] and *nnz; fill writes the columns and values.  n < 2^24. */
int spc5_hip_synth_banded_count(int64_t n, int64_t bw, uint64_t thr, uint64_t seed, int* d_rowptr, int64_t* nnz,
                                void* stream);
int spc5_hip_synth_banded_fill(int64_t n, int64_t bw, uint64_t thr, uint64_t seed, const int* d_rowptr, int* d_col,
                               double* d_val, void* stream);

/* Read-dominated streaming probe (no counterpart in the reference; a
 * measurement aid): every wave reads one contiguous tile of tile_bytes of
 * d_src with coalesced non-temporal 8-byte loads (3 KiB in flight per
 * wave), as a product's waves stream their tiles, then the waves write write_bytes of
 * d_dst (which must hold at least 16 bytes even when write_bytes is 0).
 * Sizes are multiples of 16, tile_bytes of 1 KiB.  With a write share of
 * about 3 % it has the HBM mix of one SpMV, so its rate is the ceiling
 * such a kernel can reach. */
int spc5_hip_stream_probe(const void* d_src, int64_t read_bytes, void* d_dst, int64_t write_bytes,
                          int64_t tile_bytes, void* stream);

/* Memory-access probe of a double beta(1,8) product (no counterpart in the
 * reference; a measurement aid).  It walks the handle's product tiles on the
 * product's grid and reads every header and value as the product does, and
 * then: mode 0, nothing more (the matrix stream alone); mode 1, x at every
 * value's column in the product's gather order; mode 2, the same gathers
 * folded into a 1024-entry window of x (all L1 hits; needs ncols >= 1024).
 * No row sums, no y: timing the three modes splits a product's time into
 * streaming, gathering and the rest.  d_out must hold 8 bytes (written only
 * on an impossible sum). */
int spc5_hip_access_probe(spc5_hip_mat_t h, int mode, const void* d_x, void* d_out, void* stream);

/* ---------------------------------------------------------------------- */
/* 3. Host-pointer helpers (used by include/spc5_hip.hpp)                 */
/* ---------------------------------------------------------------------- */

/* Block count of a host CSR matrix, computed on the GPU.
 * Replaces SPC5_<r>rVc_block_count (src/spc5.hpp:276-317, 510-513,
 * 635-638, 764-767, 890-893). */
int spc5_hip_host_block_count(int vs, int r, int64_t nrows, int64_t nnz, const int* rowptr, const int* colidx,
                              int64_t* nblocks);

/* Host CSR -> beta(r,VS) host arrays, converted on the GPU (byte-identical
 * to the reference).  rowsSize_out receives ceil(N/r)+1 entries; the header
 * buffer must hold *nblocks*(4+r*MB) bytes (the call fails, reporting the
 * size needed, when headers_capacity is smaller).
 * Replaces core_CSR_to_SPC5_rVc / CSR_to_SPC5_1rVc (src/spc5.hpp:165-228,
 * 455-488). */
int spc5_hip_host_convert(int value_bytes, int vs, int mask_bytes, int r, int64_t nrows, int64_t nnz,
                          const int* rowptr, const int* colidx, const void* values_in, int* rowsSize_out,
                          unsigned char* headers_out, int64_t headers_capacity, void* values_out,
                          int64_t* nblocks);

/* Device-resident copy of host arrays in the reference layout, with its
 * plan; the handle owns the device copies (freed by spc5_hip_mat_destroy).
 * Counterpart of SPC5_<r>rVc_split_omp (src/spc5.hpp:551-554 etc.): the
 * one-off preparation before repeated products. */
int spc5_hip_mat_create_from_host(int value_type, int vs, int r, int64_t nrows, int64_t ncols, int64_t nnz,
                                  int64_t nblocks, const int* rowsSize, const unsigned char* headers,
                                  const void* values, spc5_hip_mat_t* out);

/* y += A x with host x (ncols entries) and y (nrows entries).
 * Counterpart of SPC5_<r>rVc_Spmv_omp (src/spc5.hpp:557-599 etc.). */
int spc5_hip_spmv_host(spc5_hip_mat_t h, const void* x, void* y);

#ifdef __cplusplus
}
#endif

#endif /* SPC5_HIP_H */
