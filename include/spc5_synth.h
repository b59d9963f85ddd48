/*
 * spc5_synth.h -- seeded workload generators (libspc5_synth.so, host code).
 *
 * Counterpart of the reference driver's builddense / x setup
 * (examples/load_mm_and_compare.cpp:431-449, 482-485): the benchmark
 * matrices and vectors, generated from splitmix64 counters so that any row
 * can be regenerated alone and the output does not depend on thread count.
 * CSR output: rowptr[N+1], col[nnz] sorted unique per row, val[nnz] double.
 */
#ifndef SPC5_SYNTH_H
#define SPC5_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* out[i] = lo + (hi-lo) * u_i, u_i = (splitmix64(seed, i) >> 11) * 2^-53 */
void spc5_synth_uniform_f64(int64_t n, uint64_t seed, double lo, double hi, double* out);
void spc5_synth_uniform_f32(int64_t n, uint64_t seed, double lo, double hi, float* out);

/* Clustered runs: row lengths lognormal (mean 32, sigma 0.8, clipped to
 * [1,512]); runs of 1..16 contiguous columns starting uniformly in
 * [i-N/64, i+N/64] (clipped); overlaps merged.  count writes rowptr and
 * returns nnz (-1 above 2^31-1); fill writes columns and U[-1,1] values. */
int64_t spc5_synth_runs_count(int64_t N, uint64_t seed, int32_t* rowptr);
void spc5_synth_runs_fill(int64_t N, uint64_t seed, uint64_t value_seed, const int32_t* rowptr, int32_t* col,
                          double* val);

/* 5-point Laplacian on an n x n grid; returns nnz (= 5n^2 - 4n). */
int64_t spc5_synth_lap2d(int64_t n, double diag, double off, int32_t* rowptr, int32_t* col, double* val);

/* 27-point stencil on an n^3 grid; returns nnz (= (3n-2)^3), -1 on overflow. */
int64_t spc5_synth_stencil27(int64_t n, double diag, double off, int32_t* rowptr, int32_t* col, double* val);

/* Banded n x n: diagonal always, off-diagonals |i-j| <= bw kept with
 * probability fill; values U[-1,1], diagonal + bw. */
int64_t spc5_synth_banded_count(int64_t n, int64_t bw, double fill, uint64_t seed, int32_t* rowptr);
void spc5_synth_banded_fill(int64_t n, int64_t bw, double fill, uint64_t seed, const int32_t* rowptr, int32_t* col,
                            double* val);

#ifdef __cplusplus
}
#endif

#endif /* SPC5_SYNTH_H */
