/*
 * spc5_oracle.c -- CPU restatement of the SPC5 beta(r,VS) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * library is compared against.  It is loaded by tests/, by the smoke() entry
 * and by bench.py's cpu_baseline leg, and by nothing else.  The product
 * library (spc5_amd/libspc5_hip.so) never links or calls it.
 *
 * Every function restates an algorithm of the reference header
 * (src/spc5.hpp of berenger-eu/spc5) in plain C; the file:line each one
 * follows is given in its comment.  How each function is pinned (golden
 * vectors produced by the compiled reference, closed forms, brute force) is
 * listed in tests/test_oracle.py.
 *
 * Formats.  VS columns per block and a mask of `msz` bytes per row:
 *   double: VS=8,  msz=1  (src/spc5.hpp:31-38, 145-148)
 *   float : VS=16, msz=2  (src/spc5.hpp:40-42, 145-148)
 * The extension formats of this project (float values with VS=8/u8 masks)
 * use the same code with other (VS, msz) pairs.
 */
#include <stdint.h>
#include <string.h>
#include <limits.h>
#include <math.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Greedy blocking of one block-row.                                          */
/* src/spc5.hpp:165-228 (core_CSR_to_SPC5_rVc) and 455-488 (r=1 variant).    */
/* For a block-row of r rows with per-row cursors, a block starts at the      */
/* smallest column still unconsumed in any row; it takes, row by row in order,*/
/* every entry with column < start+VS; the values of a block are emitted row  */
/* after row.  The header is int32 start (host byte order) followed by the r  */
/* masks, msz bytes each, without alignment (src/spc5.hpp:215-216).           */
/* ------------------------------------------------------------------------ */

static int64_t convert_impl(int r, int vs, int msz, int vsz,
                            int nrows, const int *rowptr, const int *col,
                            const unsigned char *vin,
                            int *bptr, unsigned char *hdr, unsigned char *vout)
{
    const int64_t hsz = 4 + (int64_t)r * msz;
    int64_t nb = 0, nv = 0;
    int cur[8], end[8];
    int br = 0;
    for (int row0 = 0; row0 < nrows; row0 += r, ++br) {
        const int nsub = (nrows - row0 < r) ? (nrows - row0) : r;
        for (int s = 0; s < nsub; ++s) {
            cur[s] = rowptr[row0 + s];
            end[s] = rowptr[row0 + s + 1];
        }
        if (bptr) bptr[br] = (int)nb;
        for (;;) {
            int start = INT_MAX, any = 0;
            for (int s = 0; s < nsub; ++s) {
                if (cur[s] < end[s]) {
                    any = 1;
                    if (col[cur[s]] < start) start = col[cur[s]];
                }
            }
            if (!any) break;
            uint32_t mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int s = 0; s < nsub; ++s) {
                while (cur[s] < end[s] && (int64_t)col[cur[s]] < (int64_t)start + vs) {
                    mask[s] |= 1u << (col[cur[s]] - start);
                    if (vout) memcpy(vout + nv * vsz, vin + (int64_t)cur[s] * vsz, (size_t)vsz);
                    ++nv;
                    ++cur[s];
                }
            }
            if (hdr) {
                unsigned char *h = hdr + nb * hsz;
                memcpy(h, &start, 4);
                for (int s = 0; s < r; ++s) {
                    for (int k = 0; k < msz; ++k)
                        h[4 + s * msz + k] = (unsigned char)((mask[s] >> (8 * k)) & 0xffu);
                }
            }
            ++nb;
        }
    }
    if (bptr) bptr[br] = (int)nb;
    return nb;
}

/* Number of blocks of the beta(r,VS) form of a CSR matrix.
 * src/spc5.hpp:276-317 (core_SPC5_block_count).  Same greedy as above. */
EXPORT int64_t spc5o_block_count(int r, int vs, int nrows, const int *rowptr, const int *col)
{
    return convert_impl(r, vs, 1, 1, nrows, rowptr, col, NULL, NULL, NULL, NULL);
}

/* CSR -> beta(r,VS).  Writes bptr[0..ceil(nrows/r)] (block-row pointers,
 * src/spc5.hpp:220-226), the packed headers and the permuted values.
 * src/spc5.hpp:165-228 / 455-488.  Returns the number of blocks. */
EXPORT int64_t spc5o_convert(int r, int vs, int msz, int vsz, int nrows,
                             const int *rowptr, const int *col, const void *vin,
                             int *bptr, unsigned char *hdr, void *vout)
{
    return convert_impl(r, vs, msz, vsz, nrows, rowptr, col,
                        (const unsigned char *)vin, bptr, hdr, (unsigned char *)vout);
}

/* ------------------------------------------------------------------------ */
/* CSR SpMV, src/spc5.hpp:116-127 (CSR_Spmv_scalar): per row, sum the        */
/* products left to right from 0, then y[row] += sum.                         */
/* Products are rounded before the add (no contraction).                      */
/* ------------------------------------------------------------------------ */
EXPORT void spc5o_csr_spmv_f64(int nrows, const int *rowptr, const int *col,
                               const double *val, const double *x, double *y)
{
    for (int i = 0; i < nrows; ++i) {
        double sum = 0.0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            volatile double prod = x[col[p]] * val[p];
            sum += prod;
        }
        y[i] += sum;
    }
}

EXPORT void spc5o_csr_spmv_f32(int nrows, const int *rowptr, const int *col,
                               const float *val, const float *x, float *y)
{
    for (int i = 0; i < nrows; ++i) {
        float sum = 0.0f;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            volatile float prod = x[col[p]] * val[p];
            sum += prod;
        }
        y[i] += sum;
    }
}

/* ------------------------------------------------------------------------ */
/* beta(r,VS) SpMV with the reference kernel's arithmetic.                    */
/* src/spc5.hpp:2309-2715 (AVX-512 core_SPC5_{1,2,4,8}rVc_Spmv_*):            */
/*  - per block-row, r rows of VS lane accumulators start at 0;               */
/*  - per block, lane k of row s gets fma(v, x[start+k], acc) where v is the  */
/*    next value of the stream if bit k of mask_s is set, else 0 (the         */
/*    masked expand-load zero-fills; the x window is always read, so x must   */
/*    be padded by VS-1 and 0*Inf gives NaN as in the reference);             */
/*  - at the end of the block-row each row's lanes are summed by the halving  */
/*    tree of _mm512_reduce_add (lane k + lane k+VS/2, then k + k+VS/4, ...)  */
/*    and added to y[row]; y must be padded by r-1 (src/spc5.hpp:2622-2629).  */
/*  - headers and values are walked sequentially; bptr only gives the count   */
/*    of blocks of each block-row (only differences are used).                */
/* ------------------------------------------------------------------------ */
#define BETA_SPMV(NAME, T, FMA)                                                      \
EXPORT void NAME(int r, int vs, int msz, long nrows, const int *bptr,                \
                 const unsigned char *hdr, const T *values, const T *x, T *y)        \
{                                                                                    \
    const int64_t hsz = 4 + (int64_t)r * msz;                                        \
    T acc[8][16];                                                                    \
    for (long row0 = 0; row0 < nrows; row0 += r) {                                   \
        const long br = row0 / r;                                                    \
        for (int s = 0; s < r; ++s)                                                  \
            for (int k = 0; k < vs; ++k) acc[s][k] = (T)0;                           \
        for (int b = bptr[br]; b < bptr[br + 1]; ++b) {                              \
            int start;                                                               \
            memcpy(&start, hdr, 4);                                                  \
            for (int s = 0; s < r; ++s) {                                            \
                uint32_t m = 0;                                                      \
                for (int k = 0; k < msz; ++k)                                        \
                    m |= (uint32_t)hdr[4 + s * msz + k] << (8 * k);                  \
                for (int k = 0; k < vs; ++k) {                                       \
                    T v = (T)0;                                                      \
                    if (m & (1u << k)) v = *values++;                                \
                    acc[s][k] = FMA(v, x[start + k], acc[s][k]);                     \
                }                                                                    \
            }                                                                        \
            hdr += hsz;                                                              \
        }                                                                            \
        for (int s = 0; s < r; ++s) {                                                \
            for (int half = vs / 2; half >= 1; half /= 2)                            \
                for (int k = 0; k < half; ++k) acc[s][k] = acc[s][k] + acc[s][k + half]; \
            y[row0 + s] += acc[s][0];                                                \
        }                                                                            \
    }                                                                                \
}
BETA_SPMV(spc5o_beta_spmv_f64, double, fma)
BETA_SPMV(spc5o_beta_spmv_f32, float, fmaf)

/* ------------------------------------------------------------------------ */
/* Decode beta(r,VS) back to (row, col, value) triplets in stream order.      */
/* This is what core_SPC5_rVc_iterate (src/spc5.hpp:256-273) is meant to do;  */
/* the reference reads the mask of the wrong row for r>1; this restatement   */
/* uses mask_s of row s.  Returns the number of entries written.              */
/* ------------------------------------------------------------------------ */
EXPORT int64_t spc5o_decode(int r, int vs, int msz, int vsz, long nrows, const int *bptr,
                            const unsigned char *hdr, const void *values,
                            int *out_row, int *out_col, void *out_val)
{
    const int64_t hsz = 4 + (int64_t)r * msz;
    const unsigned char *vp = (const unsigned char *)values;
    int64_t n = 0;
    for (long row0 = 0; row0 < nrows; row0 += r) {
        const long br = row0 / r;
        for (int b = bptr[br]; b < bptr[br + 1]; ++b) {
            int start;
            memcpy(&start, hdr, 4);
            for (int s = 0; s < r; ++s) {
                uint32_t m = 0;
                for (int k = 0; k < msz; ++k) m |= (uint32_t)hdr[4 + s * msz + k] << (8 * k);
                for (int k = 0; k < vs; ++k) {
                    if (m & (1u << k)) {
                        out_row[n] = (int)(row0 + s);
                        out_col[n] = start + k;
                        memcpy((unsigned char *)out_val + n * vsz, vp, (size_t)vsz);
                        vp += vsz;
                        ++n;
                    }
                }
            }
            hdr += hsz;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* OpenMP interval split rule, src/spc5.hpp:346-386                           */
/* (core_SPC5_rVc_threadsplit, serial part).  Block-rows are walked in order; */
/* a new interval starts at row idxRow (not the first row, not past the last  */
/* thread) when |(t+1)*blocks/T - bptr[b]| < |(t+1)*blocks/T - bptr[b+1]|.   */
/* Writes start_row[T], nrows_out[T], value_offset[T].                         */
/* ------------------------------------------------------------------------ */
EXPORT void spc5o_thread_split(int r, int msz, int nrows, int nblocks, const int *bptr,
                               const unsigned char *hdr, int nthreads,
                               int *start_row, int *nrows_out, int *value_offset)
{
    const int64_t hsz = 4 + (int64_t)r * msz;
    const double per = (double)nblocks / (double)nthreads;
    int t = 0, nval = 0;
    start_row[0] = 0;
    value_offset[0] = 0;
    for (int row = 0; row < nrows; row += r) {
        const int b = row / r;
        if (t != nthreads - 1 && row != 0) {
            const double target = (double)(t + 1) * per;
            if (fabs(target - bptr[b]) < fabs(target - bptr[b + 1])) {
                nrows_out[t] = row - start_row[t];
                ++t;
                start_row[t] = row;
                value_offset[t] = nval;
            }
        }
        for (int blk = bptr[b]; blk < bptr[b + 1]; ++blk) {
            for (int s = 0; s < r; ++s) {
                uint32_t m = 0;
                for (int k = 0; k < msz; ++k) m |= (uint32_t)hdr[blk * hsz + 4 + s * msz + k] << (8 * k);
                nval += __builtin_popcount(m);
            }
        }
    }
    nrows_out[t] = nrows - start_row[t];
    for (++t; t < nthreads; ++t) {
        start_row[t] = nrows;
        value_offset[t] = nval;
        nrows_out[t] = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* Format-selection model, src/spc5.hpp:1151-1223 (SPC5_find_best) and       */
/* 1259-1304 (SPC5_find_best_omp<double>).  The speed estimate is a cubic    */
/* (sequential) or a bilinear form in threads (OpenMP) of the average number */
/* of values per block, evaluated for r = 1, 2, 4 only; the first strictly   */
/* larger estimate wins.  Coefficient tables: src/spc5.hpp:1163-1212 and      */
/* 1276-1294.  Returns the winning r (0 when none beats 0) and the estimate.  */
/* ------------------------------------------------------------------------ */
static const double FB_SEQ_F64[3][4] = {
    {2.186490e-03, -7.931533e-02, 8.719444e-01, 3.712380e-01},
    {9.822706e-04, -4.500002e-02, 6.714119e-01, 2.413985e-01},
    {1.320496e-04, -1.222130e-02, 3.520409e-01, 3.603195e-01}};
static const double FB_SEQ_F32[3][4] = {
    {-4.039852e-03, 7.997606e-02, 6.892119e-02, 1.508645e+00},
    {-9.195940e-04, 4.442050e-03, 5.273045e-01, 1.101610e+00},
    {-4.464933e-05, 1.867266e-03, 1.520910e-01, 1.186131e+00}};
static const double FB_OMP_F64[3][4] = {
    {4.3764, 0.0981, 1.8728, 0.0721},
    {5.2172, 0.1435, 0.8331, 0.0556},
    {5.3212, 0.1921, 0.5201, 0.0263}};

EXPORT int spc5o_find_best(int is_double, int nrows, const int *rowptr, const int *col,
                           int nnz, int nthreads, double *estimate)
{
    const int rs[3] = {1, 2, 4};
    const int vs = is_double ? 8 : 16;
    double best = 0.0;
    int best_r = 0;
    for (int i = 0; i < 3; ++i) {
        const int64_t nb = spc5o_block_count(rs[i], vs, nrows, rowptr, col);
        const double a = (double)nnz / (double)nb;
        double est;
        if (nthreads > 0) {
            const double *c = FB_OMP_F64[i];
            est = c[0] + c[1] * nthreads + c[2] * a + c[3] * nthreads * a;
        } else {
            const double *c = is_double ? FB_SEQ_F64[i] : FB_SEQ_F32[i];
            est = c[0] * (a * a * a) + c[1] * (a * a) + c[2] * a + c[3];
        }
        if (est > best) {
            best = est;
            best_r = rs[i];
        }
    }
    *estimate = best;
    return best_r;
}

/* ------------------------------------------------------------------------ */
/* COO (sorted by row then column) -> CSR, src/spc5.hpp:77-104: histogram of  */
/* row counts into rowptr[i+1], then prefix sum; values and columns copied in */
/* input order.  Sorting (src/spc5.hpp:107-114) is done by the caller.        */
/* ------------------------------------------------------------------------ */
EXPORT void spc5o_coo_sorted_to_csr(int nrows, int nnz, const int *ii, const int *jj,
                                    const double *vv, int *rowptr, int *col, double *val)
{
    for (int i = 0; i <= nrows; ++i) rowptr[i] = 0;
    for (int p = 0; p < nnz; ++p) {
        val[p] = vv[p];
        col[p] = jj[p];
        rowptr[ii[p] + 1] += 1;
    }
    for (int i = 2; i <= nrows; ++i) rowptr[i] += rowptr[i - 1];
}
