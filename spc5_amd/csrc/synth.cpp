// synth.cpp -- seeded, portable generators of the benchmark matrices and
// vectors (host C++/OpenMP).  Every random number comes from splitmix64 on a
// counter derived from (seed, row, index), so a row can be regenerated
// alone, results do not depend on the thread count, and the CPU baseline and
// the GPU path see the same bytes.  Uniform reals are u = (z >> 11) * 2^-53,
// mapped to lo + (hi - lo) * u.
#include "../../include/spc5_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t a, uint64_t b = 0) : s(splitmix64(splitmix64(seed ^ 0x5bd1e995ull) + a) ^ (b * 0x2545f4914f6cdd1dull)) {}
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

inline double uniform_at(uint64_t seed, uint64_t i, double lo, double hi) {
    const uint64_t z = splitmix64(splitmix64(seed ^ 0x9fb21c651e98df25ull) + i);
    return lo + (hi - lo) * ((double)(z >> 11) * 0x1.0p-53);
}

// Sorted unique columns of row i of the clustered-runs matrix.
void runs_row(int64_t N, uint64_t seed, int64_t i, std::vector<int>& cols) {
    Rng g(seed, (uint64_t)i);
    // Row length: lognormal with mean 32 and sigma 0.8, clipped to [1, 512].
    const double sigma = 0.8;
    const double mu = std::log(32.0) - 0.5 * sigma * sigma;
    double u1 = g.unit(), u2 = g.unit();
    if (u1 < 1e-300) u1 = 1e-300;
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    double Ld = std::round(std::exp(mu + sigma * z));
    const int L = (int)std::min(512.0, std::max(1.0, Ld));
    const int64_t half = N / 64;
    cols.clear();
    int placed = 0;
    while (placed < L) {
        const int run = 1 + (int)(g.next() % 16);
        int64_t c0 = i - half + (int64_t)(g.next() % (uint64_t)(2 * half + 1));
        c0 = std::max<int64_t>(0, std::min<int64_t>(N - 1, c0));
        for (int k = 0; k < run && placed < L; ++k, ++placed) {
            const int64_t c = c0 + k;
            if (c < N) cols.push_back((int)c);
        }
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
}

}  // namespace

extern "C" {

void spc5_synth_uniform_f64(int64_t n, uint64_t seed, double lo, double hi, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = uniform_at(seed, (uint64_t)i, lo, hi);
}

void spc5_synth_uniform_f32(int64_t n, uint64_t seed, double lo, double hi, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = (float)uniform_at(seed, (uint64_t)i, lo, hi);
}

int64_t spc5_synth_runs_count(int64_t N, uint64_t seed, int32_t* rowptr) {
    rowptr[0] = 0;
#pragma omp parallel
    {
        std::vector<int> cols;
        cols.reserve(1024);
#pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < N; ++i) {
            runs_row(N, seed, i, cols);
            rowptr[i + 1] = (int32_t)cols.size();
        }
    }
    int64_t acc = 0;
    for (int64_t i = 1; i <= N; ++i) {
        acc += rowptr[i];
        if (acc > INT32_MAX) return -1;
        rowptr[i] = (int32_t)acc;
    }
    return acc;
}

void spc5_synth_runs_fill(int64_t N, uint64_t seed, uint64_t value_seed, const int32_t* rowptr, int32_t* col,
                          double* val) {
#pragma omp parallel
    {
        std::vector<int> cols;
        cols.reserve(1024);
#pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < N; ++i) {
            runs_row(N, seed, i, cols);
            const int64_t p0 = rowptr[i];
            for (size_t k = 0; k < cols.size(); ++k) {
                col[p0 + (int64_t)k] = cols[k];
                val[p0 + (int64_t)k] = uniform_at(value_seed, (uint64_t)(p0 + (int64_t)k), -1.0, 1.0);
            }
        }
    }
}

int64_t spc5_synth_lap2d(int64_t n, double diag, double off, int32_t* rowptr, int32_t* col, double* val) {
    const int64_t N = n * n;
    rowptr[0] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const int64_t gi = i / n, gj = i % n;
        rowptr[i + 1] = (int32_t)(1 + (gi > 0) + (gi < n - 1) + (gj > 0) + (gj < n - 1));
    }
    for (int64_t i = 1; i <= N; ++i) rowptr[i] += rowptr[i - 1];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const int64_t gi = i / n, gj = i % n;
        int64_t p = rowptr[i];
        if (gi > 0) { col[p] = (int32_t)(i - n); val[p++] = off; }
        if (gj > 0) { col[p] = (int32_t)(i - 1); val[p++] = off; }
        col[p] = (int32_t)i; val[p++] = diag;
        if (gj < n - 1) { col[p] = (int32_t)(i + 1); val[p++] = off; }
        if (gi < n - 1) { col[p] = (int32_t)(i + n); val[p++] = off; }
    }
    return rowptr[N];
}

int64_t spc5_synth_stencil27(int64_t n, double diag, double off, int32_t* rowptr, int32_t* col, double* val) {
    const int64_t N = n * n * n;
    if ((3 * n - 2) * (3 * n - 2) * (3 * n - 2) > INT32_MAX) return -1;
    rowptr[0] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const int64_t z = i / (n * n), y = (i / n) % n, x = i % n;
        const int lz = 1 + (z > 0) + (z < n - 1), ly = 1 + (y > 0) + (y < n - 1), lx = 1 + (x > 0) + (x < n - 1);
        rowptr[i + 1] = lz * ly * lx;
    }
    for (int64_t i = 1; i <= N; ++i) rowptr[i] += rowptr[i - 1];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const int64_t z = i / (n * n), y = (i / n) % n, x = i % n;
        int64_t p = rowptr[i];
        for (int dz = -1; dz <= 1; ++dz) {
            if (z + dz < 0 || z + dz >= n) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                if (y + dy < 0 || y + dy >= n) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if (x + dx < 0 || x + dx >= n) continue;
                    const int64_t c = i + dz * n * n + dy * n + dx;
                    col[p] = (int32_t)c;
                    val[p++] = (dz == 0 && dy == 0 && dx == 0) ? diag : off;
                }
            }
        }
    }
    return rowptr[N];
}

// Banded matrix: |i-j| <= bw, each off-diagonal entry kept with probability
// `fill`, values U[-1,1], diagonal U[-1,1] + bw (strict dominance when
// fill*2bw < bw+... is not required; the diagonal shift follows the config).
int64_t spc5_synth_banded_count(int64_t n, int64_t bw, double fill, uint64_t seed, int32_t* rowptr) {
    rowptr[0] = 0;
    const uint64_t thr = (uint64_t)(fill * 18446744073709551616.0 >= 18446744073709551615.0 ? ~0ull
                                                                                             : fill * 18446744073709551616.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int c = 0;
        const int64_t j0 = std::max<int64_t>(0, i - bw), j1 = std::min<int64_t>(n - 1, i + bw);
        for (int64_t j = j0; j <= j1; ++j) {
            if (j == i || splitmix64(seed * 0x100000001b3ull ^ ((uint64_t)i << 24 ^ (uint64_t)j)) < thr) ++c;
        }
        rowptr[i + 1] = c;
    }
    int64_t acc = 0;
    for (int64_t i = 1; i <= n; ++i) {
        acc += rowptr[i];
        if (acc > INT32_MAX) return -1;
        rowptr[i] = (int32_t)acc;
    }
    return acc;
}

void spc5_synth_banded_fill(int64_t n, int64_t bw, double fill, uint64_t seed, const int32_t* rowptr, int32_t* col,
                            double* val) {
    const uint64_t thr = (uint64_t)(fill * 18446744073709551616.0 >= 18446744073709551615.0 ? ~0ull
                                                                                             : fill * 18446744073709551616.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int64_t p = rowptr[i];
        const int64_t j0 = std::max<int64_t>(0, i - bw), j1 = std::min<int64_t>(n - 1, i + bw);
        for (int64_t j = j0; j <= j1; ++j) {
            if (j == i || splitmix64(seed * 0x100000001b3ull ^ ((uint64_t)i << 24 ^ (uint64_t)j)) < thr) {
                col[p] = (int32_t)j;
                double v = uniform_at(seed + 1, (uint64_t)p, -1.0, 1.0);
                if (j == i) v += (double)bw;
                val[p++] = v;
            }
        }
    }
}

}  // extern "C"
