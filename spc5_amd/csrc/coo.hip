// coo.hip -- COO -> CSR on the GPU and the MatrixMarket coordinate reader.
//
// spc5_hip_coo_to_csr does what the reference's COO_unsorted_to_CSR /
// COO_sorted_to_CSR do (src/spc5.hpp:77-114) with the GPU: entries are
// ordered by (row, column) with a stable radix sort of 64-bit keys
// (row << 32 | column; hipCUB), the row pointers come from a binary search
// of every row's first key in the sorted keys, and columns and values are
// gathered through the sort's permutation.  Duplicate (row, column) entries
// are kept, in input order, as the reference keeps them.  With `sorted` set
// the sort is skipped and the order is checked instead (the reference
// trusts it).  Every index is range-checked.
//
// spc5_hip_mm_info / spc5_hip_mm_read parse a MatrixMarket coordinate file
// (real, integer or pattern; general, symmetric or skew-symmetric) into
// 0-based COO arrays with the mirrored entries of symmetric files appended
// after each stored one, as examples/load_mm_and_compare.cpp:287-428 reads
// it (the sort that follows there is spc5_hip_coo_to_csr here).
#include "spc5_mat.hpp"

#include <hipcub/hipcub.hpp>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace spc5;

namespace {

__global__ void k_coo_keys(const int* __restrict__ ri, const int* __restrict__ cj, int64_t nnz, int64_t nrows,
                           int64_t ncols, unsigned long long* __restrict__ keys, int* __restrict__ perm,
                           int* __restrict__ err) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    const int r = ri[t], c = cj[t];
    if (r < 0 || (int64_t)r >= nrows || c < 0 || (int64_t)c >= ncols) atomicOr(err, 1);
    keys[t] = ((unsigned long long)(unsigned)r << 32) | (unsigned)c;
    perm[t] = (int)t;
}

// err |= 2 where keys[t-1] > keys[t] (input said to be sorted but is not).
__global__ void k_check_sorted(const unsigned long long* __restrict__ keys, int64_t nnz, int* __restrict__ err) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (t >= nnz) return;
    if (keys[t - 1] > keys[t]) atomicOr(err, 2);
}

// rowptr[r] = number of keys below (r << 32), r = 0 .. nrows.
__global__ void k_rowptr_search(const unsigned long long* __restrict__ keys, int64_t nnz, int64_t nrows,
                                int* __restrict__ rowptr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    const unsigned long long key = (unsigned long long)r << 32;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    rowptr[r] = (int)lo;
}

template <class T>
__global__ void k_coo_gather(const unsigned long long* __restrict__ keys, const int* __restrict__ perm,
                             const T* __restrict__ v, int64_t nnz, int* __restrict__ col, T* __restrict__ val) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    col[t] = (int)(keys[t] & 0xffffffffull);
    val[t] = v[perm ? perm[t] : t];
}

unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

int coo_to_csr(int value_bytes, int64_t nrows, int64_t ncols, int64_t nnz, bool sorted, const int* d_i,
               const int* d_j, const void* d_v, int* d_rowptr, int* d_col, void* d_val, hipStream_t s) {
    DeviceBuffer err;
    SPC5_HIP_TRY(err.alloc(sizeof(int)));
    SPC5_HIP_TRY(hipMemsetAsync(err.p, 0, sizeof(int), s));
    const size_t n = (size_t)std::max<int64_t>(nnz, 1);
    DeviceBuffer keys_a, keys_b, perm_a, perm_b;
    SPC5_HIP_TRY(keys_a.alloc(sizeof(unsigned long long) * n));
    SPC5_HIP_TRY(perm_a.alloc(sizeof(int) * n));
    unsigned long long* keys = keys_a.as<unsigned long long>();
    int* perm = perm_a.as<int>();
    if (nnz > 0) {
        hipLaunchKernelGGL(k_coo_keys, dim3(grid_of(nnz)), dim3(256), 0, s, d_i, d_j, nnz, nrows, ncols, keys, perm,
                           err.as<int>());
        SPC5_HIP_TRY(hipGetLastError());
    }
    if (sorted) {
        if (nnz > 1) {
            hipLaunchKernelGGL(k_check_sorted, dim3(grid_of(nnz - 1)), dim3(256), 0, s, keys, nnz, err.as<int>());
            SPC5_HIP_TRY(hipGetLastError());
        }
        perm = nullptr;
    } else if (nnz > 1) {
        SPC5_HIP_TRY(keys_b.alloc(sizeof(unsigned long long) * n));
        SPC5_HIP_TRY(perm_b.alloc(sizeof(int) * n));
        int end_bit = 32;
        while (end_bit < 64 && ((unsigned long long)nrows >> (end_bit - 32)) != 0) ++end_bit;
        size_t tmp_bytes = 0;
        SPC5_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_b.as<unsigned long long>(),
                                                        perm, perm_b.as<int>(), (int)nnz, 0, end_bit, s));
        DeviceBuffer tmp;
        SPC5_HIP_TRY(tmp.alloc(std::max<size_t>(tmp_bytes, 1)));
        SPC5_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys, keys_b.as<unsigned long long>(), perm,
                                                        perm_b.as<int>(), (int)nnz, 0, end_bit, s));
        keys = keys_b.as<unsigned long long>();
        perm = perm_b.as<int>();
        SPC5_HIP_TRY(hipStreamSynchronize(s));  // the sort's scratch is freed on return
    }
    hipLaunchKernelGGL(k_rowptr_search, dim3(grid_of(nrows + 1)), dim3(256), 0, s, keys, nnz, nrows, d_rowptr);
    SPC5_HIP_TRY(hipGetLastError());
    if (nnz > 0) {
        if (value_bytes == 8)
            hipLaunchKernelGGL(k_coo_gather<double>, dim3(grid_of(nnz)), dim3(256), 0, s, keys, perm,
                               reinterpret_cast<const double*>(d_v), nnz, d_col, reinterpret_cast<double*>(d_val));
        else
            hipLaunchKernelGGL(k_coo_gather<float>, dim3(grid_of(nnz)), dim3(256), 0, s, keys, perm,
                               reinterpret_cast<const float*>(d_v), nnz, d_col, reinterpret_cast<float*>(d_val));
        SPC5_HIP_TRY(hipGetLastError());
    }
    int herr = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    if (herr & 1) {
        set_error("COO index out of range (row not in [0, nrows) or column not in [0, ncols))");
        return SPC5_HIP_ERR_INVALID;
    }
    if (herr & 2) {
        set_error("COO entries not sorted by (row, column)");
        return SPC5_HIP_ERR_INVALID;
    }
    return SPC5_HIP_OK;
}

// ---- MatrixMarket ----------------------------------------------------------

struct MMFile {
    std::string text;
    size_t pos = 0;
    int64_t nrows = 0, ncols = 0, nstored = 0;
    bool symmetric = false, skew = false, pattern = false, integer = false;
};

bool read_file(const char* path, std::string& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    return got == out.size();
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

int mm_open(const char* path, MMFile& mm) {
    SPC5_CHECK_ARG(path != nullptr, "null path");
    if (!read_file(path, mm.text)) {
        set_error(std::string("cannot read ") + path);
        return SPC5_HIP_ERR_INVALID;
    }
    const std::string& t = mm.text;
    size_t eol = t.find('\n');
    const std::string head = lower(t.substr(0, eol));
    char w[5][64] = {};
    if (std::sscanf(head.c_str(), "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5 ||
        std::strcmp(w[0], "%%matrixmarket") != 0 || std::strcmp(w[1], "matrix") != 0 ||
        std::strcmp(w[2], "coordinate") != 0) {
        set_error("not a MatrixMarket coordinate matrix header");
        return SPC5_HIP_ERR_INVALID;
    }
    const std::string field = w[3], sym = w[4];
    if (field != "real" && field != "integer" && field != "pattern") {
        set_error("only real, integer and pattern MatrixMarket matrices are supported (" + field + " given)");
        return SPC5_HIP_ERR_INVALID;
    }
    if (sym != "general" && sym != "symmetric" && sym != "skew-symmetric") {
        set_error("unsupported MatrixMarket symmetry " + sym);
        return SPC5_HIP_ERR_INVALID;
    }
    mm.pattern = field == "pattern";
    mm.integer = field == "integer";
    mm.symmetric = sym == "symmetric";
    mm.skew = sym == "skew-symmetric";
    size_t p = eol == std::string::npos ? t.size() : eol + 1;
    while (p < t.size() && t[p] == '%') {  // comment lines
        const size_t e = t.find('\n', p);
        p = e == std::string::npos ? t.size() : e + 1;
    }
    char* end = nullptr;
    const char* c = t.c_str() + p;
    mm.nrows = std::strtoll(c, &end, 10);
    mm.ncols = std::strtoll(end, &end, 10);
    mm.nstored = std::strtoll(end, &end, 10);
    if (end == c || mm.nrows < 0 || mm.ncols < 0 || mm.nstored < 0) {
        set_error("bad MatrixMarket size line");
        return SPC5_HIP_ERR_INVALID;
    }
    mm.pos = (size_t)(end - t.c_str());
    return SPC5_HIP_OK;
}

// Walks the stored entries; emit(i, j, v) for each entry of the matrix
// (mirrored entries right after their stored one).  Returns the count, or -1
// on a malformed or out-of-range entry.
template <class Emit>
int64_t mm_walk(const MMFile& mm, Emit&& emit) {
    const char* c = mm.text.c_str() + mm.pos;
    int64_t count = 0;
    for (int64_t k = 0; k < mm.nstored; ++k) {
        char* end = nullptr;
        const long long row = std::strtoll(c, &end, 10);
        if (end == c) return -1;
        c = end;
        const long long col = std::strtoll(c, &end, 10);
        if (end == c) return -1;
        c = end;
        double v = 1.0;
        if (!mm.pattern) {
            v = mm.integer ? (double)std::strtoll(c, &end, 10) : std::strtod(c, &end);
            if (end == c) return -1;
            c = end;
        }
        if (row < 1 || row > mm.nrows || col < 1 || col > mm.ncols) return -1;
        const int i = (int)(row - 1), j = (int)(col - 1);
        emit(count++, i, j, v);
        if ((mm.symmetric || mm.skew) && i != j) emit(count++, j, i, mm.skew ? -v : v);
    }
    return count;
}

}  // namespace

extern "C" {

int spc5_hip_coo_to_csr(int value_bytes, int64_t nrows, int64_t ncols, int64_t nnz, int sorted, const int* d_i,
                        const int* d_j, const void* d_v, int* d_rowptr, int* d_col, void* d_val, void* stream) {
    SPC5_CHECK_ARG(value_bytes == 8 || value_bytes == 4, "value_bytes must be 8 or 4");
    SPC5_CHECK_ARG(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative size");
    SPC5_CHECK_ARG(nrows < ((int64_t)1 << 31) && ncols < ((int64_t)1 << 31) && nnz < ((int64_t)1 << 31),
                   "sizes must fit the 32-bit CSR arrays");
    SPC5_CHECK_ARG(d_rowptr != nullptr, "null rowptr");
    SPC5_CHECK_ARG(nnz == 0 || (d_i && d_j && d_v && d_col && d_val), "null COO or CSR array");
    return coo_to_csr(value_bytes, nrows, ncols, nnz, sorted != 0, d_i, d_j, d_v, d_rowptr, d_col, d_val,
                      as_stream(stream));
}

int spc5_hip_mm_info(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nentries) {
    MMFile mm;
    const int rc = mm_open(path, mm);
    if (rc) return rc;
    const int64_t n = mm_walk(mm, [](int64_t, int, int, double) {});
    if (n < 0) {
        set_error("malformed or out-of-range MatrixMarket entry");
        return SPC5_HIP_ERR_INVALID;
    }
    if (nrows) *nrows = mm.nrows;
    if (ncols) *ncols = mm.ncols;
    if (nentries) *nentries = n;
    return SPC5_HIP_OK;
}

int spc5_hip_mm_read(const char* path, int64_t capacity, int* ri, int* cj, double* v, int64_t* nread) {
    SPC5_CHECK_ARG(capacity == 0 || (ri && cj && v), "null output array");
    MMFile mm;
    const int rc = mm_open(path, mm);
    if (rc) return rc;
    bool overflow = false;
    const int64_t n = mm_walk(mm, [&](int64_t k, int i, int j, double x) {
        if (k >= capacity) {
            overflow = true;
            return;
        }
        ri[k] = i;
        cj[k] = j;
        v[k] = x;
    });
    if (n < 0) {
        set_error("malformed or out-of-range MatrixMarket entry");
        return SPC5_HIP_ERR_INVALID;
    }
    if (overflow) {
        set_error("capacity below the number of entries (see spc5_hip_mm_info)");
        return SPC5_HIP_ERR_INVALID;
    }
    if (nread) *nread = n;
    return SPC5_HIP_OK;
}

}  // extern "C"
