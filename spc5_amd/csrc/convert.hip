// convert.hip -- CSR -> beta(r,VS) on the GPU, byte-identical to the
// reference converter (src/spc5.hpp:165-228 and 455-488).
//
// The reference walks each block-row with one cursor per row: a block starts
// at the smallest unconsumed column of the block-row and takes, row by row,
// every entry with column < start+VS.  Block-rows are independent, so the
// GPU runs one lane per block-row in two passes:
//   count : blocks per block-row -> exclusive scan -> block-row pointers
//           (the reference's rowsSize after conversion) and the block count;
//   fill  : the same walk again, writing each block's packed header at
//           (pointer + k) * (4 + r*MB) bytes and its values at the block-row's
//           value offset, which is the CSR row pointer of its first row
//           (conversion only permutes values inside a block-row).
#include "spc5_common.hpp"

#include <mutex>

#include <climits>

namespace spc5 {

namespace {

// Checks a CSR row pointer: starts at 0, non-decreasing.
__global__ void k_check_rowptr(const int* __restrict__ rowptr, int64_t nrows, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nrows) return;
    bool bad = false;
    if (i == 0) bad = rowptr[0] != 0;
    if (i < nrows) bad |= rowptr[i + 1] < rowptr[i];
    if (bad) atomicOr(err, 1);
}

template <int R>
struct Cursors {
    int cur[R];
    int end[R];
    int head[R];  // column at the cursor, INT_MAX when the row is consumed

    __device__ __forceinline__ void init(int64_t row0, int64_t nrows, const int* rowptr, const int* col) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (row0 + s < nrows) {
                cur[s] = rowptr[row0 + s];
                end[s] = rowptr[row0 + s + 1];
            } else {
                cur[s] = end[s] = 0;
            }
            head[s] = cur[s] < end[s] ? col[cur[s]] : INT_MAX;
        }
    }
    // Smallest unconsumed column; false when every row is consumed.
    __device__ __forceinline__ bool next_start(int& start) const {
        bool any = false;
        start = INT_MAX;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (cur[s] < end[s]) {
                any = true;
                start = head[s] < start ? head[s] : start;
            }
        }
        return any;
    }
};

// Blocks per block-row.  Also checks the precondition of the format
// (src/spc5.hpp:165-228 assumes it without a check): columns are
// non-negative and strictly increasing inside every row.  A violation sets
// bit 0 of *err; the fill pass is then never run.
template <int R, int VS>
__global__ void k_convert_count(int64_t nrows, int64_t nbr, const int* __restrict__ rowptr,
                                const int* __restrict__ col, int* __restrict__ cnt_out, int* __restrict__ err) {
    const int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= nbr) return;
    Cursors<R> c;
    c.init(br * R, nrows, rowptr, col);
    bool bad = false;
#pragma unroll
    for (int s = 0; s < R; ++s) bad |= c.head[s] < 0;
    int nb = 0;
    int start;
    while (c.next_start(start)) {
        ++nb;
        const int64_t lim = (int64_t)start + VS;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            while (c.cur[s] < c.end[s] && c.head[s] < lim) {
                const int prev = c.head[s];
                ++c.cur[s];
                c.head[s] = c.cur[s] < c.end[s] ? col[c.cur[s]] : INT_MAX;
                bad |= c.cur[s] < c.end[s] && c.head[s] <= prev;
            }
        }
    }
    cnt_out[br] = nb;
    if (bad) atomicOr(err, 1);
}

template <class W, int R, int VS, int MB>
__global__ void k_convert_fill(int64_t nrows, int64_t nbr, const int* __restrict__ rowptr,
                               const int* __restrict__ col, const W* __restrict__ vin,
                               const int* __restrict__ bptr, int64_t nblocks, unsigned char* __restrict__ hdr,
                               W* __restrict__ vout) {
    constexpr int H = 4 + R * MB;
    constexpr int NW = (R * MB + 7) / 8;  // 64-bit words of the combined mask
    const int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= nbr) return;
    const int64_t row0 = br * R;
    Cursors<R> c;
    c.init(row0, nrows, rowptr, col);
    int64_t bo = bptr[br];
    const int64_t bend = bptr[br + 1] < nblocks ? bptr[br + 1] : nblocks;
    int64_t vo = rowptr[row0];
    int start;
    while (c.next_start(start)) {
        if (bo >= bend) break;  // inconsistent pointers: never write past the block-row
        const int64_t lim = (int64_t)start + VS;
        unsigned long long mw[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) mw[k] = 0ull;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            while (c.cur[s] < c.end[s] && c.head[s] < lim) {
                // In range for sorted columns (checked by the count pass); the
                // guard only keeps a direct call on unsorted input in bounds.
                const unsigned off = (unsigned)(c.head[s] - start);
                const int bit = s * VS + (int)(off < (unsigned)VS ? off : 0u);
                mw[bit >> 6] |= 1ull << (bit & 63);
                vout[vo++] = vin[c.cur[s]];
                ++c.cur[s];
                c.head[s] = c.cur[s] < c.end[s] ? col[c.cur[s]] : INT_MAX;
            }
        }
        unsigned char* h = hdr + bo * H;
        __builtin_memcpy(h, &start, 4);
        __builtin_memcpy(h + 4, mw, R * MB);
        ++bo;
    }
}

template <int R, int VS>
void launch_count(int64_t nrows, int64_t nbr, const int* rowptr, const int* col, int* out, int* err,
                  hipStream_t s) {
    const unsigned g = (unsigned)((nbr + 255) / 256);
    hipLaunchKernelGGL((k_convert_count<R, VS>), dim3(g), dim3(256), 0, s, nrows, nbr, rowptr, col, out, err);
}

template <class W, int VS, int MB>
int launch_fill_r(int r, int64_t nrows, int64_t nbr, const int* rowptr, const int* col, const void* vin,
                  const int* bptr, int64_t nblocks, unsigned char* hdr, void* vout, hipStream_t s) {
    const unsigned g = (unsigned)((nbr + 255) / 256);
    const W* vi = reinterpret_cast<const W*>(vin);
    W* vo = reinterpret_cast<W*>(vout);
    switch (r) {
    case 1:
        hipLaunchKernelGGL((k_convert_fill<W, 1, VS, MB>), dim3(g), dim3(256), 0, s, nrows, nbr, rowptr, col, vi, bptr,
                           nblocks, hdr, vo);
        break;
    case 2:
        hipLaunchKernelGGL((k_convert_fill<W, 2, VS, MB>), dim3(g), dim3(256), 0, s, nrows, nbr, rowptr, col, vi, bptr,
                           nblocks, hdr, vo);
        break;
    case 4:
        hipLaunchKernelGGL((k_convert_fill<W, 4, VS, MB>), dim3(g), dim3(256), 0, s, nrows, nbr, rowptr, col, vi, bptr,
                           nblocks, hdr, vo);
        break;
    case 8:
        hipLaunchKernelGGL((k_convert_fill<W, 8, VS, MB>), dim3(g), dim3(256), 0, s, nrows, nbr, rowptr, col, vi, bptr,
                           nblocks, hdr, vo);
        break;
    default: set_error("r must be 1, 2, 4 or 8"); return SPC5_HIP_ERR_INVALID;
    }
    SPC5_HIP_TRY(hipGetLastError());
    return SPC5_HIP_OK;
}

template <class W>
int launch_fill_w(int vs, int mb, int r, int64_t nrows, int64_t nbr, const int* rowptr, const int* col,
                  const void* vin, const int* bptr, int64_t nblocks, unsigned char* hdr, void* vout, hipStream_t s) {
    if (vs == 8 && mb == 1) return launch_fill_r<W, 8, 1>(r, nrows, nbr, rowptr, col, vin, bptr, nblocks, hdr, vout, s);
    if (vs == 16 && mb == 2)
        return launch_fill_r<W, 16, 2>(r, nrows, nbr, rowptr, col, vin, bptr, nblocks, hdr, vout, s);
    set_error("(VS, mask bytes) must be (8, 1) or (16, 2)");
    return SPC5_HIP_ERR_INVALID;
}

// Scratch of one count pass: words[0] flags a bad row pointer, words[1]
// unsorted or negative columns, words[2] receives the block total; the scan
// of the block counts uses partials[0 .. partial_cap).
struct CountScratch {
    int* words = nullptr;
    int* partials = nullptr;
    int64_t partial_cap = 0;
};

// Fixed device scratch of spc5_hip_convert_count_noalloc, one copy per GPU
// (a file-scope __device__ array): kScratchWords flag/total words, then the
// scan partials.  Host code uses the address hipGetSymbolAddress returns on
// each device, never the array's name.
constexpr int kScratchInts = 16384;  // 64 KiB
constexpr int kScratchWords = 64;    // keeps the partials 256-byte aligned
constexpr int kMaxDevices = 64;
__device__ __align__(256) int scratch[kScratchInts];

std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (looked up once
// per device, with that device current).
int scratch_address(int** out) {
    int dev = -1;
    SPC5_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index out of range for the fixed scratch table");
        return SPC5_HIP_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        SPC5_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) {
            set_error("hipGetSymbolAddress returned no address for the scratch array");
            return SPC5_HIP_ERR_HIP;
        }
        g_scratch_ptr[dev] = reinterpret_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SPC5_HIP_OK;
}

// Reads n flag/total words back (synchronises the stream).
int read_words(const int* d_words, int* h, int n, hipStream_t s) {
    SPC5_HIP_TRY(hipMemcpyAsync(h, d_words, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return SPC5_HIP_OK;
}

// Count pass shared by spc5_hip_convert_count and its no-allocation variant:
// row pointer check, block count per block-row, exclusive scan into
// d_rowsSize_out[0..nbr], total into *nblocks.  Allocates nothing itself.
int convert_count_impl(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx, int* d_rowsSize_out,
                       int64_t* nblocks, const CountScratch& ws, hipStream_t s) {
    *nblocks = 0;
    const int64_t nbr = (nrows + r - 1) / r;
    if (scan_partials_needed(nbr + 1) > ws.partial_cap) {
        set_error("too many block-rows for the scratch space (" + std::to_string(nbr) + " > " +
                  std::to_string(ws.partial_cap * 2048 - 1) + ")");
        return SPC5_HIP_ERR_INVALID;
    }
    // flags and total start at zero: a fill kernel, not a memset node
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ws.words, (int64_t)3, 0);
    SPC5_HIP_TRY(hipGetLastError());
    {
        const unsigned g = (unsigned)((nrows + 1 + 255) / 256);
        hipLaunchKernelGGL(k_check_rowptr, dim3(g), dim3(256), 0, s, d_rowptr, nrows, ws.words);
        SPC5_HIP_TRY(hipGetLastError());
        int h = 0;
        int rc = read_words(ws.words, &h, 1, s);
        if (rc) return rc;
        // the count pass reads columns at the row pointers: only once they are valid
        if (h) {
            set_error("CSR row pointer must start at 0 and be non-decreasing");
            return SPC5_HIP_ERR_INVALID;
        }
    }
    if (nbr > 0 && d_colidx == nullptr) {
        // A matrix without entries may come without a column array (an
        // empty tensor's storage); the count pass then reads no column.
        int nnz_h = 0;
        SPC5_HIP_TRY(hipMemcpyAsync(&nnz_h, d_rowptr + nrows, sizeof(int), hipMemcpyDeviceToHost, s));
        SPC5_HIP_TRY(hipStreamSynchronize(s));
        SPC5_CHECK_ARG(nnz_h == 0, "colidx is null");
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, d_rowsSize_out + nbr, (int64_t)1, 0);
    SPC5_HIP_TRY(hipGetLastError());
    if (nbr > 0) {
        const int key = r * 100 + vs;
        switch (key) {
        case 108: launch_count<1, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 208: launch_count<2, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 408: launch_count<4, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 808: launch_count<8, 8>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 116: launch_count<1, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 216: launch_count<2, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 416: launch_count<4, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        case 816: launch_count<8, 16>(nrows, nbr, d_rowptr, d_colidx, d_rowsSize_out, ws.words + 1, s); break;
        }
        SPC5_HIP_TRY(hipGetLastError());
    }
    int rc = scan_exclusive_i32_ws(d_rowsSize_out, nbr + 1, ws.words + 2, ws.partials, s);
    if (rc) return rc;
    int h[3] = {0, 0, 0};
    rc = read_words(ws.words, h, 3, s);
    if (rc) return rc;
    if (h[1]) {
        set_error("CSR columns must be non-negative and strictly increasing inside each row");
        return SPC5_HIP_ERR_INVALID;
    }
    *nblocks = h[2];
    return SPC5_HIP_OK;
}

int check_convert_count_args(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_rowsSize_out,
                             const int64_t* nblocks) {
    SPC5_CHECK_ARG(nblocks != nullptr, "nblocks is null");
    SPC5_CHECK_ARG(r == 1 || r == 2 || r == 4 || r == 8, "r must be 1, 2, 4 or 8");
    SPC5_CHECK_ARG(vs == 8 || vs == 16, "VS must be 8 or 16");
    SPC5_CHECK_ARG(nrows >= 0 && nrows <= INT32_MAX, "nrows out of range");
    SPC5_CHECK_ARG(d_rowptr != nullptr && d_rowsSize_out != nullptr, "null pointer");
    SPC5_CHECK_ARG((const void*)d_rowptr != (const void*)d_rowsSize_out, "rowsSize_out must not alias rowptr");
    return SPC5_HIP_OK;
}

}  // namespace

}  // namespace spc5

using namespace spc5;

extern "C" {

int spc5_hip_convert_count(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx,
                           int* d_rowsSize_out, int64_t* nblocks, void* stream) {
    int rc = check_convert_count_args(vs, r, nrows, d_rowptr, d_rowsSize_out, nblocks);
    if (rc) return rc;
    *nblocks = 0;
    hipStream_t s = as_stream(stream);
    const int64_t nbr = (nrows + r - 1) / r;
    CountScratch ws;
    ws.partial_cap = scan_partials_needed(nbr + 1);
    DeviceBuffer buf;
    SPC5_HIP_TRY(buf.alloc(sizeof(int) * (size_t)(kScratchWords + ws.partial_cap)));
    ws.words = buf.as<int>();
    ws.partials = ws.words + kScratchWords;
    rc = convert_count_impl(vs, r, nrows, d_rowptr, d_colidx, d_rowsSize_out, nblocks, ws, s);
    // the buffer is freed on return: every launch that uses it has ended
    // (each path out of the pass ends with a stream synchronisation)
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    return rc;
}

int spc5_hip_convert_count_noalloc(int vs, int r, int64_t nrows, const int* d_rowptr, const int* d_colidx,
                                   int* d_rowsSize_out, int64_t* nblocks, void* stream) {
    int rc = check_convert_count_args(vs, r, nrows, d_rowptr, d_rowsSize_out, nblocks);
    if (rc) return rc;
    *nblocks = 0;
    CountScratch ws;
    rc = scratch_address(&ws.words);
    if (rc) return rc;
    ws.partials = ws.words + kScratchWords;
    ws.partial_cap = kScratchInts - kScratchWords;
    return convert_count_impl(vs, r, nrows, d_rowptr, d_colidx, d_rowsSize_out, nblocks, ws, as_stream(stream));
}

int spc5_hip_convert_fill(int value_bytes, int vs, int mask_bytes, int r, int64_t nrows, const int* d_rowptr,
                          const int* d_colidx, const void* d_values_in, const int* d_rowsSize,
                          unsigned char* d_headers_out, void* d_values_out, void* stream) {
    SPC5_CHECK_ARG(value_bytes == 8 || value_bytes == 4 || value_bytes == 2, "value_bytes must be 8, 4 or 2");
    SPC5_CHECK_ARG(nrows >= 0 && nrows <= INT32_MAX, "nrows out of range");
    SPC5_CHECK_ARG(d_rowptr != nullptr && d_rowsSize != nullptr, "null pointer");
    hipStream_t s = as_stream(stream);
    const int64_t nbr = (nrows + r - 1) / r;
    if (nbr == 0) return SPC5_HIP_OK;
    int nb = 0;
    SPC5_HIP_TRY(hipMemcpyAsync(&nb, d_rowsSize + nbr, sizeof(int), hipMemcpyDeviceToHost, s));
    SPC5_HIP_TRY(hipStreamSynchronize(s));
    if (nb == 0) return SPC5_HIP_OK;
    SPC5_CHECK_ARG(d_headers_out != nullptr && d_values_out != nullptr && d_values_in != nullptr,
                   "null output/value pointer");
    switch (value_bytes) {
    case 8:
        return launch_fill_w<unsigned long long>(vs, mask_bytes, r, nrows, nbr, d_rowptr, d_colidx, d_values_in,
                                                 d_rowsSize, nb, d_headers_out, d_values_out, s);
    case 4:
        return launch_fill_w<unsigned int>(vs, mask_bytes, r, nrows, nbr, d_rowptr, d_colidx, d_values_in,
                                           d_rowsSize, nb, d_headers_out, d_values_out, s);
    default:
        return launch_fill_w<unsigned short>(vs, mask_bytes, r, nrows, nbr, d_rowptr, d_colidx, d_values_in,
                                             d_rowsSize, nb, d_headers_out, d_values_out, s);
    }
}

}  // extern "C"
