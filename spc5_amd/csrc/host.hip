// host.hip -- host-pointer entry points used by the C++ drop-in header
// (include/spc5_hip.hpp): conversion of a host CSR matrix on the GPU, and
// device-resident matrices created from host arrays.  They let a program
// written against the reference's spc5.hpp use the GPU without including
// any HIP header.
#include "spc5_mat.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

using namespace spc5;

namespace {

struct HostCsrOnDevice {
    DeviceBuffer rowptr, col, val;
    int upload(int64_t nrows, int64_t nnz, const int* rp, const int* c, const void* v, size_t vbytes) {
        SPC5_HIP_TRY(rowptr.alloc(sizeof(int) * (size_t)(nrows + 1)));
        SPC5_HIP_TRY(col.alloc(sizeof(int) * (size_t)nnz));
        SPC5_HIP_TRY(hipMemcpy(rowptr.p, rp, sizeof(int) * (size_t)(nrows + 1), hipMemcpyHostToDevice));
        if (nnz > 0) SPC5_HIP_TRY(hipMemcpy(col.p, c, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice));
        if (v) {
            SPC5_HIP_TRY(val.alloc(vbytes * (size_t)nnz));
            if (nnz > 0) SPC5_HIP_TRY(hipMemcpy(val.p, v, vbytes * (size_t)nnz, hipMemcpyHostToDevice));
        }
        return SPC5_HIP_OK;
    }
};

int check_host_csr(int64_t nrows, int64_t nnz, const int* rp) {
    SPC5_CHECK_ARG(nrows >= 0 && nrows <= INT32_MAX && nnz >= 0 && nnz <= INT32_MAX, "sizes out of range");
    SPC5_CHECK_ARG(rp != nullptr, "rowptr is null");
    SPC5_CHECK_ARG(rp[0] == 0 && (int64_t)rp[nrows] == nnz, "rowptr must start at 0 and end at nnz");
    return SPC5_HIP_OK;
}

}  // namespace

extern "C" {

int spc5_hip_host_block_count(int vs, int r, int64_t nrows, int64_t nnz, const int* rowptr, const int* colidx,
                              int64_t* nblocks) {
    int rc = check_host_csr(nrows, nnz, rowptr);
    if (rc) return rc;
    HostCsrOnDevice d;
    rc = d.upload(nrows, nnz, rowptr, colidx, nullptr, 0);
    if (rc) return rc;
    DeviceBuffer rs;
    SPC5_HIP_TRY(rs.alloc(sizeof(int) * (size_t)((nrows + r - 1) / r + 1)));
    return spc5_hip_convert_count(vs, r, nrows, d.rowptr.as<int>(), d.col.as<int>(), rs.as<int>(), nblocks, nullptr);
}

int spc5_hip_host_convert(int value_bytes, int vs, int mask_bytes, int r, int64_t nrows, int64_t nnz,
                          const int* rowptr, const int* colidx, const void* values_in, int* rowsSize_out,
                          unsigned char* headers_out, int64_t headers_capacity, void* values_out,
                          int64_t* nblocks) {
    int rc = check_host_csr(nrows, nnz, rowptr);
    if (rc) return rc;
    SPC5_CHECK_ARG(value_bytes == 8 || value_bytes == 4 || value_bytes == 2, "value_bytes must be 8, 4 or 2");
    SPC5_CHECK_ARG(rowsSize_out != nullptr && nblocks != nullptr, "null output");
    HostCsrOnDevice d;
    rc = d.upload(nrows, nnz, rowptr, colidx, values_in, (size_t)value_bytes);
    if (rc) return rc;
    const int64_t nbr = (nrows + r - 1) / r;
    DeviceBuffer rs;
    SPC5_HIP_TRY(rs.alloc(sizeof(int) * (size_t)(nbr + 1)));
    rc = spc5_hip_convert_count(vs, r, nrows, d.rowptr.as<int>(), d.col.as<int>(), rs.as<int>(), nblocks, nullptr);
    if (rc) return rc;
    const int64_t hbytes = *nblocks * (4 + (int64_t)r * mask_bytes);
    if (hbytes > headers_capacity) {
        set_error("headers_capacity too small: need " + std::to_string(hbytes) + " bytes");
        return SPC5_HIP_ERR_INVALID;
    }
    DeviceBuffer hdr, vout;
    SPC5_HIP_TRY(hdr.alloc((size_t)hbytes));
    SPC5_HIP_TRY(vout.alloc((size_t)value_bytes * (size_t)nnz));
    rc = spc5_hip_convert_fill(value_bytes, vs, mask_bytes, r, nrows, d.rowptr.as<int>(), d.col.as<int>(), d.val.p,
                               rs.as<int>(), hdr.as<unsigned char>(), vout.p, nullptr);
    if (rc) return rc;
    SPC5_HIP_TRY(hipMemcpy(rowsSize_out, rs.p, sizeof(int) * (size_t)(nbr + 1), hipMemcpyDeviceToHost));
    if (hbytes > 0) SPC5_HIP_TRY(hipMemcpy(headers_out, hdr.p, (size_t)hbytes, hipMemcpyDeviceToHost));
    if (nnz > 0) SPC5_HIP_TRY(hipMemcpy(values_out, vout.p, (size_t)value_bytes * (size_t)nnz, hipMemcpyDeviceToHost));
    return SPC5_HIP_OK;
}

int spc5_hip_mat_create_from_host(int value_type, int vs, int r, int64_t nrows, int64_t ncols, int64_t nnz,
                                  int64_t nblocks, const int* rowsSize, const unsigned char* headers,
                                  const void* values, spc5_hip_mat_t* out) {
    SPC5_CHECK_ARG(out != nullptr, "out is null");
    *out = nullptr;
    SPC5_CHECK_ARG(r == 1 || r == 2 || r == 4 || r == 8, "r must be 1, 2, 4 or 8");
    const int mb = vs == 16 ? 2 : 1;
    SPC5_CHECK_ARG(valid_format(value_type, vs, mb), "unsupported (value type, VS) combination");
    SPC5_CHECK_ARG(nrows >= 0 && nnz >= 0 && nblocks >= 0, "negative size");
    const int64_t nbr = (nrows + r - 1) / r;
    auto rs = std::make_shared<DeviceBuffer>();
    auto hdr = std::make_shared<DeviceBuffer>();
    auto val = std::make_shared<DeviceBuffer>();
    const size_t hb = (size_t)nblocks * (size_t)header_bytes(r, mb);
    const size_t vb = (size_t)nnz * value_bytes(value_type);
    SPC5_HIP_TRY(rs->alloc(sizeof(int) * (size_t)(nbr + 1)));
    SPC5_HIP_TRY(hdr->alloc(hb));
    SPC5_HIP_TRY(val->alloc(vb));
    // block pointers relative to the first one (only differences are meaningful)
    std::vector<int> bp((size_t)nbr + 1);
    for (int64_t i = 0; i <= nbr; ++i) bp[(size_t)i] = rowsSize[i] - rowsSize[0];
    SPC5_HIP_TRY(hipMemcpy(rs->p, bp.data(), sizeof(int) * (size_t)(nbr + 1), hipMemcpyHostToDevice));
    if (hb) SPC5_HIP_TRY(hipMemcpy(hdr->p, headers, hb, hipMemcpyHostToDevice));
    if (vb) SPC5_HIP_TRY(hipMemcpy(val->p, values, vb, hipMemcpyHostToDevice));
    spc5_hip_mat_t h = nullptr;
    int rc = spc5_hip_mat_create(value_type, vs, r, nrows, ncols, nnz, nblocks, rs->as<int>(),
                                 hdr->as<unsigned char>(), val->p, nullptr, &h);
    if (rc) return rc;
    h->owned.push_back(rs);
    h->owned.push_back(hdr);
    h->owned.push_back(val);
    *out = h;
    return SPC5_HIP_OK;
}

int spc5_hip_spmv_host(spc5_hip_mat_t h, const void* x, void* y) {
    SPC5_CHECK_ARG(h != nullptr, "null handle");
    if (h->nrows == 0) return SPC5_HIP_OK;
    SPC5_CHECK_ARG(x != nullptr && y != nullptr, "x or y is null");
    const size_t es = h->value_type == SPC5_HIP_F64 ? 8 : 4;
    DeviceBuffer dx, dy;
    SPC5_HIP_TRY(dx.alloc(es * (size_t)h->ncols));
    SPC5_HIP_TRY(dy.alloc(es * (size_t)h->nrows));
    SPC5_HIP_TRY(hipMemcpy(dx.p, x, es * (size_t)h->ncols, hipMemcpyHostToDevice));
    SPC5_HIP_TRY(hipMemcpy(dy.p, y, es * (size_t)h->nrows, hipMemcpyHostToDevice));
    int rc = spc5_hip_spmv(h, dx.p, dy.p, nullptr);
    if (rc) return rc;
    SPC5_HIP_TRY(hipMemcpy(y, dy.p, es * (size_t)h->nrows, hipMemcpyDeviceToHost));
    return SPC5_HIP_OK;
}

// ---- serialisation ----------------------------------------------------------
// File layout, little-endian:
//   char   magic[8] = "SPC5BRVC"; uint32 version = 1;
//   int32  value_type, vs, mask_bytes, r, reserved (0);
//   int64  nrows, ncols, nnz, nblocks;
//   int32  rowsSize[ceil(nrows/r) + 1]     (block-row pointers, first 0)
//   uint8  headers[nblocks * (4 + r * mask_bytes)]
//   values[nnz]                            (8, 4 or 2 bytes each)

int spc5_hip_mat_save(spc5_hip_mat_t h, const char* path) {
    SPC5_CHECK_ARG(h != nullptr && path != nullptr, "null handle or path");
    const int64_t nbr = h->nbr;
    std::vector<int> bp((size_t)nbr + 1);
    std::vector<unsigned char> hdr((size_t)h->nblocks * (size_t)header_bytes(h->r, h->mb));
    std::vector<unsigned char> val((size_t)h->nnz * value_bytes(h->value_type));
    SPC5_HIP_TRY(hipDeviceSynchronize());
    SPC5_HIP_TRY(hipMemcpy(bp.data(), h->bptr, sizeof(int) * bp.size(), hipMemcpyDeviceToHost));
    if (!hdr.empty()) SPC5_HIP_TRY(hipMemcpy(hdr.data(), h->hdr, hdr.size(), hipMemcpyDeviceToHost));
    if (!val.empty()) SPC5_HIP_TRY(hipMemcpy(val.data(), h->values, val.size(), hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error(std::string("cannot write ") + path);
        return SPC5_HIP_ERR_INVALID;
    }
    const uint32_t version = 1;
    const int32_t meta[5] = {h->value_type, h->vs, h->mb, h->r, 0};
    const int64_t dims[4] = {h->nrows, h->ncols, h->nnz, h->nblocks};
    bool ok = std::fwrite("SPC5BRVC", 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 &&
              std::fwrite(meta, 4, 5, f) == 5 && std::fwrite(dims, 8, 4, f) == 4 &&
              std::fwrite(bp.data(), 4, bp.size(), f) == bp.size() &&
              std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
              std::fwrite(val.data(), 1, val.size(), f) == val.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        set_error(std::string("short write to ") + path);
        return SPC5_HIP_ERR_INVALID;
    }
    return SPC5_HIP_OK;
}

int spc5_hip_mat_load(const char* path, spc5_hip_mat_t* out) {
    SPC5_CHECK_ARG(path != nullptr && out != nullptr, "null path or out");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string("cannot read ") + path);
        return SPC5_HIP_ERR_INVALID;
    }
    char magic[8];
    uint32_t version = 0;
    int32_t meta[5];
    int64_t dims[4];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SPC5BRVC", 8) == 0 &&
              std::fread(&version, 4, 1, f) == 1 && version == 1 && std::fread(meta, 4, 5, f) == 5 &&
              std::fread(dims, 8, 4, f) == 4;
    const int vt = meta[0], vs = meta[1], mb = meta[2], r = meta[3];
    ok = ok && (r == 1 || r == 2 || r == 4 || r == 8) && valid_format(vt, vs, mb) && dims[0] >= 0 && dims[1] >= 0 &&
         dims[2] >= 0 && dims[3] >= 0 && dims[2] <= INT32_MAX && dims[3] <= INT32_MAX && dims[0] <= INT32_MAX;
    if (!ok) {
        std::fclose(f);
        set_error(std::string("not a beta(r,VS) matrix file: ") + path);
        return SPC5_HIP_ERR_INVALID;
    }
    const int64_t nbr = (dims[0] + r - 1) / r;
    std::vector<int> bp((size_t)nbr + 1);
    std::vector<unsigned char> hdr((size_t)dims[3] * (size_t)header_bytes(r, mb));
    std::vector<unsigned char> val((size_t)dims[2] * value_bytes(vt));
    ok = std::fread(bp.data(), 4, bp.size(), f) == bp.size() && std::fread(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
         std::fread(val.data(), 1, val.size(), f) == val.size();
    std::fclose(f);
    if (!ok) {
        set_error(std::string("truncated matrix file: ") + path);
        return SPC5_HIP_ERR_INVALID;
    }
    return spc5_hip_mat_create_from_host(vt, vs, r, dims[0], dims[1], dims[2], dims[3], bp.data(), hdr.data(),
                                         val.data(), out);
}

}  // extern "C"
