// spmm.hip -- packing of the dose-loading matrix and its voxel-major copy for D^T X products.
//
// Host side of kernels_spmm.hip.  The ensemble keeps one CSR per beam with local row offsets
// (MatCSREnsemble, IMRTOpt/src/IMRTSparseMatEns.cu:10-336); both operations first build a
// global row pointer over all matrices (rows in matrix order, entries already contiguous).
#include <hip/hip_runtime.h>

#include <vector>

#include "engine.h"

using namespace fdi;

struct fd_spmm {
    int device = 0;
    int vtype = FD_VAL_F32;
    uint64_t rows = 0, cols = 0, nnz = 0;
    DevBuf<uint64_t> colptr;  // cols + 1
    DevBuf<uint32_t> crow;    // nnz, row of each entry, ascending within a column
    DevBuf<uint8_t> cval;     // nnz values (4 or 2 bytes each)
    DevBuf<float> scale;      // rows (16-bit types)
};

namespace {

// Global row pointer of the ensemble (R + 1 entries) into rowptr.
int build_rowptr(const fd_dosemat* m, DevBuf<uint64_t>& rowptr, uint64_t* R_out, hipStream_t s) {
    const int nm = (int)m->rows.size();
    std::vector<uint64_t> first((size_t)nm + 1, 0);
    for (int k = 0; k < nm; k++) first[k + 1] = first[k] + m->rows[k];
    const uint64_t R = first[nm];
    *R_out = R;
    FD_TRY(rowptr.reserve(R + 1));
    if (nm == 0) {
        FD_HIP(launch_fill_u64(rowptr.ptr, 1, 0, s));
        return FD_OK;
    }
    DevBuf<uint64_t> table;  // row_first[nm + 1], off_off[nm], nnz_off[nm]
    FD_TRY(table.reserve(3 * (size_t)nm + 1));
    std::vector<uint64_t> host(3 * (size_t)nm + 1);
    std::copy(first.begin(), first.end(), host.begin());
    std::copy(m->off_off.begin(), m->off_off.end(), host.begin() + nm + 1);
    std::copy(m->nnz_off.begin(), m->nnz_off.end(), host.begin() + 2 * nm + 1);
    FD_HIP(hipMemcpyAsync(table.ptr, host.data(), host.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    hipError_t e = launch_global_rows(nm, table.ptr, table.ptr + nm + 1, table.ptr + 2 * nm + 1, m->offsets.ptr, R,
                                      m->used_nnz, rowptr.ptr, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    table.release();
    if (e != hipSuccess) {
        set_error(std::string("global rows: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

bool valid_type(int t) { return t == FD_VAL_F32 || t == FD_VAL_BF16 || t == FD_VAL_FP16; }

}  // namespace

extern "C" {

int fd_dosemat_pack(fd_dosemat* m, int value_type, void* stream) {
    if (!m || (value_type != FD_VAL_BF16 && value_type != FD_VAL_FP16)) {
        set_error("null matrix or value type not FD_VAL_BF16 / FD_VAL_FP16");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    DevBuf<uint64_t> rowptr;
    uint64_t R = 0;
    int r = build_rowptr(m, rowptr, &R, s);
    if (r == FD_OK) r = m->values16.reserve(std::max<uint64_t>(m->used_nnz, 1));
    if (r == FD_OK) r = m->row_scale.reserve(std::max<uint64_t>(R, 1));
    if (r == FD_OK) {
        hipError_t e = launch_pack_rows(R, rowptr.ptr, m->values.ptr, value_type, m->row_scale.ptr, m->values16.ptr, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error(std::string("pack: ") + hipGetErrorString(e));
            r = FD_ERR_HIP;
        }
    }
    rowptr.release();
    if (r == FD_OK) m->packed = value_type;
    return r;
}

int fd_dosemat_copy_packed(const fd_dosemat* m, float* scale, uint16_t* values) {
    if (!m || m->packed < 0) {
        set_error("matrix not packed (fd_dosemat_pack)");
        return FD_ERR_ARG;
    }
    FD_HIP(hipDeviceSynchronize());
    const uint64_t R = m->total_rows();
    if (scale && R) FD_HIP(hipMemcpy(scale, m->row_scale.ptr, R * sizeof(float), hipMemcpyDeviceToHost));
    if (values && m->used_nnz)
        FD_HIP(hipMemcpy(values, m->values16.ptr, m->used_nnz * sizeof(uint16_t), hipMemcpyDeviceToHost));
    return FD_OK;
}

int fd_spmm_create(const fd_dosemat* m, int value_type, fd_spmm** out, void* stream) {
    if (!m || !out || !valid_type(value_type)) {
        set_error("null argument or unknown value type");
        return FD_ERR_ARG;
    }
    *out = nullptr;
    if (m->n_cols >= (1ull << 40)) {
        set_error("too many columns");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    fd_spmm* p = new fd_spmm();
    FD_HIP(hipGetDevice(&p->device));
    p->vtype = value_type;
    p->cols = m->n_cols;
    p->nnz = m->used_nnz;
    const int vbytes = value_type == FD_VAL_F32 ? 4 : 2;
    DevBuf<uint64_t> rowptr;
    DevBuf<uint32_t> count;
    DevBuf<uint16_t> q;
    auto fail = [&](int r) {
        rowptr.release();
        count.release();
        q.release();
        fd_spmm_destroy(p);
        return r;
    };
    int r = build_rowptr(m, rowptr, &p->rows, s);
    if (r != FD_OK) return fail(r);
    if (p->rows >= (1ull << 32)) {
        set_error("more than 2^32 rows");
        return fail(FD_ERR_ARG);
    }
    if ((r = p->colptr.reserve(p->cols + 1)) != FD_OK || (r = p->crow.reserve(std::max<uint64_t>(p->nnz, 1))) != FD_OK ||
        (r = p->cval.reserve(std::max<uint64_t>(p->nnz, 1) * vbytes)) != FD_OK ||
        (r = p->scale.reserve(std::max<uint64_t>(p->rows, 1))) != FD_OK ||
        (r = count.reserve(p->cols)) != FD_OK)
        return fail(r);
    const void* src = m->values.ptr;
    if (vbytes == 2) {
        if ((r = q.reserve(std::max<uint64_t>(p->nnz, 1))) != FD_OK) return fail(r);
        src = q.ptr;
    }
    hipError_t e = launch_pack_rows(p->rows, rowptr.ptr, m->values.ptr, value_type, p->scale.ptr,
                                    vbytes == 2 ? q.ptr : nullptr, s);
    if (e == hipSuccess) e = launch_fill_u32(count.ptr, p->cols, 0u, s);
    if (e == hipSuccess) e = launch_col_count(p->nnz, m->columns.ptr, count.ptr, s);
    DevBuf<uint64_t> scan_tmp;
    if (e == hipSuccess && scan_tmp.reserve(scan_scratch_size(p->cols)) != FD_OK) e = hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = launch_scan_u32(count.ptr, p->cols, p->colptr.ptr, p->colptr.ptr + p->cols, scan_tmp.ptr, s);
    if (e == hipSuccess) e = launch_fill_u32(count.ptr, p->cols, 0u, s);
    if (e == hipSuccess)
        e = launch_col_scatter(p->rows, rowptr.ptr, m->columns.ptr, src, vbytes, p->colptr.ptr, count.ptr,
                               p->crow.ptr, p->cval.ptr, s);
    if (e == hipSuccess) e = launch_col_sort(p->cols, p->colptr.ptr, p->crow.ptr, p->cval.ptr, vbytes, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    scan_tmp.release();
    if (e != hipSuccess) {
        set_error(std::string("spmm build: ") + hipGetErrorString(e));
        return fail(FD_ERR_HIP);
    }
    rowptr.release();
    count.release();
    q.release();
    *out = p;
    return FD_OK;
}

int fd_spmm_info(const fd_spmm* p, uint64_t* rows, uint64_t* cols, uint64_t* nnz) {
    if (!p) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (rows) *rows = p->rows;
    if (cols) *cols = p->cols;
    if (nnz) *nnz = p->nnz;
    return FD_OK;
}

int fd_spmm_run(fd_spmm* p, const float* x_dev, int k, float* y_dev, void* stream) {
    if (!p || k <= 0 || !y_dev || (p->rows && !x_dev)) {
        set_error("null argument or k <= 0");
        return FD_ERR_ARG;
    }
    // k % 4 == 0 takes 16-byte loads and stores
    if (k % 4 == 0 && ((uintptr_t)x_dev % 16 || (uintptr_t)y_dev % 16)) {
        set_error("x_dev and y_dev must be 16-byte aligned");
        return FD_ERR_ARG;
    }
    FD_HIP(hipSetDevice(p->device));
    hipError_t e = launch_spmm(p->vtype, p->cols, p->colptr.ptr, p->crow.ptr, p->cval.ptr, p->scale.ptr, x_dev, k,
                               y_dev, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error(std::string("spmm: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

int fd_spmm_destroy(fd_spmm* p) {
    if (!p) return FD_OK;
    p->colptr.release();
    p->crow.release();
    p->cval.release();
    p->scale.release();
    delete p;
    return FD_OK;
}

}  // extern "C"
