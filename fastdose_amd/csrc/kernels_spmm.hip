// kernels_spmm.hip -- 16-bit packing of the dose-loading matrix and D^T X products (gfx950).
//
// The optimiser of the reference multiplies the dose-loading matrix by fluence vectors
// (dose = D^T x, IMRTOpt/src/IMRTOptimize.cu:604-625 through cuSPARSE; IMRTDebug.cu:211-315).
// Here the ensemble of per-beam CSR matrices (rows = beamlets, columns = voxels) is:
//   - packed: per row a FP32 scale (the row maximum) and values q = v / scale rounded to
//     BF16 or FP16 (round to nearest even);
//   - transposed to a voxel-major copy (CSC): a column's entries sorted by row, so every product
//     sums in a fixed order and is bitwise reproducible;
//   - multiplied: one wavefront per voxel, lanes along the k right-hand sides (4 per lane,
//     16-byte loads of X rows and one coalesced 16-byte store per lane of Y), FP32 accumulation.
// The product is HBM-bound on writing Y (n_voxels x k floats); D and X are read once per column
// entry from L2.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fdi {
namespace {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // finite inputs only
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_fp16_rne(float f) {
    const _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float fp16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// rowptr[r] (global, r in [0, R]) from the ensemble's per-matrix local offsets.
__global__ void global_rows_kernel(int n_mat, const uint64_t* __restrict__ row_first,
                                   const uint64_t* __restrict__ off_off, const uint64_t* __restrict__ nnz_off,
                                   const uint64_t* __restrict__ offsets, uint64_t R, uint64_t total_nnz,
                                   uint64_t* __restrict__ rowptr) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > R) return;
    if (r == R) {
        rowptr[R] = total_nnz;
        return;
    }
    int lo = 0, hi = n_mat - 1;  // matrix k with row_first[k] <= r < row_first[k + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_first[mid] <= r) lo = mid; else hi = mid - 1;
    }
    rowptr[r] = nnz_off[lo] + offsets[off_off[lo] + (r - row_first[lo])];
}

// one wavefront per row: scale = row maximum; q = v / scale in 16 bits (scale 0 for an empty row)
__global__ __launch_bounds__(256) void pack_rows_kernel(uint64_t R, const uint64_t* __restrict__ rowptr,
                                                        const float* __restrict__ values, int vtype,
                                                        float* __restrict__ scale, uint16_t* __restrict__ q) {
    const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= R) return;
    const uint64_t a = rowptr[r], b = rowptr[r + 1];
    float m = 0.0f;
    for (uint64_t e = a + lane; e < b; e += 64) m = fmaxf(m, values[e]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) scale[r] = m;
    if (!q) return;
    for (uint64_t e = a + lane; e < b; e += 64) {
        const float t = values[e] / m;
        q[e] = vtype == FD_VAL_BF16 ? f32_to_bf16_rne(t) : f32_to_fp16_rne(t);
    }
}

__global__ void col_count_kernel(uint64_t nnz, const uint64_t* __restrict__ columns, uint32_t* __restrict__ count) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&count[columns[e]], 1u);
}

// one wavefront per row: entries to their column segment (order inside a column fixed later)
template <typename V>
__global__ __launch_bounds__(256) void col_scatter_kernel(uint64_t R, const uint64_t* __restrict__ rowptr,
                                                          const uint64_t* __restrict__ columns,
                                                          const V* __restrict__ vals,
                                                          const uint64_t* __restrict__ colptr,
                                                          uint32_t* __restrict__ cursor, uint32_t* __restrict__ crow,
                                                          V* __restrict__ cval) {
    const uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= R) return;
    for (uint64_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += 64) {
        const uint64_t c = columns[e];
        const uint64_t pos = colptr[c] + atomicAdd(&cursor[c], 1u);
        crow[pos] = (uint32_t)r;
        cval[pos] = vals[e];
    }
}

// one thread per column: insertion sort of its (row, value) entries by row
template <typename V>
__global__ void col_sort_kernel(uint64_t C, const uint64_t* __restrict__ colptr, uint32_t* __restrict__ crow,
                                V* __restrict__ cval) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint64_t a = colptr[c], b = colptr[c + 1];
    for (uint64_t i = a + 1; i < b; i++) {
        const uint32_t kr = crow[i];
        const V kv = cval[i];
        uint64_t j = i;
        while (j > a && crow[j - 1] > kr) {
            crow[j] = crow[j - 1];
            cval[j] = cval[j - 1];
            j--;
        }
        crow[j] = kr;
        cval[j] = kv;
    }
}

template <int VT>
__device__ __forceinline__ float entry_value(const void* cval, uint64_t e, const float* scale, uint32_t r) {
    if (VT == FD_VAL_F32) return ((const float*)cval)[e];
    const uint16_t h = ((const uint16_t*)cval)[e];
    return (VT == FD_VAL_BF16 ? bf16_to_f32(h) : fp16_to_f32(h)) * scale[r];
}

typedef float v4f __attribute__((ext_vector_type(4)));

// 64-bit value of lane `src` (wave-uniform result)
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Y[c][j] = sum over column c's entries (r, v) of v * X[r][j].  A wavefront takes tiles of 64
// consecutive columns: one coalesced load brings the tile's column pointers into lanes, then the
// wavefront walks the columns (pointers read across lanes), lanes along j (4 per lane, 16-byte
// X loads), and streams each Y row out with one 16-byte store per lane (empty columns store zeros).
template <int VT>
__global__ __launch_bounds__(256) void spmm_kernel(uint64_t C, const uint64_t* __restrict__ colptr,
                                                   const uint32_t* __restrict__ crow, const void* __restrict__ cval,
                                                   const float* __restrict__ scale, const float* __restrict__ X,
                                                   int k, float* __restrict__ Y) {
    const int lane = threadIdx.x & 63;
    const bool vec = (k & 3) == 0;  // rows of X and Y 16-byte aligned
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t tiles = (C + 63) / 64;
    for (uint64_t tile = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; tile < tiles; tile += waves) {
        const uint64_t c0 = tile * 64;
        const uint64_t cl = c0 + lane < C ? c0 + lane : C;
        const uint64_t lo = colptr[cl];
        const uint64_t hi = colptr[cl + 1 <= C ? cl + 1 : C];
        const int ncol = (int)(C - c0 < 64 ? C - c0 : 64);
        for (int jc = 0; jc < ncol; jc++) {
            const uint64_t c = c0 + jc;
            const uint64_t a = lane_u64(lo, jc), e_end = lane_u64(hi, jc);
            for (int j0 = 0; j0 < k; j0 += 256) {
                const int j = j0 + 4 * lane;
                if (vec && j + 3 < k) {
                    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    for (uint64_t e = a; e < e_end; e++) {
                        const uint32_t r = crow[e];
                        const float v = entry_value<VT>(cval, e, scale, r);
                        const float4 x = *(const float4*)(X + (uint64_t)r * k + j);
                        acc.x = fmaf(v, x.x, acc.x);
                        acc.y = fmaf(v, x.y, acc.y);
                        acc.z = fmaf(v, x.z, acc.z);
                        acc.w = fmaf(v, x.w, acc.w);
                    }
                    const v4f o = {acc.x, acc.y, acc.z, acc.w};
                    __builtin_nontemporal_store(o, (v4f*)(Y + c * (uint64_t)k + j));
                } else if (j < k) {
                    float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    for (uint64_t e = a; e < e_end; e++) {
                        const uint32_t r = crow[e];
                        const float v = entry_value<VT>(cval, e, scale, r);
                        for (int i = 0; i < 4 && j + i < k; i++) t[i] = fmaf(v, X[(uint64_t)r * k + j + i], t[i]);
                    }
                    for (int i = 0; i < 4 && j + i < k; i++) Y[c * (uint64_t)k + j + i] = t[i];
                }
            }
        }
    }
}

unsigned blocks_for(uint64_t threads, unsigned cap = 1u << 20) {
    const uint64_t b = (threads + 255) / 256;
    return (unsigned)(b < cap ? (b ? b : 1) : cap);
}

}  // namespace

hipError_t launch_global_rows(int n_mat, const uint64_t* row_first, const uint64_t* off_off, const uint64_t* nnz_off,
                              const uint64_t* offsets, uint64_t R, uint64_t total_nnz, uint64_t* rowptr,
                              hipStream_t stream) {
    hipLaunchKernelGGL(global_rows_kernel, dim3(blocks_for(R + 1, 0xFFFFFFFFu)), dim3(256), 0, stream, n_mat,
                       row_first, off_off, nnz_off, offsets, R, total_nnz, rowptr);
    return hipGetLastError();
}

hipError_t launch_pack_rows(uint64_t R, const uint64_t* rowptr, const float* values, int vtype, float* scale,
                            uint16_t* q, hipStream_t stream) {
    if (R == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_rows_kernel, dim3(blocks_for(R * 64, 0xFFFFFFFFu)), dim3(256), 0, stream, R, rowptr,
                       values, vtype, scale, q);
    return hipGetLastError();
}

hipError_t launch_col_count(uint64_t nnz, const uint64_t* columns, uint32_t* count, hipStream_t stream) {
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(col_count_kernel, dim3(blocks_for(nnz, 65536)), dim3(256), 0, stream, nnz, columns, count);
    return hipGetLastError();
}

hipError_t launch_col_scatter(uint64_t R, const uint64_t* rowptr, const uint64_t* columns, const void* vals,
                              int vbytes, const uint64_t* colptr, uint32_t* cursor, uint32_t* crow, void* cval,
                              hipStream_t stream) {
    if (R == 0) return hipSuccess;
    const dim3 g(blocks_for(R * 64, 0xFFFFFFFFu)), b(256);
    if (vbytes == 4)
        hipLaunchKernelGGL(col_scatter_kernel<float>, g, b, 0, stream, R, rowptr, columns, (const float*)vals, colptr,
                           cursor, crow, (float*)cval);
    else
        hipLaunchKernelGGL(col_scatter_kernel<uint16_t>, g, b, 0, stream, R, rowptr, columns, (const uint16_t*)vals,
                           colptr, cursor, crow, (uint16_t*)cval);
    return hipGetLastError();
}

hipError_t launch_col_sort(uint64_t C, const uint64_t* colptr, uint32_t* crow, void* cval, int vbytes,
                           hipStream_t stream) {
    if (C == 0) return hipSuccess;
    const dim3 g(blocks_for(C, 0xFFFFFFFFu)), b(256);
    if (vbytes == 4)
        hipLaunchKernelGGL(col_sort_kernel<float>, g, b, 0, stream, C, colptr, crow, (float*)cval);
    else
        hipLaunchKernelGGL(col_sort_kernel<uint16_t>, g, b, 0, stream, C, colptr, crow, (uint16_t*)cval);
    return hipGetLastError();
}

hipError_t launch_spmm(int vtype, uint64_t C, const uint64_t* colptr, const uint32_t* crow, const void* cval,
                       const float* scale, const float* X, int k, float* Y, hipStream_t stream) {
    if (C == 0 || k <= 0) return hipSuccess;
    const dim3 g(blocks_for((C + 63) / 64 * 64, 4096)), b(256);
    if (vtype == FD_VAL_F32)
        hipLaunchKernelGGL(spmm_kernel<FD_VAL_F32>, g, b, 0, stream, C, colptr, crow, cval, scale, X, k, Y);
    else if (vtype == FD_VAL_BF16)
        hipLaunchKernelGGL(spmm_kernel<FD_VAL_BF16>, g, b, 0, stream, C, colptr, crow, cval, scale, X, k, Y);
    else
        hipLaunchKernelGGL(spmm_kernel<FD_VAL_FP16>, g, b, 0, stream, C, colptr, crow, cval, scale, X, k, Y);
    return hipGetLastError();
}

}  // namespace fdi
