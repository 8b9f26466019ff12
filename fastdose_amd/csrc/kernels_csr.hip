// kernels_csr.hip -- prefix scans and dense -> CSR conversion (gfx950).
//
// fd_dense_to_csr stands in for cusparseDenseToSparse in MatCSREnsemble::addMat
// (IMRTOpt/src/IMRTSparseMatEns.cu:78-161): row-major input, exact zeros dropped, ascending
// columns, 0-based 64-bit offsets.  Count pass (one workgroup per row) -> scan -> ordered fill
// (one workgroup per row, wavefront ballots keep the column order).
#include <algorithm>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fdi {

namespace {

constexpr int kScanThreads = 1024;

// Single-workgroup exclusive scan: thread i owns a contiguous chunk of the input.
__global__ __launch_bounds__(kScanThreads) void scan_u32_kernel(const uint32_t* __restrict__ count, uint64_t N,
                                                                uint64_t* __restrict__ off,
                                                                uint64_t* __restrict__ total) {
    __shared__ uint64_t s_part[kScanThreads];
    const uint64_t chunk = (N + kScanThreads - 1) / kScanThreads;
    const uint64_t lo = (uint64_t)threadIdx.x * chunk;
    const uint64_t hi = lo + chunk < N ? lo + chunk : N;
    uint64_t sum = 0;
    for (uint64_t i = lo; i < hi; i++) sum += count[i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele over the 1024 partial sums
    for (int o = 1; o < kScanThreads; o <<= 1) {
        uint64_t v = threadIdx.x >= (unsigned)o ? s_part[threadIdx.x - o] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = s_part[threadIdx.x] - sum;
    for (uint64_t i = lo; i < hi; i++) {
        off[i] = run;
        run += count[i];
    }
    if (threadIdx.x == kScanThreads - 1) total[0] = s_part[kScanThreads - 1];
}

__global__ __launch_bounds__(256) void dense_count_kernel(const float* __restrict__ dense, uint64_t cols,
                                                          uint32_t* __restrict__ row_count) {
    __shared__ uint32_t s_w[4];
    const float* row = dense + (uint64_t)blockIdx.x * cols;
    uint32_t c = 0;
    for (uint64_t j = threadIdx.x; j < cols; j += blockDim.x) c += row[j] != 0.0f ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) row_count[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(256) void dense_fill_kernel(const float* __restrict__ dense, uint64_t cols,
                                                         const uint64_t* __restrict__ row_off,
                                                         uint64_t* __restrict__ columns,
                                                         float* __restrict__ values) {
    __shared__ uint32_t s_w[5];
    const float* row = dense + (uint64_t)blockIdx.x * cols;
    uint64_t out = row_off[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t j0 = 0; j0 < cols; j0 += blockDim.x) {
        const uint64_t j = j0 + threadIdx.x;
        const float v = j < cols ? row[j] : 0.0f;
        const bool nz = v != 0.0f;
        const uint64_t mask = __ballot(nz);
        const uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t wbase = 0;
        for (int w = 0; w < wave; w++) wbase += s_w[w];
        const uint32_t tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (nz) {
            columns[out + wbase + before] = j;
            values[out + wbase + before] = v;
        }
        out += tot;
        __syncthreads();
    }
}

__global__ void write_offsets_kernel(const uint64_t* __restrict__ row_off, const uint64_t* __restrict__ total,
                                     uint64_t rows, uint64_t* __restrict__ offsets) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) offsets[r] = row_off[r];
    if (r == rows) offsets[rows] = total[0];
}

}  // namespace

hipError_t launch_scan_u32(const uint32_t* count, uint64_t N, uint64_t* off, uint64_t* total,
                           hipStream_t stream) {
    hipLaunchKernelGGL(scan_u32_kernel, dim3(1), dim3(kScanThreads), 0, stream, count, N, off, total);
    return hipGetLastError();
}

hipError_t launch_dense_count(const float* dense, uint64_t rows, uint64_t cols, uint32_t* row_count,
                              hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(dense_count_kernel, dim3((unsigned)rows), dim3(256), 0, stream, dense, cols, row_count);
    return hipGetLastError();
}

hipError_t launch_dense_fill(const float* dense, uint64_t rows, uint64_t cols, const uint64_t* row_off,
                             uint64_t* columns, float* values, hipStream_t stream) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(dense_fill_kernel, dim3((unsigned)rows), dim3(256), 0, stream, dense, cols, row_off, columns,
                       values);
    return hipGetLastError();
}

// Plain fill kernels: clearing through a kernel rather than a memset node keeps every stage
// capturable into a HIP graph.
template <typename T>
__global__ void fill_kernel(T* p, uint64_t n, T v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_write_offsets(const uint64_t* row_off, const uint64_t* total, uint64_t rows, uint64_t* offsets,
                                hipStream_t stream) {
    hipLaunchKernelGGL(write_offsets_kernel, dim3((unsigned)((rows + 1 + 255) / 256)), dim3(256), 0, stream,
                       row_off, total, rows, offsets);
    return hipGetLastError();
}

}  // namespace fdi
