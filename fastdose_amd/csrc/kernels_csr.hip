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

// Three-pass scan for large inputs: tile sums, a scan of the tile sums, tile scans.  A tile is
// 4096 elements, 16 consecutive ones per thread.
constexpr int kTileThreads = 256;
constexpr int kTileItems = 16;
constexpr uint64_t kTile = (uint64_t)kTileThreads * kTileItems;

__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* s_wave, uint64_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int w = 0; w < nw; w++) {
            const uint64_t t = s_wave[w];
            s_wave[w] = acc;
            acc += t;
        }
        s_wave[nw] = acc;
    }
    __syncthreads();
    const uint64_t res = s_wave[wave] + x - v;
    total = s_wave[nw];
    __syncthreads();
    return res;
}

__global__ __launch_bounds__(kTileThreads) void scan_tile_sums_kernel(const uint32_t* __restrict__ count, uint64_t N,
                                                                      uint64_t* __restrict__ sums) {
    __shared__ uint64_t s_wave[kTileThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kTileItems;
    uint64_t v = 0;
#pragma unroll
    for (int i = 0; i < kTileItems; i++)
        if (base + i < N) v += count[base + i];
    uint64_t total;
    block_excl_scan_u64(v, s_wave, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of the tile sums in place (one workgroup), grand total to total[0]
__global__ __launch_bounds__(1024) void scan_sums_kernel(uint64_t* __restrict__ sums, uint64_t nb,
                                                         uint64_t* __restrict__ total) {
    __shared__ uint64_t s_wave[1024 / 64 + 1];
    uint64_t carry = 0;
    for (uint64_t i0 = 0; i0 < nb; i0 += blockDim.x) {
        const uint64_t i = i0 + threadIdx.x;
        const uint64_t v = i < nb ? sums[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan_u64(v, s_wave, tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) total[0] = carry;
}

__global__ __launch_bounds__(kTileThreads) void scan_tiles_kernel(const uint32_t* __restrict__ count, uint64_t N,
                                                                  const uint64_t* __restrict__ sums,
                                                                  uint64_t* __restrict__ off) {
    __shared__ uint64_t s_wave[kTileThreads / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kTileItems;
    uint32_t c[kTileItems];
    uint64_t v = 0;
#pragma unroll
    for (int i = 0; i < kTileItems; i++) {
        c[i] = base + i < N ? count[base + i] : 0u;
        v += c[i];
    }
    uint64_t total;
    uint64_t run = sums[blockIdx.x] + block_excl_scan_u64(v, s_wave, total);
#pragma unroll
    for (int i = 0; i < kTileItems; i++) {
        if (base + i < N) off[base + i] = run;
        run += c[i];
    }
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

uint64_t scan_scratch_size(uint64_t N) { return (N + kTile - 1) / kTile + 1; }

hipError_t launch_scan_u32(const uint32_t* count, uint64_t N, uint64_t* off, uint64_t* total, uint64_t* scratch,
                           hipStream_t stream) {
    if (!scratch || N <= kTile) {
        hipLaunchKernelGGL(scan_u32_kernel, dim3(1), dim3(kScanThreads), 0, stream, count, N, off, total);
        return hipGetLastError();
    }
    const uint64_t nb = (N + kTile - 1) / kTile;
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3((unsigned)nb), dim3(kTileThreads), 0, stream, count, N, scratch);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, stream, scratch, nb, total);
    hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)nb), dim3(kTileThreads), 0, stream, count, N, scratch, off);
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
