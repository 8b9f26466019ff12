// noalloc.hip -- fd_dense_to_csr_noalloc: dense -> CSR with no allocation by the library.
//
// The same kernels as fd_dense_to_csr (count -> scan -> ordered fill -> offsets), but the per-row
// counts and row offsets live in one file-scope device array instead of hipMalloc'd buffers, and the
// nonzero total comes back through a plain (pageable) host variable instead of pinned memory.
// Each device that runs the library holds its own copy of the array until the process ends; its
// device address is looked up once per device (hipGetSymbolAddress) and only that address is ever
// handed to kernels.  Every call on one device shares the array: calls that overlap in time on
// one device race (the entry point synchronises its stream before it returns, so calls issued one
// after another from any streams do not).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.h"
#include "fastdose_hip.h"

namespace fdi {
namespace {

constexpr uint64_t kNoAllocMaxRows = FD_NOALLOC_MAX_ROWS;
// row offsets: u64 [kNoAllocMaxRows + 1]; then per-row counts: u32 [kNoAllocMaxRows], 256-aligned
constexpr size_t kOffBytes = (kNoAllocMaxRows + 1) * sizeof(uint64_t);
constexpr size_t kCntStart = (kOffBytes + 255) / 256 * 256;
constexpr size_t kScratchBytes = kCntStart + kNoAllocMaxRows * sizeof(uint32_t);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` on each device, once looked up

// Device address of this device's copy of the scratch array (looked up on its first use here).
int scratch_address(unsigned char** out) {
    int dev = 0;
    FD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal beyond the allocation-free entry point's table");
        return FD_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FD_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = (unsigned char*)g_scratch_dev[dev];
    return FD_OK;
}

}  // namespace
}  // namespace fdi

using namespace fdi;

extern "C" int fd_dense_to_csr_noalloc(const float* dense_dev, uint64_t rows, uint64_t cols, fd_csr* out,
                                       void* stream) {
    if (!out || !out->offsets || (rows > 0 && cols > 0 && !dense_dev)) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (rows > kNoAllocMaxRows) {
        set_error("fd_dense_to_csr_noalloc takes at most FD_NOALLOC_MAX_ROWS (1048576) rows; use fd_dense_to_csr");
        return FD_ERR_ARG;
    }
    if (cols >= (1ull << 32)) {
        set_error("dense matrix too large for the per-row u32 counts");
        return FD_ERR_ARG;
    }
    unsigned char* base = nullptr;
    FD_TRY(scratch_address(&base));
    uint64_t* off = (uint64_t*)base;
    uint32_t* cnt = (uint32_t*)(base + kCntStart);
    hipStream_t s = (hipStream_t)stream;
    out->rows = rows;
    // every count is written by the count kernel (or the fill), the scan overwrites every offset
    // and the total: nothing left by an earlier call is read
    if (rows > 0 && cols > 0) {
        FD_HIP(launch_dense_count(dense_dev, rows, cols, cnt, s));
    } else if (rows > 0) {
        FD_HIP(launch_fill_u32(cnt, rows, 0, s));
    }
    FD_HIP(launch_scan_u32(cnt, rows, off, off + rows, nullptr, s));
    uint64_t total = 0;
    FD_HIP(hipMemcpyAsync(&total, off + rows, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FD_HIP(hipStreamSynchronize(s));
    out->nnz = total;
    if (total > out->capacity || (total > 0 && (!out->columns || !out->values))) {
        set_error("CSR capacity too small; nnz reported in out->nnz");
        return FD_ERR_CAPACITY;
    }
    if (rows > 0 && cols > 0 && total > 0) FD_HIP(launch_dense_fill(dense_dev, rows, cols, off, out->columns, out->values, s));
    FD_HIP(launch_write_offsets(off, off + rows, rows, out->offsets, s));
    FD_HIP(hipStreamSynchronize(s));
    return FD_OK;
}
