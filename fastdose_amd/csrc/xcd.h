// xcd.h -- workgroup order for the 8 XCDs of an MI355X.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fdi {

// The dispatcher hands workgroups to the 8 XCDs (each with its own L2) round-robin by linear id.
// This maps linear id `id` of `nb` workgroups to a logical index so that every XCD runs one
// contiguous range of logical indices: work with neighbouring indices (the planes of a beamlet,
// adjacent beamlets) then reads its data through one L2 instead of eight.  A bijection of [0, nb)
// for any nb.
__device__ __forceinline__ uint32_t xcd_logical_id(uint32_t id, uint32_t nb) {
    constexpr uint32_t kXcd = 8;
    const uint32_t q = nb / kXcd, r = nb % kXcd, x = id % kXcd, k = id / kXcd;
    return x * q + min(x, r) + k;
}

}  // namespace fdi
