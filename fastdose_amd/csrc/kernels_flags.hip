// kernels_flags.hip -- beamlet / PTV intersection flags on the device (gfx950).
//
// d_beamletFlagInit (IMRTOpt/src/IMRTBeamBundle.cu:130-316): a beamlet of a bundle is active if
// any of its 3 x 3 lateral sample rays, at any of the central beam's long_dim depths, lands where
// the linearly filtered PTV mask is positive.  One wavefront per (beamlet, bundle): lanes take the
// (depth, sample) points in the host loop's order 64 at a time and stop at the first chunk with a
// hit (the flag is an OR, so the early exit does not change it).  The point arithmetic is the host
// code's FP32 sequence (explicit fmaf, IEEE division, no contraction), so flags are bit-identical
// to fd_beamlet_flags_host.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fdi {
namespace {

constexpr int kFlagSS = 3;

// Linear-filtered binary mask > 0 at unnormalised texture coordinates (texel i spans [i, i+1),
// border 0): true iff a texel with nonzero weight is set.
__device__ bool mask_hit_dev(const uint8_t* __restrict__ mask, int X, int Y, int Z, float cx, float cy, float cz) {
    const float t[3] = {cx - 0.5f, cy - 0.5f, cz - 0.5f};
    const int dim[3] = {X, Y, Z};
    int base[3], reach[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float f = floorf(t[d]);
        if (!(f > -2.0f && f < (float)dim[d] + 1.0f)) return false;
        base[d] = (int)f;
        reach[d] = (t[d] - f) > 0.0f ? 2 : 1;
    }
    for (int k = 0; k < reach[2]; k++) {
        const int z = base[2] + k;
        if (z < 0 || z >= Z) continue;
        for (int j = 0; j < reach[1]; j++) {
            const int y = base[1] + j;
            if (y < 0 || y >= Y) continue;
            for (int i = 0; i < reach[0]; i++) {
                const int x = base[0] + i;
                if (x < 0 || x >= X) continue;
                if (mask[(size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z)]) return true;
            }
        }
    }
    return false;
}

__global__ __launch_bounds__(64) void beamlet_flags_kernel(const FlagBeamDev* __restrict__ beams,
                                                           const uint8_t* __restrict__ mask, int X, int Y, int Z,
                                                           float vx, float vy, float vz,
                                                           uint8_t* __restrict__ flags) {
    const FlagBeamDev& d = beams[blockIdx.y];
    const int idx = blockIdx.x;
    if (idx >= d.fx * d.fy) return;
    const int bx = idx % d.fx, by = idx / d.fx;
    const int total = d.L * kFlagSS * kFlagSS;
    bool on = false;
    for (int t0 = 0; t0 < total && !on; t0 += 64) {
        const int t = t0 + (int)threadIdx.x;
        bool hit = false;
        if (t < total) {
            const uint32_t i = (uint32_t)(t / (kFlagSS * kFlagSS));
            const int r = t - (int)i * kFlagSS * kFlagSS;
            const int j = r / kFlagSS, k = r - j * kFlagSS;
            const float depth = d.lim_min + (float)i * d.ls;
            const float scale = depth / d.sad;
            const float dx = ((float)bx + (j + 0.5f) / (float)kFlagSS - (float)d.fx * 0.5f) * d.bs[0];
            const float dy = d.sad;
            const float dz = ((float)by + (k + 0.5f) / (float)kFlagSS - (float)d.fy * 0.5f) * d.bs[1];
            const float px = fmaf(d.Ri[2], dz, fmaf(d.Ri[1], dy, d.Ri[0] * dx));
            const float py = fmaf(d.Ri[5], dz, fmaf(d.Ri[4], dy, d.Ri[3] * dx));
            const float pz = fmaf(d.Ri[8], dz, fmaf(d.Ri[7], dy, d.Ri[6] * dx));
            hit = mask_hit_dev(mask, X, Y, Z, fmaf(px, scale, d.src[0]) / vx, fmaf(py, scale, d.src[1]) / vy,
                               fmaf(pz, scale, d.src[2]) / vz);
        }
        on = __any(hit);
    }
    if (threadIdx.x == 0) flags[d.out_off + idx] = on ? 1 : 0;
}

}  // namespace

hipError_t launch_beamlet_flags(const FlagBeamDev* beams, int n, int max_pixels, const uint8_t* mask,
                                const fd_volume& vol, uint8_t* flags, hipStream_t stream) {
    if (n <= 0 || max_pixels <= 0) return hipSuccess;
    dim3 grid((unsigned)max_pixels, (unsigned)n), block(64);
    hipLaunchKernelGGL(beamlet_flags_kernel, grid, block, 0, stream, beams, mask, (int)vol.dim[0], (int)vol.dim[1],
                       (int)vol.dim[2], vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2], flags);
    return hipGetLastError();
}

}  // namespace fdi
