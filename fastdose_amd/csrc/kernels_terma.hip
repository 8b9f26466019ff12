// kernels_terma.hip -- Terma ray trace in beam's-eye view (gfx950).
//
// Follows FastDoseLib/src/TermaComp.cu:179-253: for BEV pixel (ix, iy) of a beamlet the ray
// direction is Ri * ((ix + .5 - 8) bs.x, sad, (iy + .5 - 8) bs.y); sample k = 4i + j + 1 of
// slice i sits at depth fraction t_k = lim_min/sad + k * ls/(4 sad); the radiological path
// accumulates rho * |step|; Terma is the substep mean of
//   fluence_pix * sum_e w_e exp(-mu_e rpl) / t_k^2,   w_e = fluence_e E_e mu_e,
// and the density output is the substep mean of rho.  The spectrum sum is read from a table over
// the radiological depth built with the context (linear interpolation, error below ~5e-6).  The
// density is sampled with a hand-written trilinear filter (texel centres at i + 0.5, zero outside
// the volume) instead of a texture unit, reading a zero-bordered copy of the density (4 x 4 x 4
// bricks, so that the 64 rays of a wavefront share cache lines whatever the beam direction)
// through a buffer resource, so that the eight corners need no bounds tests.
//
// The radiological path is a prefix sum along each ray, which one lane per ray would walk alone
// (256 rays per beamlet: one wavefront per SIMD on small batches).  Each ray is cut into 1, 2 or
// 4 segments of slices (a launch parameter), one wavefront group per segment: pass 1 writes the
// mean density of its slices, sums rho |step| over its segment and (FP32 path) stores the running
// in-segment sums of the four samples of each slice in a scratch buffer; after a barrier each
// segment adds the sum of the segments before it (in order) and pass 2 reads the stored sums back
// for the Terma.  (The FP64 path walks its samples again instead of storing doubles.)  Segment 0
// reproduces the sequential accumulation exactly; later segments differ from it by the rounding of
// one regrouped sum.  Rays with zero fluence skip pass 2's spectrum sum: their Terma is exactly 0.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fdi {

namespace {

constexpr int kTermaSeg = 4;  // most ray segments per beamlet (wavefront groups of the workgroup)

// Trilinear density at voxel-unit position (cx, cy, cz) (texel centres at i + 0.5, zero outside
// the volume) from the zero-bordered, bricked copy (nbx, nby: bricks along x and y).
__device__ __forceinline__ float density_at(__amdgpu_buffer_rsrc_t rs, int X, int Y, int Z, uint32_t nbx,
                                            uint32_t nby, float cx, float cy, float cz) {
    const float xb = cx - 0.5f, yb = cy - 0.5f, zb = cz - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    // outside: every corner is in the zero border or beyond (and the offsets stay in range)
    if (!(fx > -2.0f && fx < (float)X && fy > -2.0f && fy < (float)Y && fz > -2.0f && fz < (float)Z)) return 0.0f;
    const float ax = xb - fx, ay = yb - fy, az = zb - fz;
    const uint32_t x = (uint32_t)((int)fx + 1), y = (uint32_t)((int)fy + 1), z = (uint32_t)((int)fz + 1);
    const uint32_t base = ((((z >> 2) * nby + (y >> 2)) * nbx + (x >> 2)) << 6) + ((z & 3) << 4) + ((y & 3) << 2) + (x & 3);
    // next voxel along each axis: inside the brick, or the first layer of the neighbouring brick
    const uint32_t dx = (x & 3) == 3 ? 64 - 3 : 1;
    const uint32_t dy = (y & 3) == 3 ? (nbx << 6) - 12 : 4;
    const uint32_t dz = (z & 3) == 3 ? ((nbx * nby) << 6) - 48 : 16;
    auto ld = [&](uint32_t i) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(i << 2), 0, 0));
    };
    const float v0 = ld(base), v1 = ld(base + dx), v2 = ld(base + dy), v3 = ld(base + dy + dx);
    const float v4 = ld(base + dz), v5 = ld(base + dz + dx), v6 = ld(base + dz + dy), v7 = ld(base + dz + dy + dx);
    const float c00 = fmaf(ax, v1 - v0, v0);
    const float c10 = fmaf(ax, v3 - v2, v2);
    const float c01 = fmaf(ax, v5 - v4, v4);
    const float c11 = fmaf(ax, v7 - v6, v6);
    const float c0 = fmaf(ay, c10 - c00, c00);
    const float c1 = fmaf(ay, c11 - c01, c01);
    return fmaf(az, c1 - c0, c0);
}

// rpl64_len: beamlets whose central-axis path (long_dim * long_spacing) exceeds this many cm
// accumulate the radiological path in FP64 (0: every beamlet; infinity: none).
__global__ __launch_bounds__(kTermaSeg * kNpix) void terma_kernel(
    const BeamletDev* __restrict__ beams, const float* __restrict__ fluence, const float* __restrict__ density_pad,
    int X, int Y, int Z, float vsx, float vsy, float vsz, SpectrumDev spec, float rpl64_len,
    float* __restrict__ terma, float* __restrict__ dens_bev, float4* __restrict__ scratch) {
    __shared__ float s_seg[kTermaSeg][kNpix];
    __shared__ double s_seg64[kTermaSeg][kNpix];
    const int b = blockIdx.x;
    const int p = threadIdx.x & (kNpix - 1);
    const int seg = __builtin_amdgcn_readfirstlane(threadIdx.x / kNpix);
    const int nseg = blockDim.x / kNpix;
    const BeamletDev& B = beams[b];
    const int ix = p & (kFmap - 1), iy = p >> 4;
    const float half = 0.5f * kFmap;

    // ray direction (BEV pixel centre at SAD) in PVCS
    const float bx = (ix + 0.5f - half) * B.bs[0];
    const float by = B.sad;
    const float bz = (iy + 0.5f - half) * B.bs[1];
    const float px = fmaf(B.Ri[2], bz, fmaf(B.Ri[1], by, B.Ri[0] * bx));
    const float py = fmaf(B.Ri[5], bz, fmaf(B.Ri[4], by, B.Ri[3] * bx));
    const float pz = fmaf(B.Ri[8], bz, fmaf(B.Ri[7], by, B.Ri[6] * bx));
    const float dt = B.ls / (B.sad * kSuperSampling);
    const float step_len = sqrtf(px * px + py * py + pz * pz) * dt;
    const float t0 = B.lim_min / B.sad;
    // positions in voxel units: c(t) = (src + p t) / vs
    const float ax = px / vsx, ay = py / vsy, az = pz / vsz;
    const float sx = B.src[0] / vsx, sy = B.src[1] / vsy, sz = B.src[2] / vsz;
    const float fl = fluence[(size_t)b * kNpix + p];
    const bool lit = fl != 0.0f;

    const uint32_t nbx = ((uint32_t)X + 5) / 4, nby = ((uint32_t)Y + 5) / 4, nbz = ((uint32_t)Z + 5) / 4;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)density_pad, (short)0, (int)(nbx * nby * nbz * 64 * sizeof(float)), 0x00020000);
    auto rho_at = [&](int k) {
        const float t = fmaf((float)k, dt, t0);
        return density_at(rs, X, Y, Z, nbx, nby, fmaf(ax, t, sx), fmaf(ay, t, sy), fmaf(az, t, sz));
    };

    float* __restrict__ T = terma + B.bev_off + p;
    float* __restrict__ R = dens_bev + B.bev_off + p;
    float4* __restrict__ Q = scratch + B.bev_off + p;  // running in-segment sums, 4 per slice
    const bool RPL64 = (float)B.L * B.ls > rpl64_len;  // uniform over the workgroup
    const int i0 = (seg * B.L) / nseg, i1 = ((seg + 1) * B.L) / nseg;

    // pass 1: mean density per slice, radiological length of the segment
    float lr = 0.0f;
    double lr64 = 0.0;
    for (int i = i0; i < i1; i++) {
        float dsum = 0.0f;
        float q[kSuperSampling];
#pragma unroll
        for (int j = 0; j < kSuperSampling; j++) {
            const float rho = rho_at(kSuperSampling * i + j + 1);
            if (RPL64)
                lr64 += (double)rho * (double)step_len;
            else
                lr = fmaf(rho, step_len, lr);
            q[j] = lr;
            dsum += rho;
        }
        R[(size_t)i * kNpix] = dsum * (1.0f / kSuperSampling);
        if (!RPL64 && lit) Q[(size_t)i * kNpix] = make_float4(q[0], q[1], q[2], q[3]);
    }
    s_seg[seg][p] = lr;
    s_seg64[seg][p] = lr64;
    __syncthreads();
    float rpl = 0.0f;
    double rpl64 = 0.0;
    for (int s2 = 0; s2 < seg; s2++) {
        rpl += s_seg[s2][p];
        rpl64 += s_seg64[s2][p];
    }

    // pass 2: Terma
    if (!lit) {
        for (int i = i0; i < i1; i++) T[(size_t)i * kNpix] = 0.0f;
        return;
    }
    for (int i = i0; i < i1; i++) {
        float tsum = 0.0f;
        float q[kSuperSampling];
        if (!RPL64) {
            const float4 qv = Q[(size_t)i * kNpix];
            q[0] = qv.x;
            q[1] = qv.y;
            q[2] = qv.z;
            q[3] = qv.w;
        }
#pragma unroll
        for (int j = 0; j < kSuperSampling; j++) {
            const int k = kSuperSampling * i + j + 1;
            const float t = fmaf((float)k, dt, t0);
            float path;
            if (RPL64) {
                rpl64 += (double)rho_at(k) * (double)step_len;
                path = (float)rpl64;
            } else {
                path = seg == 0 ? q[j] : rpl + q[j];
            }
            // sum_e w_e exp(-mu_e path): linear interpolation in the spectrum table
            const float x = path * spec.tab_inv_h;
            int it = (int)x;
            it = it < spec.tab_n - 2 ? it : spec.tab_n - 2;
            const float2 fe = reinterpret_cast<const float2*>(spec.tab)[it];
            const float acc = fmaf(x - (float)it, fe.y, fe.x);
            tsum = fmaf(acc, __builtin_amdgcn_rcpf(t * t), tsum);
        }
        T[(size_t)i * kNpix] = fl * tsum * (1.0f / kSuperSampling);
    }
}

}  // namespace

hipError_t launch_terma(const BeamletDev* beams, int n, const float* fluence, const float* density_pad,
                        const fd_volume& vol, const SpectrumDev& spec, float rpl64_len, float* terma,
                        float* dens_bev, float* scratch, int segments, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (segments != 1 && segments != 2 && segments != kTermaSeg) return hipErrorInvalidValue;
    dim3 grid(n), block(segments * kNpix);
    const int X = vol.dim[0], Y = vol.dim[1], Z = vol.dim[2];
    hipLaunchKernelGGL(terma_kernel, grid, block, 0, stream, beams, fluence, density_pad, X, Y, Z, vol.voxel_size[0],
                       vol.voxel_size[1], vol.voxel_size[2], spec, rpl64_len, terma, dens_bev,
                       reinterpret_cast<float4*>(scratch));
    return hipGetLastError();
}

}  // namespace fdi
