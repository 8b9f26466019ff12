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
// the volume) instead of a texture unit.  It reads per-voxel corner records built once per context
// (the eight corners of the voxel's trilinear cell with a zero border, in 4 x 4 x 4 bricks): one
// 32-byte read per sample instead of eight gathers, no bounds tests, and rays of any direction
// share cache lines.
//
// The radiological path is a prefix sum along each ray, which one lane per ray would walk alone
// (256 rays per beamlet: one wavefront per SIMD on small batches).  Each ray is cut into 1, 2, 4,
// 8 or 16 segments of slices (a launch parameter; with 8 or 16, a beamlet's rays take two
// workgroups of 128 or four of 64), one wavefront group per segment: pass 1 writes the
// mean density of its slices, sums rho |step| over its segment and (FP32 path) stores the running
// in-segment sums of the four samples of each slice in a scratch buffer; after a barrier each
// segment adds the sum of the segments before it (in order) and pass 2 reads the stored sums back
// for the Terma.  (The FP64 path walks its samples again instead of storing doubles.)  Segment 0
// reproduces the sequential accumulation exactly; later segments differ from it by the rounding of
// one regrouped sum.  Rays with zero fluence skip pass 2's spectrum sum: their Terma is exactly 0.
#include <algorithm>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fdi {

namespace {

constexpr int kTermaThreads = 1024;  // workgroup size limit: ray segments x rays per workgroup
constexpr int kTermaMaxSeg = 16;    // most ray segments (16: four workgroups of 64 rays per beamlet)

// Corner records: for every base voxel (x, y, z) in [0, X] x [0, Y] x [0, Z] of the density with a
// one-voxel zero border, its eight trilinear corners c0..c7 (x..x+1, y..y+1, z..z+1; x fastest) as
// two float4 (c0, c1 - c0, c2, c3 - c2), (c4, c5 - c4, c6, c7 - c6): the x differences the first
// interpolation step needs, rounded as it would round them.  Records in 4 x 4 x 4 bricks: record
// ((z/4 nby + y/4) nbx + x/4) 64 + 16 (z%4) + 4 (y%4) + x%4, nb* = ceil((dim + 1) / 4).  A sample
// reads one 32-byte record.
__device__ __forceinline__ uint64_t corner_record(uint32_t x, uint32_t y, uint32_t z, uint32_t nbx, uint32_t nby) {
    return ((((uint64_t)(z >> 2) * nby + (y >> 2)) * nbx + (x >> 2)) << 6) + ((z & 3) << 4) + ((y & 3) << 2) + (x & 3);
}

__global__ __launch_bounds__(256) void expand_density_kernel(const float* __restrict__ density, int X, int Y, int Z,
                                                            uint32_t nbx, uint32_t nby, float4* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nx = (uint64_t)X + 1, ny = (uint64_t)Y + 1, nz = (uint64_t)Z + 1;
    if (i >= nx * ny * nz) return;
    const uint32_t x = (uint32_t)(i % nx), y = (uint32_t)((i / nx) % ny), z = (uint32_t)(i / (nx * ny));
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        // padded coordinate x + (c & 1) is original coordinate x + (c & 1) - 1
        const int ox = (int)x + (c & 1) - 1, oy = (int)y + ((c >> 1) & 1) - 1, oz = (int)z + (c >> 2) - 1;
        const bool in = ox >= 0 && ox < X && oy >= 0 && oy < Y && oz >= 0 && oz < Z;
        v[c] = in ? density[((size_t)oz * Y + oy) * X + ox] : 0.0f;
    }
    const uint64_t r = corner_record(x, y, z, nbx, nby);
    out[2 * r] = make_float4(v[0], v[1] - v[0], v[2], v[3] - v[2]);
    out[2 * r + 1] = make_float4(v[4], v[5] - v[4], v[6], v[7] - v[6]);
}

// Trilinear density at voxel-unit position (cx, cy, cz) (texel centres at i + 0.5, zero outside
// the volume) from the corner records (nbx, nby: bricks along x and y).  The last record read stays
// in registers: consecutive samples of a ray are a fraction of a voxel apart (four per slice), so
// most fall in the base voxel of the one before and need no load; a lane loads only when its base
// voxel changes.  (r, lo, hi): the record held, r = ~0 for none.
__device__ __forceinline__ float density_at(const float4* __restrict__ rec, int X, int Y, int Z, uint32_t nbx,
                                            uint32_t nby, float cx, float cy, float cz, uint64_t& r, float4& lo,
                                            float4& hi) {
    const float xb = cx - 0.5f, yb = cy - 0.5f, zb = cz - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    // outside: every corner is in the zero border or beyond (and the record index stays in range)
    if (!(fx > -2.0f && fx < (float)X && fy > -2.0f && fy < (float)Y && fz > -2.0f && fz < (float)Z)) return 0.0f;
    const float ax = xb - fx, ay = yb - fy, az = zb - fz;
    const uint64_t rn =
        corner_record((uint32_t)((int)fx + 1), (uint32_t)((int)fy + 1), (uint32_t)((int)fz + 1), nbx, nby);
    if (rn != r) {
        lo = rec[2 * rn];
        hi = rec[2 * rn + 1];
        r = rn;
    }
    const float c00 = fmaf(ax, lo.y, lo.x);
    const float c10 = fmaf(ax, lo.w, lo.z);
    const float c01 = fmaf(ax, hi.y, hi.x);
    const float c11 = fmaf(ax, hi.w, hi.z);
    const float c0 = fmaf(ay, c10 - c00, c00);
    const float c1 = fmaf(ay, c11 - c01, c01);
    return fmaf(az, c1 - c0, c0);
}

// rpl64_len: beamlets whose central-axis path (long_dim * long_spacing) exceeds this many cm
// accumulate the radiological path in FP64 (0: every beamlet; infinity: none).
// Workgroup (beamlet blockIdx.x, ray block blockIdx.y): `rays` consecutive BEV pixels (a multiple
// of 64) times blockDim.x / rays segments.
__global__ __launch_bounds__(kTermaThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void terma_kernel(
    const BeamletDev* __restrict__ beams, const float* __restrict__ fluence, const float* __restrict__ density_pad,
    int X, int Y, int Z, float vsx, float vsy, float vsz, SpectrumDev spec, float rpl64_len,
    float* __restrict__ terma, float* __restrict__ dens_bev, float4* __restrict__ scratch, int rays) {
    __shared__ float s_seg[kTermaThreads];
    __shared__ double s_seg64[kTermaThreads];
    const int b = blockIdx.x;
    const int lr = threadIdx.x % rays;
    const int p = blockIdx.y * rays + lr;
    const int seg = __builtin_amdgcn_readfirstlane(threadIdx.x / rays);
    const int nseg = blockDim.x / rays;
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

    const uint32_t nbx = ((uint32_t)X + 4) / 4, nby = ((uint32_t)Y + 4) / 4;
    const float4* __restrict__ rec = reinterpret_cast<const float4*>(density_pad);
    uint64_t rc = ~uint64_t(0);
    float4 rlo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rhi = rlo;
    auto rho_at = [&](int k) {
        const float t = fmaf((float)k, dt, t0);
        return density_at(rec, X, Y, Z, nbx, nby, fmaf(ax, t, sx), fmaf(ay, t, sy), fmaf(az, t, sz), rc, rlo, rhi);
    };
    // the FP64 path's second walk keeps no record in registers (fewer live registers there)
    auto rho_fresh = [&](int k) {
        const float t = fmaf((float)k, dt, t0);
        uint64_t r = ~uint64_t(0);
        float4 lo, hi;
        return density_at(rec, X, Y, Z, nbx, nby, fmaf(ax, t, sx), fmaf(ay, t, sy), fmaf(az, t, sz), r, lo, hi);
    };

    float* __restrict__ T = terma + B.bev_off + p;
    float* __restrict__ R = dens_bev + B.bev_off + p;
    float4* __restrict__ Q = scratch + B.bev_off + p;  // running in-segment sums, 4 per slice
    const bool RPL64 = (float)B.L * B.ls > rpl64_len;  // uniform over the workgroup
    const int i0 = (seg * B.L) / nseg, i1 = ((seg + 1) * B.L) / nseg;

    // pass 1: mean density per slice, radiological length of the segment
    float lpath = 0.0f;
    double lpath64 = 0.0;
    for (int i = i0; i < i1; i++) {
        float dsum = 0.0f;
        float q[kSuperSampling];
#pragma unroll
        for (int j = 0; j < kSuperSampling; j++) {
            const float rho = rho_at(kSuperSampling * i + j + 1);
            if (RPL64)
                lpath64 += (double)rho * (double)step_len;
            else
                lpath = fmaf(rho, step_len, lpath);
            q[j] = lpath;
            dsum += rho;
        }
        R[(size_t)i * kNpix] = dsum * (1.0f / kSuperSampling);
        if (!RPL64 && lit) Q[(size_t)i * kNpix] = make_float4(q[0], q[1], q[2], q[3]);
    }
    s_seg[threadIdx.x] = lpath;
    s_seg64[threadIdx.x] = lpath64;
    __syncthreads();
    float rpl = 0.0f;
    double rpl64 = 0.0;
    for (int s2 = 0; s2 < seg; s2++) {
        rpl += s_seg[s2 * rays + lr];
        rpl64 += s_seg64[s2 * rays + lr];
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
                rpl64 += (double)rho_fresh(k) * (double)step_len;
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

uint64_t density_records_floats(const fd_volume& vol) {
    const uint64_t nbx = (vol.dim[0] + 4) / 4, nby = (vol.dim[1] + 4) / 4, nbz = (vol.dim[2] + 4) / 4;
    return nbx * nby * nbz * 64 * 8;
}

hipError_t launch_expand_density(const float* density, const fd_volume& vol, float* records, hipStream_t stream) {
    const int X = vol.dim[0], Y = vol.dim[1], Z = vol.dim[2];
    const uint64_t n = (uint64_t)(X + 1) * (Y + 1) * (Z + 1);
    const uint32_t nbx = ((uint32_t)X + 4) / 4, nby = ((uint32_t)Y + 4) / 4;
    hipLaunchKernelGGL(expand_density_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, density, X, Y,
                       Z, nbx, nby, reinterpret_cast<float4*>(records));
    return hipGetLastError();
}

hipError_t launch_terma(const BeamletDev* beams, int n, const float* fluence, const float* density_pad,
                        const fd_volume& vol, const SpectrumDev& spec, float rpl64_len, float* terma,
                        float* dens_bev, float* scratch, int segments, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (segments != 1 && segments != 2 && segments != 4 && segments != 8 && segments != kTermaMaxSeg)
        return hipErrorInvalidValue;
    const int rays = std::min(kNpix, kTermaThreads / segments);
    dim3 grid(n, kNpix / rays), block(segments * rays);
    const int X = vol.dim[0], Y = vol.dim[1], Z = vol.dim[2];
    hipLaunchKernelGGL(terma_kernel, grid, block, 0, stream, beams, fluence, density_pad, X, Y, Z, vol.voxel_size[0],
                       vol.voxel_size[1], vol.voxel_size[2], spec, rpl64_len, terma, dens_bev,
                       reinterpret_cast<float4*>(scratch), rays);
    return hipGetLastError();
}

}  // namespace fdi
