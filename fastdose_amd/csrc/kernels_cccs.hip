// kernels_cccs.hip -- collapsed-cone convolution on the BEV grid (gfx950).
//
// Algorithm (FastDoseLib/src/DoseComp.cu:217-392, maths in KernelGen/derivation.md:94-291).
// For every forward zenith cone theta and azimuth phi, the 256 BEV rays of a beamlet march
// through the grid in lockstep: they share one segment sequence (the lattice is shift
// invariant, lateral coordinates wrap with period 16), each ray adding its lane offset.  For a
// segment of radiological length p through voxel m of the current slice:
//     dose[m] += (A/a ((1-g) T_m + g X_A) + B/b ((1-g) T_m + g X_B)) p,   cumu[m] += p
//     X <- T_m (1 - e^{-a p}) + e^{-a p} X          (same for B, b)
// with g = 1 as the reference computes it (integer division in d_calcg) or the exact
// (1-e^{-x})/x, and X reset to 0 when a ray crosses the BEV lateral edge.  At the end of a
// slice the dose summed over all phi is divided by (cumu of the last phi + 1e-4).
//
// Mapping to CDNA4.  Workgroup = (beamlet, forward cone), one wavefront per phi; a wavefront
// carries all 256 rays of its cone, 4 per lane, with X_A/X_B in registers.  The Terma and
// density slab of the current slice are staged once in LDS for all azimuths; each wavefront
// scatters into its own LDS dose slab with LDS float atomics (the per-segment lane -> voxel map
// is a wrapped shift, so collisions only occur at FP32 rounding edges and are still added), so
// no two wavefronts touch the same accumulator.  Three barriers per slice.  The march lattice
// (next point, slice end, wrapped indices, edge crossings) is computed in FP32 with the exact
// operation order of the CPU oracle; only the physics uses fast transcendentals.
//
// The slice loop follows the reference: the slice processed next is round(z) of the phi-0
// march, and every azimuth marches from its own position; slices the phi-0 march skips get no
// dose from this cone.  Partial doses of each cone go to `partial`; a second kernel sums the
// cones in a fixed order so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace fdi {

namespace {

constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ int wrap16(float m) {
    // fmodf(m, 16) (exact: m - 16 trunc(m/16)), shifted to [0, 16), floored; 16 wraps to 0
    float r = m - 16.0f * truncf(m * 0.0625f);
    r = (r < 0.0f) ? r + 16.0f : r;
    return ((int)floorf(r)) & 15;
}

template <bool EXACT_G>
__global__ __launch_bounds__(1024) void cccs_kernel(const BeamletDev* __restrict__ beams, int n,
                                                    const ConeTable* __restrict__ cones, int n_fwd,
                                                    const float* __restrict__ terma,
                                                    const float* __restrict__ dens_bev,
                                                    float* __restrict__ partial, uint64_t stride) {
    __shared__ float sT[kNpix];
    __shared__ float sR[kNpix];
    __shared__ float sC[kNpix];
    __shared__ float sD[kMaxPhi * kNpix];
    __shared__ int s_slice;
    __shared__ int s_next;

    const int t = blockIdx.x / n;  // forward-cone slot, widest cone first
    const int b = blockIdx.x - t * n;
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int n_phi = nthreads >> 6;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const BeamletDev& B = beams[b];
    const int L = B.L;

    // cone constants
    const float A_a = cones->A_a[t], B_b = cones->B_b[t];
    const float na = -cones->a[t] * kLog2e, nb = -cones->b[t] * kLog2e;
    const float ca = cones->a[t], cb = cones->b[t];

    // grid direction of (theta, phi_wave): unit vector scaled to voxel units, renormalised
    const float ux = cones->ux[t][wave], uy = cones->uy[t][wave], uz = cones->uz[t][wave];
    const float vx = ux / B.bs[0], vy = uy / B.bs[1], vz = uz / B.ls;
    const float rn = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
    const float dx = vx * rn, dy = vy * rn, dz = vz * rn;
    const float ix_ = 1.0f / dx, iy_ = 1.0f / dy, iz_ = 1.0f / dz;

    // segment-length constants: physical lateral offset = (coord - 8) bs f(z),
    // f(z) = (z ls + lim_min) / sad
    const float inv_sad = 1.0f / B.sad;
    const float bsx = B.bs[0], bsy = B.bs[1], ls = B.ls, lim_min = B.lim_min;

    const int ix = lane & 15;
    const int iy0 = lane >> 4;
    const float fix = (float)ix;

    float XA[kRaysPerLane], XB[kRaysPerLane];
#pragma unroll
    for (int r = 0; r < kRaysPerLane; r++) {
        XA[r] = 0.0f;
        XB[r] = 0.0f;
    }
    float bxp = 0.5f, byp = 0.5f, bzp = 0.0f;  // march position of this azimuth

    const float* __restrict__ Tg = terma + B.bev_off;
    const float* __restrict__ Rg = dens_bev + B.bev_off;
    float* __restrict__ P = partial + ((uint64_t)b * n_fwd + t) * stride;
    float* __restrict__ myD = sD + wave * kNpix;
    const bool last_phi = wave == n_phi - 1;

    if (tid == 0) s_slice = 0;
    __syncthreads();
    int written = 0;  // slices [0, written) of P are final
    while (true) {
        const int s = s_slice;
        if (s >= L) break;
        for (int z = written; z < s; z++)  // slices the phi-0 march stepped over
            for (int i = tid; i < kNpix; i += nthreads) P[(size_t)z * kNpix + i] = 0.0f;
        for (int i = tid; i < kNpix; i += nthreads) {
            sT[i] = Tg[(size_t)s * kNpix + i];
            sR[i] = Rg[(size_t)s * kNpix + i];
            sC[i] = 0.0f;
        }
        for (int i = tid; i < n_phi * kNpix; i += nthreads) sD[i] = 0.0f;
        __syncthreads();

        for (int it = 0; it < kMaxStepsPerSlice; it++) {
            // next lattice point (d_nextPoint, DoseComp.cu:76-111)
            const float sx = dx > 0.0f ? (ceilf(bxp + kLargerEps) - bxp) * ix_
                                       : (dx < 0.0f ? (floorf(bxp - kLargerEps) - bxp) * ix_ : INFINITY);
            const float sy = dy > 0.0f ? (ceilf(byp + kLargerEps) - byp) * iy_
                                       : (dy < 0.0f ? (floorf(byp - kLargerEps) - byp) * iy_ : INFINITY);
            const float sz = dz > 0.0f ? (ceilf(bzp + kLargerEps) - bzp) * iz_
                                       : (dz < 0.0f ? (floorf(bzp - kLargerEps) - bzp) * iz_ : INFINITY);
            const float st = fminf(fminf(sx, sy), sz);
            float nx = fmaf(st, dx, bxp), ny = fmaf(st, dy, byp), nz = fmaf(st, dz, bzp);
            const bool flag = fabsf(nz - roundf(nz)) < kEps;
            if (flag) {
                nx = fmaf(sz, dx, bxp);
                ny = fmaf(sz, dy, byp);
                nz = roundf(fmaf(sz, dz, bzp));
            }

            // segment length terms shared by all rays
            const float dzs = (nz - bzp) * ls;
            const float fe = fmaf(nz, ls, lim_min) * inv_sad;
            const float df = dzs * inv_sad;
            const float gx = bsx * df, gy = bsy * df;
            const float axc = fmaf(gx, bxp - 8.0f, bsx * (nx - bxp) * fe);
            const float ayc = fmaf(gy, byp - 8.0f, bsy * (ny - byp) * fe);
            const float dz2 = dzs * dzs;

            // lane x: wrapped voxel column and edge crossing
            const float btx = bxp + fix, ntx = nx + fix;
            const int mx = wrap16((btx + ntx) * 0.5f);
            const bool cross_x = floorf(btx * 0.0625f) != floorf(ntx * 0.0625f);
            const float ddx = fmaf(gx, fix, axc);
            const float dxz2 = fmaf(ddx, ddx, dz2);

#pragma unroll
            for (int r = 0; r < kRaysPerLane; r++) {
                const int iy = iy0 + 4 * r;
                const float fiy = (float)iy;
                const float bty = byp + fiy, nty = ny + fiy;
                const int my = wrap16((bty + nty) * 0.5f);
                const bool cross = cross_x || (floorf(bty * 0.0625f) != floorf(nty * 0.0625f));
                const int m = mx + (my << 4);
                const float Tm = sT[m];
                const float Rm = sR[m];
                const float ddy = fmaf(gy, fiy, ayc);
                const float seg = __builtin_sqrtf(fmaf(ddy, ddy, dxz2));
                const float p = seg * Rm;
                // 1 - e^{-ap}: series below 1/128 (exact 0 at p = 0, no cancellation), else exp2
                const float ap = ca * p, bp = cb * p;
                const float oma = ap < 0.0078125f ? ap * fmaf(-ap, fmaf(-ap, 1.0f / 6.0f, 0.5f), 1.0f)
                                                  : 1.0f - exp2f(na * p);
                const float omb = bp < 0.0078125f ? bp * fmaf(-bp, fmaf(-bp, 1.0f / 6.0f, 0.5f), 1.0f)
                                                  : 1.0f - exp2f(nb * p);
                float dose;
                if (EXACT_G) {
                    const float ga = ap > 0.0f ? oma / ap : 1.0f;
                    const float gb = bp > 0.0f ? omb / bp : 1.0f;
                    dose = fmaf(A_a, fmaf(ga, XA[r] - Tm, Tm), B_b * fmaf(gb, XB[r] - Tm, Tm)) * p;
                } else {
                    dose = fmaf(A_a, XA[r], B_b * XB[r]) * p;
                }
                atomicAdd(&myD[m], dose);
                if (last_phi) atomicAdd(&sC[m], p);
                const float xa = fmaf(oma, Tm - XA[r], XA[r]);
                const float xb = fmaf(omb, Tm - XB[r], XB[r]);
                XA[r] = cross ? 0.0f : xa;
                XB[r] = cross ? 0.0f : xb;
            }
            bxp = nx;
            byp = ny;
            bzp = nz;
            if (flag) break;
        }
        if (tid == 0) s_next = max((int)roundf(bzp), s + 1);  // always advances
        __syncthreads();
        for (int i = tid; i < kNpix; i += nthreads) {
            float sum = 0.0f;
            for (int f = 0; f < n_phi; f++) sum += sD[f * kNpix + i];
            P[(size_t)s * kNpix + i] = sum / (sC[i] + kEps);
        }
        written = s + 1;
        if (tid == 0) s_slice = s_next;
        __syncthreads();
    }
    for (int z = written; z < L; z++)
        for (int i = tid; i < kNpix; i += nthreads) P[(size_t)z * kNpix + i] = 0.0f;
}

// dose[b][s][p] = sum over forward cones of partial[b][t][s][p], cone slots in order.
__global__ __launch_bounds__(256) void cone_sum_kernel(const BeamletDev* __restrict__ beams, int n_fwd,
                                                       const float* __restrict__ partial,
                                                       float* __restrict__ dose, uint64_t stride) {
    const int b = blockIdx.y;
    const BeamletDev& B = beams[b];
    const uint64_t count = (uint64_t)B.L * kNpix;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float* src = partial + (uint64_t)b * n_fwd * stride + i;
    float acc = 0.0f;
    for (int t = 0; t < n_fwd; t++) acc += src[(uint64_t)t * stride];
    dose[B.bev_off + i] = acc;
}

}  // namespace

hipError_t launch_cccs(const BeamletDev* beams, int n, int max_L, const ConeTable* cones_dev, int n_fwd,
                       int n_phi, bool exact_g, const float* terma, const float* dens_bev, float* partial,
                       float* dose, uint64_t stride, hipStream_t stream, hipEvent_t* ev) {
    if (n <= 0) return hipSuccess;
    if (ev) (void)hipEventRecord(ev[0], stream);
    if (n_fwd > 0) {
        dim3 grid((unsigned)(n * n_fwd)), block(n_phi * kWave);
        if (exact_g)
            hipLaunchKernelGGL(cccs_kernel<true>, grid, block, 0, stream, beams, n, cones_dev, n_fwd, terma,
                               dens_bev, partial, stride);
        else
            hipLaunchKernelGGL(cccs_kernel<false>, grid, block, 0, stream, beams, n, cones_dev, n_fwd, terma,
                               dens_bev, partial, stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], stream);
    dim3 g2((unsigned)(((uint64_t)max_L * kNpix + 255) / 256), (unsigned)n), b2(256);
    hipLaunchKernelGGL(cone_sum_kernel, g2, b2, 0, stream, beams, n_fwd, partial, dose, stride);
    if (ev) (void)hipEventRecord(ev[2], stream);
    return hipGetLastError();
}

}  // namespace fdi
