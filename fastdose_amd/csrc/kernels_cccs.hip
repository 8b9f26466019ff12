// kernels_cccs.hip -- collapsed-cone convolution on the BEV grid (gfx950).
//
// Algorithm (FastDoseLib/src/DoseComp.cu:217-392, maths in KernelGen/derivation.md:94-291).
// For every forward zenith cone theta and azimuth phi, the 256 BEV rays of a beamlet march
// through the grid in lockstep: they share one segment sequence and each ray sits at the shared
// lattice point plus its integer lane offset (lateral coordinates wrap with period 16).  For a
// segment of radiological length p through voxel m of the current slice:
//     dose[m] += (A/a ((1-g) T_m + g X_A) + B/b ((1-g) T_m + g X_B)) p,   cumu[m] += p
//     X <- T_m (1 - e^{-a p}) + e^{-a p} X          (same for B, b)
// with g = 1 as the reference computes it (integer division in d_calcg) or the exact
// (1-e^{-x})/x, and X reset to 0 when a ray crosses the BEV lateral edge.  At the end of a
// slice the dose summed over all phi is divided by (cumu of the last phi + 1e-4).
//
// Mapping to CDNA4.  Workgroup = (beamlet, forward cone), one wavefront per azimuth.  Because the
// lane -> voxel map of a segment is a wrapped shift of the shared midpoint voxel, the wavefront
// keeps everything voxel-indexed in registers: lane l owns voxels (x = l & 15, y = 4 (l >> 4) + r),
// r = 0..3, and holds their Terma, density, dose and cumulative p for the slice plus the X_A/X_B
// of the ray currently in each voxel.  When the shared midpoint moves to a neighbouring voxel the
// X state is rotated with it: along x by a DPP rotation inside each 16-lane row, along y by a
// register rotation across the four slots plus one cross-row move.  The inner loop therefore
// touches no LDS memory; LDS only stages the next slice's Terma/density slab (prefetched during
// the current slice) and the per-azimuth doses for the end-of-slice combine (double-buffered,
// one barrier per slice).  The march itself (next point, slice end, midpoint voxel, edge
// crossings) is FP32 with the exact operation order of the CPU oracle.
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
constexpr int kSlots = kRaysPerLane;  // voxels (rows) per lane

// Distance to the next voxel boundary along one axis, in units of the direction's parameter
// (d_nextPoint, DoseComp.cu:76-111): ceil(c + 1e-3) - c going up, floor(c - 1e-3) - c going
// down, infinite for a zero direction component.  With sgn = +1 / -1 for d > 0 / otherwise this
// is (sgn ceil(sgn c + 1e-3) - c) / d: sgn c + 1e-3 rounds to -(c - 1e-3) exactly when sgn = -1
// (round-to-nearest is symmetric), so each operation rounds exactly as the two-branch form; for
// d = +-0 the product is +inf.
__device__ __forceinline__ float axis_step(float c, float sgn, float inv) {
    return fmaf(sgn, ceilf(fmaf(sgn, c, kLargerEps)), -c) * inv;
}

// A wave-uniform float kept in a scalar register.
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// value of the neighbouring lane in the same 16-lane row (DPP row rotation)
__device__ __forceinline__ float row_from_left(float v) {   // dst[i] = src[i - 1]
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x121, 0xF, 0xF, false));
}
__device__ __forceinline__ float row_from_right(float v) {  // dst[i] = src[i + 1]
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x12F, 0xF, 0xF, false));
}
// value of lane src_lane
__device__ __forceinline__ float wave_from(float v, int src_lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
}

// X state follows the shared midpoint voxel: new[x] = old[x - 1] (dx = +1) and so on.
template <int N>
__device__ __forceinline__ void shift_x(float (&XA)[N], float (&XB)[N], int d) {
    for (; d > 0; d--)
#pragma unroll
        for (int r = 0; r < N; r++) {
            XA[r] = row_from_left(XA[r]);
            XB[r] = row_from_left(XB[r]);
        }
    for (; d < 0; d++)
#pragma unroll
        for (int r = 0; r < N; r++) {
            XA[r] = row_from_right(XA[r]);
            XB[r] = row_from_right(XB[r]);
        }
}
template <int N>
__device__ __forceinline__ void shift_y(float (&XA)[N], float (&XB)[N], int d, int up_lane, int down_lane) {
    for (; d > 0; d--) {  // new[y] = old[y - 1]
        const float ta = wave_from(XA[N - 1], up_lane);
        const float tb = wave_from(XB[N - 1], up_lane);
#pragma unroll
        for (int r = N - 1; r > 0; r--) {
            XA[r] = XA[r - 1];
            XB[r] = XB[r - 1];
        }
        XA[0] = ta;
        XB[0] = tb;
    }
    for (; d < 0; d++) {  // new[y] = old[y + 1]
        const float ta = wave_from(XA[0], down_lane);
        const float tb = wave_from(XB[0], down_lane);
#pragma unroll
        for (int r = 0; r < N - 1; r++) {
            XA[r] = XA[r + 1];
            XB[r] = XB[r + 1];
        }
        XA[N - 1] = ta;
        XB[N - 1] = tb;
    }
}

// Lattice coordinate (within the 16-wide period) of the one ray that crosses the lateral edge in
// a step from integer cell fb to fn (|fn - fb| <= 1): the ray with offset i crosses iff
// floor((fb + i) / 16) != floor((fn + i) / 16), i.e. i = -max(fb, fn) mod 16.  -1: none.
__device__ __forceinline__ int crossing_offset(int fb, int fn) {
    return fb == fn ? -1 : (-(fb > fn ? fb : fn)) & 15;
}

// Body of the CCCS kernel; CUMU: this wavefront carries the last azimuth and accumulates the
// cumulative p used for the slice normalisation (the other wavefronts skip it).
template <bool EXACT_G, bool CUMU>
__device__ __forceinline__ void cccs_body(const BeamletDev* __restrict__ beams, int n,
                                          const ConeTable* __restrict__ cones, int n_fwd,
                                          const float* __restrict__ terma, const float* __restrict__ dens_bev,
                                          float* __restrict__ partial, uint64_t stride, float* smem,
                                          int* s_next) {
    float* sTR = smem;             // [2 buffers][T | rho][256]
    float* sC = smem + 4 * kNpix;  // [2 buffers][256]
    float* sD = smem + 6 * kNpix;  // [2 buffers][n_phi][256]

    const int t = blockIdx.x / n;  // forward-cone slot, widest cone first
    const int b = blockIdx.x - t * n;
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int n_phi = nthreads >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const BeamletDev& B = beams[b];
    const int L = B.L;

    // cone constants
    const float A_a = uniform(cones->A_a[t]), B_b = uniform(cones->B_b[t]);
    const float na = uniform(-cones->a[t] * kLog2e), nb = uniform(-cones->b[t] * kLog2e);
    const float ca = uniform(cones->a[t]), cb = uniform(cones->b[t]);

    // grid direction of (theta, phi_wave): unit vector scaled to voxel units, renormalised
    const float ux = cones->ux[t][wave], uy = cones->uy[t][wave], uz = cones->uz[t][wave];
    const float vxd = ux / B.bs[0], vyd = uy / B.bs[1], vzd = uz / B.ls;
    const float rn = 1.0f / sqrtf(vxd * vxd + vyd * vyd + vzd * vzd);
    const float dx = uniform(vxd * rn), dy = uniform(vyd * rn), dz = uniform(vzd * rn);
    const float ix_ = uniform(1.0f / dx), iy_ = uniform(1.0f / dy), iz_ = uniform(1.0f / dz);
    const float sgx = dx > 0.0f ? 1.0f : -1.0f, sgy = dy > 0.0f ? 1.0f : -1.0f, sgz = dz > 0.0f ? 1.0f : -1.0f;

    const float inv_sad = uniform(1.0f / B.sad);
    const float bsx = uniform(B.bs[0]), bsy = uniform(B.bs[1]), ls = uniform(B.ls), lim_min = uniform(B.lim_min);

    // my voxels: x = vx, y = 4 g + r; flat index vbase + 16 r
    const int vx = lane & 15;
    const int g = lane >> 4;
    const int vbase = vx + 64 * g;
    const int up_lane = (lane + 48) & 63;    // lane (x, g - 1)
    const int down_lane = (lane + 16) & 63;  // lane (x, g + 1)

    float XA[kSlots], XB[kSlots];
#pragma unroll
    for (int r = 0; r < kSlots; r++) {
        XA[r] = 0.0f;
        XB[r] = 0.0f;
    }
    float bxp = 0.5f, byp = 0.5f, bzp = 0.0f;  // march position of this azimuth
    int mx_prev = 0, my_prev = 0;              // midpoint voxel the X state is aligned to
    int fbx = 0, fby = 0;                      // integer cells of the march position

    const float* __restrict__ Tg = terma + B.bev_off;
    const float* __restrict__ Rg = dens_bev + B.bev_off;
    float* __restrict__ P = partial + ((uint64_t)b * n_fwd + t) * stride;
    const bool last_phi = wave == n_phi - 1;
    const bool two = nthreads < 2 * kNpix;         // each thread prefetches two slab values
    const bool prefetch = nthreads >= kNpix;        // smaller workgroups reload each slab

    // slab of slice 0 into buffer 0 (T at [0, 256), rho at [256, 512))
    for (int i = tid; i < 2 * kNpix; i += nthreads) sTR[i] = i < kNpix ? Tg[i] : Rg[i - kNpix];
    __syncthreads();

    int s = 0;
    int buf = 0;
    int written = 0;  // slices [0, written) of P are final
    while (s < L) {
        // prefetch the next slice's slab (the usual successor; a skipped slice reloads below)
        const int sp = s + 1 < L ? s + 1 : s;
        float pf0 = 0.0f, pf1 = 0.0f;
        if (prefetch) {
            const int i0 = tid, i1 = tid + nthreads;
            if (i0 < 2 * kNpix) pf0 = i0 < kNpix ? Tg[(size_t)sp * kNpix + i0] : Rg[(size_t)sp * kNpix + i0 - kNpix];
            if (two && i1 < 2 * kNpix)
                pf1 = i1 < kNpix ? Tg[(size_t)sp * kNpix + i1] : Rg[(size_t)sp * kNpix + i1 - kNpix];
        }
        // this slice's Terma and density of my voxels
        const float* slab = sTR + buf * 2 * kNpix;
        float T[kSlots], R[kSlots], D[kSlots], Cp[kSlots];
#pragma unroll
        for (int r = 0; r < kSlots; r++) {
            T[r] = slab[vbase + 16 * r];
            R[r] = slab[kNpix + vbase + 16 * r];
            D[r] = 0.0f;
            Cp[r] = 0.0f;
        }

        for (int it = 0; it < kMaxStepsPerSlice; it++) {
            // next lattice point (d_nextPoint, DoseComp.cu:76-111)
            const float sx = axis_step(bxp, sgx, ix_);
            const float sy = axis_step(byp, sgy, iy_);
            const float sz = axis_step(bzp, sgz, iz_);
            const float st = fminf(fminf(sx, sy), sz);
            float nx = fmaf(st, dx, bxp), ny = fmaf(st, dy, byp), nz = fmaf(st, dz, bzp);
            // rint == round here: a tie (.5) is never within 1e-4 of an integer, and the snap is
            // only applied to values within 1e-4 of one
            // wave-uniform: taken as a scalar branch (one step per slice)
            const bool flag = __builtin_amdgcn_readfirstlane(fabsf(nz - rintf(nz)) < kEps ? 1 : 0) != 0;
            if (flag) {
                nx = fmaf(sz, dx, bxp);
                ny = fmaf(sz, dy, byp);
                nz = rintf(fmaf(sz, dz, bzp));
            }

            // shared midpoint voxel (wave-uniform); rotate the X state to follow it
            const int mx = __builtin_amdgcn_readfirstlane((int)floorf((bxp + nx) * 0.5f));
            const int my = __builtin_amdgcn_readfirstlane((int)floorf((byp + ny) * 0.5f));
            const int ddx = ((mx - mx_prev + 8) & 15) - 8;
            const int ddy = ((my - my_prev + 8) & 15) - 8;
            mx_prev = mx;
            my_prev = my;
            if (ddx == 1) {
#pragma unroll
                for (int r = 0; r < kSlots; r++) {
                    XA[r] = row_from_left(XA[r]);
                    XB[r] = row_from_left(XB[r]);
                }
            } else if (ddx == -1) {
#pragma unroll
                for (int r = 0; r < kSlots; r++) {
                    XA[r] = row_from_right(XA[r]);
                    XB[r] = row_from_right(XB[r]);
                }
            } else if (ddx != 0) {
                shift_x(XA, XB, ddx);
            }
            if (ddy != 0) shift_y(XA, XB, ddy, up_lane, down_lane);

            // segment length terms shared by all rays: physical lateral offset of a ray at lattice
            // coordinate c is (c - 8) bs f(z), f(z) = (z ls + lim_min) / sad (unwrapped)
            const float dzs = (nz - bzp) * ls;
            const float fe = fmaf(nz, ls, lim_min) * inv_sad;
            const float df = dzs * inv_sad;
            const float gx = bsx * df, gy = bsy * df;
            const float axc = fmaf(gx, bxp - 8.0f, bsx * (nx - bxp) * fe);
            const float ayc = fmaf(gy, byp - 8.0f, bsy * (ny - byp) * fe);
            const float dz2 = dzs * dzs;
            const int fnx = __builtin_amdgcn_readfirstlane((int)floorf(nx));
            const int fny = __builtin_amdgcn_readfirstlane((int)floorf(ny));

            // ray in my column: lane offset, lateral length term
            const int ix = (vx - mx) & 15;
            const float ddxl = fmaf(gx, (float)ix, axc);
            const float dxz2 = fmaf(ddxl, ddxl, dz2);
#pragma unroll
            for (int r = 0; r < kSlots; r++) {
                const int iy = (4 * g + r - my) & 15;
                const float ddyl = fmaf(gy, (float)iy, ayc);
                const float p = __builtin_amdgcn_sqrtf(fmaf(ddyl, ddyl, dxz2)) * R[r];
                // e^{-ap}, e^{-bp}
                const float ea = __builtin_amdgcn_exp2f(na * p);
                const float eb = __builtin_amdgcn_exp2f(nb * p);
                if (EXACT_G) {
                    const float oma = 1.0f - ea, omb = 1.0f - eb;
                    const float ap = ca * p, bp = cb * p;
                    const float ga = ap > 0.0f ? oma * __builtin_amdgcn_rcpf(ap) : 1.0f;
                    const float gb = bp > 0.0f ? omb * __builtin_amdgcn_rcpf(bp) : 1.0f;
                    D[r] = fmaf(fmaf(A_a, fmaf(ga, XA[r] - T[r], T[r]), B_b * fmaf(gb, XB[r] - T[r], T[r])), p,
                                D[r]);
                } else {
                    D[r] = fmaf(fmaf(A_a, XA[r], B_b * XB[r]), p, D[r]);
                }
                if (CUMU) Cp[r] += p;
                // X <- T (1 - e) + e X = T - e (T - X)
                XA[r] = fmaf(-ea, T[r] - XA[r], T[r]);
                XB[r] = fmaf(-eb, T[r] - XB[r], T[r]);
            }

            // rays that crossed the lateral BEV edge restart with X = 0 (DoseComp.cu:372-377).  A step
            // that moves one cell crosses the edge with one ray column and one ray row at most.
            if (abs(fnx - fbx) <= 1 && abs(fny - fby) <= 1) {
                const int cxo = crossing_offset(fbx, fnx), cyo = crossing_offset(fby, fny);
                if (cxo >= 0) {
                    const bool m = vx == ((mx + cxo) & 15);
#pragma unroll
                    for (int r = 0; r < kSlots; r++) {
                        XA[r] = m ? 0.0f : XA[r];
                        XB[r] = m ? 0.0f : XB[r];
                    }
                }
                if (cyo >= 0) {
                    const int yrow = (my + cyo) & 15;  // voxel row of the crossing rays
                    const bool m = g == (yrow >> 2);
                    const int rr = yrow & 3;
#pragma unroll
                    for (int r = 0; r < kSlots; r++)
                        if (r == rr) {
                            XA[r] = m ? 0.0f : XA[r];
                            XB[r] = m ? 0.0f : XB[r];
                        }
                }
            } else {
                // general test, ray by ray
                const bool cx = ((fbx + ix) >> 4) != ((fnx + ix) >> 4);
#pragma unroll
                for (int r = 0; r < kSlots; r++) {
                    const int iy = (4 * g + r - my) & 15;
                    const bool c = cx || (((fby + iy) >> 4) != ((fny + iy) >> 4));
                    XA[r] = c ? 0.0f : XA[r];
                    XB[r] = c ? 0.0f : XB[r];
                }
            }
            bxp = nx;
            byp = ny;
            bzp = nz;
            fbx = fnx;
            fby = fny;
            if (flag) break;
        }

        // publish this azimuth's slice dose, the next slab and the next slice index
        float* myD = sD + (buf * n_phi + wave) * kNpix;
#pragma unroll
        for (int r = 0; r < kSlots; r++) myD[vbase + 16 * r] = D[r];
        if (last_phi) {
#pragma unroll
            for (int r = 0; r < kSlots; r++) sC[buf * kNpix + vbase + 16 * r] = Cp[r];
        }
        {
            float* nslab = sTR + (buf ^ 1) * 2 * kNpix;
            const int i0 = tid, i1 = tid + nthreads;
            if (i0 < 2 * kNpix) nslab[i0] = pf0;
            if (two && i1 < 2 * kNpix) nslab[i1] = pf1;
        }
        if (tid == 0) s_next[buf] = max((int)rintf(bzp), s + 1);  // always advances
        __syncthreads();
        const int sn = s_next[buf];
        // combine slice s: sum over azimuths / (cumulative p of the last azimuth + 1e-4)
        for (int i = tid; i < kNpix; i += nthreads) {
            float sum = 0.0f;
            for (int f = 0; f < n_phi; f++) sum += sD[(buf * n_phi + f) * kNpix + i];
            P[(size_t)s * kNpix + i] = sum / (sC[buf * kNpix + i] + kEps);
        }
        for (int z = s + 1; z < sn && z < L; z++)  // slices the phi-0 march stepped over
            for (int i = tid; i < kNpix; i += nthreads) P[(size_t)z * kNpix + i] = 0.0f;
        written = sn < L ? sn : L;
        if ((sn != sp || !prefetch) && sn < L) {
            // the phi-0 march skipped a slice (or no prefetch): load the slab, then barrier
            __syncthreads();
            float* nslab = sTR + (buf ^ 1) * 2 * kNpix;
            for (int i = tid; i < 2 * kNpix; i += nthreads)
                nslab[i] = i < kNpix ? Tg[(size_t)sn * kNpix + i] : Rg[(size_t)sn * kNpix + i - kNpix];
            __syncthreads();
        }
        s = sn;
        buf ^= 1;
    }
    for (int z = written; z < L; z++)
        for (int i = tid; i < kNpix; i += nthreads) P[(size_t)z * kNpix + i] = 0.0f;
}

template <bool EXACT_G>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6, 8))) void cccs_kernel(
    const BeamletDev* __restrict__ beams, int n, const ConeTable* __restrict__ cones, int n_fwd,
    const float* __restrict__ terma, const float* __restrict__ dens_bev, float* __restrict__ partial,
    uint64_t stride) {
    // one LDS allocation shared by both instantiations of the body
    extern __shared__ float smem[];
    __shared__ int s_next[2];
    if ((threadIdx.x >> 6) == (blockDim.x >> 6) - 1)
        cccs_body<EXACT_G, true>(beams, n, cones, n_fwd, terma, dens_bev, partial, stride, smem, s_next);
    else
        cccs_body<EXACT_G, false>(beams, n, cones, n_fwd, terma, dens_bev, partial, stride, smem, s_next);
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
        const size_t shm = (size_t)(6 + 2 * n_phi) * kNpix * sizeof(float);
        if (exact_g)
            hipLaunchKernelGGL(cccs_kernel<true>, grid, block, shm, stream, beams, n, cones_dev, n_fwd, terma,
                               dens_bev, partial, stride);
        else
            hipLaunchKernelGGL(cccs_kernel<false>, grid, block, shm, stream, beams, n, cones_dev, n_fwd, terma,
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
