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

constexpr int kSlots = kRaysPerLane;  // voxels (rows) per lane

// Per-cone partial dose of beamlet B for forward-cone slot t: the beamlet's n_fwd volumes of
// L x 256 floats start at n_fwd x its BEV offset, so the partial buffer has the BEV layout of the
// batch scaled by the cone count (n_fwd x the BEV floats of the batch).
__device__ __forceinline__ float* cone_partial(float* partial, const BeamletDev& B, int n_fwd, int t) {
    return partial + (uint64_t)B.bev_off * (uint64_t)n_fwd + (uint64_t)t * (uint64_t)B.L * kNpix;
}

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
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x121, 0xF, 0xF, false));
}
__device__ __forceinline__ float row_from_right(float v) {  // dst[i] = src[i + 1]
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x12F, 0xF, 0xF, false));
}
// the same shifts without the wrap: the lane that would receive the other end of its row gets 0
__device__ __forceinline__ float row_from_left_zero(float v) {   // dst[i] = i % 16 ? src[i - 1] : 0
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_from_right_zero(float v) {  // dst[i] = i % 16 != 15 ? src[i + 1] : 0
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x101, 0xF, 0xF, true));
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

// In-place slot rotations of a voxel-indexed array (four slots = four voxel rows of the lane):
// slots_down: v[r] <- v[r - 1], v[0] <- in; slots_up: v[r] <- v[r + 1], v[3] <- in.  Written as
// register moves on operands tied to their own registers, so that the steps without a y shift
// keep the array where it is (a phi over a permutation of the slots would otherwise make the
// compiler copy the array on every step).
__device__ __forceinline__ void slots_down(float (&v)[4], float in) {
    asm volatile(
        "v_mov_b32 %3, %2\n\t"
        "v_mov_b32 %2, %1\n\t"
        "v_mov_b32 %1, %0\n\t"
        "v_mov_b32 %0, %4"
        : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3])
        : "v"(in));
}
__device__ __forceinline__ void slots_up(float (&v)[4], float in) {
    asm volatile(
        "v_mov_b32 %0, %1\n\t"
        "v_mov_b32 %1, %2\n\t"
        "v_mov_b32 %2, %3\n\t"
        "v_mov_b32 %3, %4"
        : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3])
        : "v"(in));
}

// The same for the X state and a third voxel-indexed array that follows the rays (their offsets).
template <int N>
__device__ __forceinline__ void shift_y3(float (&XA)[N], float (&XB)[N], float (&Z)[N], int d, int up_lane,
                                         int down_lane) {
    for (; d > 0; d--) {  // new[y] = old[y - 1]
        const float ta = wave_from(XA[N - 1], up_lane);
        const float tb = wave_from(XB[N - 1], up_lane);
        const float tz = wave_from(Z[N - 1], up_lane);
#pragma unroll
        for (int r = N - 1; r > 0; r--) {
            XA[r] = XA[r - 1];
            XB[r] = XB[r - 1];
            Z[r] = Z[r - 1];
        }
        XA[0] = ta;
        XB[0] = tb;
        Z[0] = tz;
    }
    for (; d < 0; d++) {  // new[y] = old[y + 1]
        const float ta = wave_from(XA[0], down_lane);
        const float tb = wave_from(XB[0], down_lane);
        const float tz = wave_from(Z[0], down_lane);
#pragma unroll
        for (int r = 0; r < N - 1; r++) {
            XA[r] = XA[r + 1];
            XB[r] = XB[r + 1];
            Z[r] = Z[r + 1];
        }
        XA[N - 1] = ta;
        XB[N - 1] = tb;
        Z[N - 1] = tz;
    }
}

// Lattice coordinate (within the 16-wide period) of the one ray that crosses the lateral edge in
// a step from integer cell fb to fn (|fn - fb| <= 1): the ray with offset i crosses iff
// floor((fb + i) / 16) != floor((fn + i) / 16), i.e. i = -max(fb, fn) mod 16.  -1: none.
__device__ __forceinline__ int crossing_offset(int fb, int fn) {
    return fb == fn ? -1 : (-(fb > fn ? fb : fn)) & 15;
}

// Segment lengths of a step without a square root per ray.  The ray with lattice offsets (i, j)
// moves by (gx i + axc, gy j + ayc, dzs), squared length q; the rays of one step differ only by
// the beam's divergence over the step, so q lies within (sqrt(2) e + e^2) s0^2 of the central
// ray's S0 = s0^2 (offsets 7.5), e = 7.5 bs / sad.  One Heron step from s0, s = q / (2 s0) + s0 / 2,
// is within (q - S0)^2 / (8 S0 q) of sqrt(q) and never below it: 2.2e-5 relative at most for
// 1.25 mm pixels at 100 cm, 9e-5 for 2.5 mm.  With k = 1 / sqrt(2 s0) the ray terms are carried
// scaled by k and the constant c = k^2 dzs^2 + s0 / 2, so s = (k ddy)^2 + ((k ddx)^2 + c): two FMAs
// on values the step computes anyway.  s0 and k come from the hardware square root and reciprocal
// square root (within an ulp or so, which moves s by about 1e-7 relative).  A step with S0 < 1e-30
// (every ray shorter than 1e-15 cm; 0 for S0 = 0) gets k = c = 0, length 0.  Returns the step's
// record (gx, gy, axc, ayc) k and c.
__device__ __forceinline__ float4 heron_record(float gx, float gy, float axc, float ayc, float dz2, float& c) {
    const float cx = fmaf(gx, 7.5f, axc), cy = fmaf(gy, 7.5f, ayc);
    const float S0 = fmaf(cx, cx, fmaf(cy, cy, dz2));
    const bool pos = S0 > 1e-30f;
    const float s0 = __builtin_amdgcn_sqrtf(pos ? S0 : 1.0f);
    const float k = pos ? __builtin_amdgcn_rsqf(2.0f * s0) : 0.0f;
    c = pos ? fmaf(k * k, dz2, 0.5f * s0) : 0.0f;
    return make_float4(gx * k, gy * k, axc * k, ayc * k);
}

// March state of one azimuth in a self-marching wavefront: grid direction, its reciprocals and
// signs, the march position, the midpoint voxel the X state is aligned to and the integer cells
// of the position.
struct SelfMarch {
    float dx, dy, dz, ix_, iy_, iz_, sgx, sgy, sgz;
    float bxp, byp, bzp;
    int mx_prev, my_prev, fbx, fby;
};

// One slice of one azimuth for a self-marching wavefront: steps until the march reaches the slice
// end; CUMU: this azimuth is the last one and accumulates the cumulative p of the slice.
template <bool EXACT_G, bool CUMU>
__device__ __forceinline__ void self_march_slice(SelfMarch& M, float (&XA)[kSlots], float (&XB)[kSlots],
                                                 const float (&T)[kSlots], const float (&R)[kSlots],
                                                 float (&D)[kSlots], float (&Cp)[kSlots], float A_a, float B_b,
                                                 float na, float nb, float ca, float cb, float bsx, float bsy,
                                                 float ls, float lim_min, float inv_sad, int vx, int g, int up_lane,
                                                 int down_lane) {
    for (int it = 0; it < kMaxStepsPerSlice; it++) {
        // next lattice point (d_nextPoint, DoseComp.cu:76-111)
        const float sx = axis_step(M.bxp, M.sgx, M.ix_);
        const float sy = axis_step(M.byp, M.sgy, M.iy_);
        const float sz = axis_step(M.bzp, M.sgz, M.iz_);
        const float st = fminf(fminf(sx, sy), sz);
        float nx = fmaf(st, M.dx, M.bxp), ny = fmaf(st, M.dy, M.byp), nz = fmaf(st, M.dz, M.bzp);
        // rint == round here: a tie (.5) is never within 1e-4 of an integer, and the snap is only
        // applied to values within 1e-4 of one; wave-uniform, taken as a scalar branch
        const bool flag = __builtin_amdgcn_readfirstlane(fabsf(nz - rintf(nz)) < kEps ? 1 : 0) != 0;
        if (flag) {
            nx = fmaf(sz, M.dx, M.bxp);
            ny = fmaf(sz, M.dy, M.byp);
            nz = rintf(fmaf(sz, M.dz, M.bzp));
        }

        // shared midpoint voxel (wave-uniform); rotate the X state to follow it
        const int mx = __builtin_amdgcn_readfirstlane((int)floorf((M.bxp + nx) * 0.5f));
        const int my = __builtin_amdgcn_readfirstlane((int)floorf((M.byp + ny) * 0.5f));
        const int ddx = ((mx - M.mx_prev + 8) & 15) - 8;
        const int ddy = ((my - M.my_prev + 8) & 15) - 8;
        M.mx_prev = mx;
        M.my_prev = my;
        if (ddx != 0) shift_x(XA, XB, ddx);
        if (ddy != 0) shift_y(XA, XB, ddy, up_lane, down_lane);

        // segment length terms shared by all rays: physical lateral offset of a ray at lattice
        // coordinate c is (c - 8) bs f(z), f(z) = (z ls + lim_min) / sad (unwrapped)
        const float dzs = (nz - M.bzp) * ls;
        const float fe = fmaf(nz, ls, lim_min) * inv_sad;
        const float df = dzs * inv_sad;
        const float gx = bsx * df, gy = bsy * df;
        const float axc = fmaf(gx, M.bxp - 8.0f, bsx * (nx - M.bxp) * fe);
        const float ayc = fmaf(gy, M.byp - 8.0f, bsy * (ny - M.byp) * fe);
        float hc;
        const float4 hr = heron_record(gx, gy, axc, ayc, dzs * dzs, hc);
        const int fnx = __builtin_amdgcn_readfirstlane((int)floorf(nx));
        const int fny = __builtin_amdgcn_readfirstlane((int)floorf(ny));

        // ray in my column: lane offset, lateral length term (scaled, heron_record)
        const int ix = (vx - mx) & 15;
        const float ddxl = fmaf(hr.x, (float)ix, hr.z);
        const float dxz2 = fmaf(ddxl, ddxl, hc);
#pragma unroll
        for (int r = 0; r < kSlots; r++) {
            const int iy = (4 * g + r - my) & 15;
            const float ddyl = fmaf(hr.y, (float)iy, hr.w);
            const float pp = fmaf(ddyl, ddyl, dxz2) * R[r];
            // e^{-ap}, e^{-bp}
            const float ea = __builtin_amdgcn_exp2f(na * pp);
            const float eb = __builtin_amdgcn_exp2f(nb * pp);
            if (EXACT_G) {
                const float oma = 1.0f - ea, omb = 1.0f - eb;
                const float ap = ca * pp, bp = cb * pp;
                const float ga = ap > 0.0f ? oma * __builtin_amdgcn_rcpf(ap) : 1.0f;
                const float gb = bp > 0.0f ? omb * __builtin_amdgcn_rcpf(bp) : 1.0f;
                D[r] = fmaf(fmaf(A_a, fmaf(ga, XA[r] - T[r], T[r]), B_b * fmaf(gb, XB[r] - T[r], T[r])), pp, D[r]);
            } else {
                D[r] = fmaf(fmaf(A_a, XA[r], B_b * XB[r]), pp, D[r]);
            }
            if (CUMU) Cp[r] += pp;
            // X <- T (1 - e) + e X = T - e (T - X)
            XA[r] = fmaf(-ea, T[r] - XA[r], T[r]);
            XB[r] = fmaf(-eb, T[r] - XB[r], T[r]);
        }

        // rays that crossed the lateral BEV edge restart with X = 0 (DoseComp.cu:372-377).  A step
        // that moves one cell crosses the edge with one ray column and one ray row at most.
        if (abs(fnx - M.fbx) <= 1 && abs(fny - M.fby) <= 1) {
            const int cxo = crossing_offset(M.fbx, fnx), cyo = crossing_offset(M.fby, fny);
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
            const bool cx = ((M.fbx + ix) >> 4) != ((fnx + ix) >> 4);
#pragma unroll
            for (int r = 0; r < kSlots; r++) {
                const int iy = (4 * g + r - my) & 15;
                const bool c = cx || (((M.fby + iy) >> 4) != ((fny + iy) >> 4));
                XA[r] = c ? 0.0f : XA[r];
                XB[r] = c ? 0.0f : XB[r];
            }
        }
        M.bxp = nx;
        M.byp = ny;
        M.bzp = nz;
        M.fbx = fnx;
        M.fby = fny;
        if (flag) break;
    }
}

// Workgroup of self-marching wavefronts (more azimuths than the march kernel's LDS holds):
// wavefront w carries azimuths APW w .. APW w + APW - 1 (those below n_phi), each with its own X
// state and march, one after the other within every slice; the wavefront holding the last azimuth
// (always the last wavefront) accumulates that azimuth's cumulative p.  APW = 2 keeps 17-24 azimuths
// within 12 wavefronts.
template <bool EXACT_G, int APW>
__device__ __forceinline__ void cccs_body(const BeamletDev* __restrict__ beams, int n,
                                          const ConeTable* __restrict__ cones, int n_fwd,
                                          const float* __restrict__ terma, const float* __restrict__ dens_bev,
                                          float* __restrict__ partial, float* smem, int* s_next) {
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int n_waves = nthreads >> 6;
    const int n_phi = cones->n_phi;
    float* sTR = smem;             // [2 buffers][T | rho][256]
    float* sC = smem + 4 * kNpix;  // [2 buffers][256]
    float* sD = smem + 6 * kNpix;  // [2 buffers][n_waves][256]

    const int t = blockIdx.x / n;  // forward-cone slot, widest cone first
    const int b = blockIdx.x - t * n;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const BeamletDev& B = beams[b];
    const int L = B.L;
    const int az0 = APW * wave;                                   // first azimuth of this wavefront
    const int n_az = n_phi - az0 < APW ? n_phi - az0 : APW;       // its azimuth count (wave-uniform)
    const int cumu_slot = (n_phi - 1) - az0;                      // slot of the last azimuth, if here
    const bool last_wave = wave == n_waves - 1;

    // cone constants
    const float A_a = uniform(cones->A_a[t]), B_b = uniform(cones->B_b[t]);
    const float na = cones->na[t], nb = cones->nb[t];  // scalar loads
    const float ca = uniform(cones->a[t]), cb = uniform(cones->b[t]);
    const float inv_sad = uniform(1.0f / B.sad);
    const float bsx = uniform(B.bs[0]), bsy = uniform(B.bs[1]), ls = uniform(B.ls), lim_min = uniform(B.lim_min);

    // march state of my azimuths; grid direction of (theta, phi): unit vector scaled to voxel units,
    // renormalised
    SelfMarch Mz[APW];
    float XA[APW][kSlots], XB[APW][kSlots];
#pragma unroll
    for (int a = 0; a < APW; a++) {
        const int f = az0 + a < n_phi ? az0 + a : az0;
        const float ux = cones->ux[t][f], uy = cones->uy[t][f], uz = cones->uz[t][f];
        const float vxd = ux / B.bs[0], vyd = uy / B.bs[1], vzd = uz / B.ls;
        const float rn = 1.0f / sqrtf(vxd * vxd + vyd * vyd + vzd * vzd);
        SelfMarch& M = Mz[a];
        M.dx = uniform(vxd * rn);
        M.dy = uniform(vyd * rn);
        M.dz = uniform(vzd * rn);
        M.ix_ = uniform(1.0f / M.dx);
        M.iy_ = uniform(1.0f / M.dy);
        M.iz_ = uniform(1.0f / M.dz);
        M.sgx = M.dx > 0.0f ? 1.0f : -1.0f;
        M.sgy = M.dy > 0.0f ? 1.0f : -1.0f;
        M.sgz = M.dz > 0.0f ? 1.0f : -1.0f;
        M.bxp = 0.5f;
        M.byp = 0.5f;
        M.bzp = 0.0f;
        M.mx_prev = M.my_prev = M.fbx = M.fby = 0;
#pragma unroll
        for (int r = 0; r < kSlots; r++) {
            XA[a][r] = 0.0f;
            XB[a][r] = 0.0f;
        }
    }

    // my voxels: x = vx, y = 4 g + r; flat index vbase + 16 r
    const int vx = lane & 15;
    const int g = lane >> 4;
    const int vbase = vx + 64 * g;
    const int up_lane = (lane + 48) & 63;    // lane (x, g - 1)
    const int down_lane = (lane + 16) & 63;  // lane (x, g + 1)

    const float* __restrict__ Tg = terma + B.bev_off;
    const float* __restrict__ Rg = dens_bev + B.bev_off;
    float* __restrict__ P = cone_partial(partial, B, n_fwd, t);
    const bool two = nthreads < 2 * kNpix;    // each thread prefetches two slab values
    const bool prefetch = nthreads >= kNpix;  // smaller workgroups reload each slab

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
#pragma unroll
        for (int a = 0; a < APW; a++) {
            if (a >= n_az) break;  // wave-uniform
            if (a == cumu_slot)
                self_march_slice<EXACT_G, true>(Mz[a], XA[a], XB[a], T, R, D, Cp, A_a, B_b, na, nb, ca, cb, bsx, bsy,
                                                ls, lim_min, inv_sad, vx, g, up_lane, down_lane);
            else
                self_march_slice<EXACT_G, false>(Mz[a], XA[a], XB[a], T, R, D, Cp, A_a, B_b, na, nb, ca, cb, bsx,
                                                 bsy, ls, lim_min, inv_sad, vx, g, up_lane, down_lane);
        }

        // publish this wavefront's slice dose (sum over its azimuths), the next slab and the next
        // slice index (the phi-0 march decides it)
        float* myD = sD + (buf * n_waves + wave) * kNpix;
#pragma unroll
        for (int r = 0; r < kSlots; r++) myD[vbase + 16 * r] = D[r];
        if (last_wave) {
#pragma unroll
            for (int r = 0; r < kSlots; r++) sC[buf * kNpix + vbase + 16 * r] = Cp[r];
        }
        {
            float* nslab = sTR + (buf ^ 1) * 2 * kNpix;
            const int i0 = tid, i1 = tid + nthreads;
            if (i0 < 2 * kNpix) nslab[i0] = pf0;
            if (two && i1 < 2 * kNpix) nslab[i1] = pf1;
        }
        if (tid == 0) s_next[buf] = max((int)rintf(Mz[0].bzp), s + 1);  // always advances
        __syncthreads();
        const int sn = s_next[buf];
        // combine slice s: sum over azimuths / (cumulative p of the last azimuth + 1e-4)
        for (int i = tid; i < kNpix; i += nthreads) {
            float sum = 0.0f;
            for (int f = 0; f < n_waves; f++) sum += sD[(buf * n_waves + f) * kNpix + i];
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

// Self-marching workgroups, one or two azimuths per wavefront (APW).
template <bool EXACT_G, int APW>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 8))) void cccs_inline_kernel(
    const BeamletDev* __restrict__ beams, int n, const ConeTable* __restrict__ cones, int n_fwd,
    const float* __restrict__ terma, const float* __restrict__ dens_bev, float* __restrict__ partial) {
    extern __shared__ float smem[];
    __shared__ int s_next[2];
    cccs_body<EXACT_G, APW>(beams, n, cones, n_fwd, terma, dens_bev, partial, smem, s_next);
}

// ---------------------------------------------------------------------------------------------
// Workgroup = n_phi compute wavefronts + one march wavefront.
//
// The march of an azimuth (next lattice point, slice-end snap, midpoint voxel, edge crossings
// and the segment-length terms shared by its 256 rays) is the same for every lane of the
// wavefront that carries the azimuth, so computing it there costs a full wavefront instruction
// per operation.  Here one extra wavefront marches all azimuths at once, one per lane, one chunk
// ahead of the compute wavefronts, and hands each step over as a 32-byte record in LDS.  A
// compute wavefront then reads one record per step (two broadcast LDS reads) and spends its
// vector instructions on the per-ray work only.
//
// A chunk holds up to kRecCap steps of every azimuth and never crosses a slice end, so the one
// barrier per chunk is the existing one per slice unless an azimuth needs more than kRecCap
// steps in a slice.

struct StepRec {
    float gx, gy;      // lateral length per unit lattice offset, bs dz ls / sad, times k
    float axc, ayc;    // lateral length of the ray at offset 0, times k
    float c;           // k^2 dz^2 + s0 / 2 (k, s0: heron_record)
    uint32_t w;        // decisions, bit fields below
    int32_t gxw, gyw;  // general edge test: (fb & 15) | ((fn - fb) << 4) per axis
};
static_assert(sizeof(StepRec) == 32, "StepRec is two 16-byte LDS accesses");
// w: [0,5) shift of the midpoint voxel along x (two's complement), [5,10) along y, [10,14) mx & 15,
//    [14,18) my & 15, bit 18 + [19,23) voxel column whose rays crossed the x edge, bit 23 + [24,28)
//    voxel row whose rays crossed the y edge, bit 28: a step of more than one cell (test each ray),
//    bit 29: the previous step's column reset is done by this step's x shift (zero fill), bits 30 /
//    31: the x / y shift is one voxel.
constexpr int kRecCap = 32;       // steps per azimuth per chunk (and records per march table block)
static_assert((kRecCap & (kRecCap - 1)) == 0, "table blocks are aligned by masking");
constexpr int kHdr = kMaxPhi + 2; // per chunk: step count of each azimuth, slice done, next slice
constexpr uint32_t kColReset = 1u << 18, kRowReset = 1u << 23, kGeneralReset = 1u << 28, kZeroFillX = 1u << 29;
// bit 30 / 31: the shift along x / y is exactly one voxel (its direction is the field's sign bit, 4 / 9)
constexpr uint32_t kUnitX = 1u << 30, kUnitY = 1u << 31;
constexpr uint32_t kShiftBits = 0x3FFu, kResetBits = kColReset | kRowReset | kGeneralReset;
constexpr int kMaxMarchPhi = 10;  // azimuths of the march kernel (64 KiB of LDS); more: inline kernel
constexpr int kMaxPhiPerWave1 = 16;  // self-marching kernel: one azimuth per wavefront up to 16, then two
// A slab in LDS: A/a T, B/b T, -a log2(e) rho and -b log2(e) rho of the slice, 256 each (read by
// the compute wavefronts when the slice starts), and rho itself (read by the slice's combine).
constexpr int kSlab = 4 * kNpix;
constexpr float kLn2 = 0.6931471805599453f;

struct CcLds {
    float4* rec;   // [2][n_phi][kRecStride]: StepRec k of an azimuth at float4 2k, 2k + 1
    float4* pos;   // [n_phi][kRecCap + 1] lattice points of the chunk being marched
    int2* carry;   // [n_phi] (padded to 2 n_phi) midpoint voxel of the last step marched
    float* slab;   // [2][kSlab], slice i in buffer i & 1
    float* rho;    // [3][256], slice i in buffer i % 3 (the combine of slice i - 1 still reads its rho
                   // while slice i + 1 is being prefetched)
    float* sC;     // [2][256], slice i in buffer i & 1
    float* sD;     // [2][n_phi][256], slice i in buffer i & 1
    int* hdr;      // [2][kHdr]
};

// Records of one azimuth are kRecStride float4 apart (one float4 of padding spreads the azimuths
// over the LDS banks when the march wavefront writes them).
constexpr int kRecStride = 2 * kRecCap + 1;

__host__ __device__ constexpr size_t march_lds_bytes(int n_phi) {
    return (size_t)2 * n_phi * kRecStride * sizeof(float4) + (size_t)n_phi * (kRecCap + 1) * sizeof(float4) +
           (size_t)2 * n_phi * sizeof(int2) + (size_t)(2 * kSlab + (3 + 2 + 2 * n_phi) * kNpix) * sizeof(float) +
           (size_t)2 * kHdr * sizeof(int);
}

static_assert(march_lds_bytes(kMaxMarchPhi) <= 65536, "march kernel LDS beyond 64 KiB");
static_assert(kMaxMarchPhi <= kTabHdr - 2, "a march table chunk header holds every azimuth's count");

__device__ __forceinline__ CcLds carve_lds(float4* base, int n_phi) {
    CcLds S;
    S.rec = base;
    S.pos = base + 2 * n_phi * kRecStride;
    S.carry = reinterpret_cast<int2*>(S.pos + n_phi * (kRecCap + 1));
    float* f = reinterpret_cast<float*>(S.carry + 2 * n_phi);
    S.slab = f;
    S.rho = f + 2 * kSlab;
    S.sC = S.rho + 3 * kNpix;
    S.sD = S.sC + 2 * kNpix;
    S.hdr = reinterpret_cast<int*>(S.sD + 2 * n_phi * kNpix);
    return S;
}

// LDS of cccs_tab_kernel: two block buffers per compute wavefront (kTabRec float4: a block of
// kRecCap records of its azimuth's stream and one more, the read past a chunk's last), two chunk
// headers, then the slabs, rho, cumulative length and per-azimuth doses as in carve_lds.  Block
// b + 1 is loaded into LDS while the chunks of block b run; chunk c's header during chunk c - 1.
constexpr int kTabRec = 2 * (kRecCap + 1);
__host__ __device__ constexpr size_t tab_lds_bytes(int n_phi) {
    return (size_t)2 * n_phi * kTabRec * sizeof(float4) + (size_t)2 * kTabHdr * sizeof(int) +
           (size_t)(2 * kSlab + (3 + 2 + 2 * n_phi) * kNpix) * sizeof(float);
}

__device__ __forceinline__ CcLds carve_lds_tab(float4* base, int n_phi) {
    CcLds S{};
    S.rec = base;
    S.hdr = reinterpret_cast<int*>(base + 2 * n_phi * kTabRec);
    float* f = reinterpret_cast<float*>(S.hdr + 2 * kTabHdr);
    S.slab = f;
    S.rho = f + 2 * kSlab;
    S.sC = S.rho + 3 * kNpix;
    S.sD = S.sC + 2 * kNpix;
    return S;
}

// What every wavefront of the workgroup shares: the beamlet's slabs and output, thread layout.
struct CcBlock {
    const float* Tg;
    const float* Rg;
    float* P;
    int L, tid, nthreads, n_phi;
    bool two, prefetch;
    float A_a, B_b, na, nb;  // cone constants the slab is scaled with
};

// Slab element e = 64 q + lane (q wave-uniform) of slice z: Terma for q < 4, density above, so
// the source pointer is a scalar and only the lane offset is a vector register.
__device__ __forceinline__ float slab_value(const CcBlock& K, int z, int q) {
    const float* src = (q < 4 ? K.Tg : K.Rg) + (uint32_t)z * kNpix + (q & 3) * kWave;
    return src[threadIdx.x & 63];
}
// Slab element e (e < 256: Terma, else density; wave-uniform) of slice ordinal si into LDS.
__device__ __forceinline__ void store_slab(const CcBlock& K, const CcLds& S, int si, int e, float v) {
    float* slab = S.slab + (si & 1) * kSlab;
    if (e < kNpix) {
        slab[e] = K.A_a * v;
        slab[kNpix + e] = K.B_b * v;
    } else {
        slab[kNpix + e] = K.na * v;  // 2 kNpix + (e - kNpix)
        slab[2 * kNpix + e] = K.nb * v;
        S.rho[(si % 3) * kNpix + e - kNpix] = v;
    }
}
__device__ __forceinline__ void slab_regs(const CcBlock& K, int z, float& pf0, float& pf1) {
    const int q0 = K.tid >> 6, q1 = (K.tid + K.nthreads) >> 6;  // wave-uniform
    if (q0 < 8) pf0 = slab_value(K, z, __builtin_amdgcn_readfirstlane(q0));
    if (K.two && q1 < 8) pf1 = slab_value(K, z, __builtin_amdgcn_readfirstlane(q1));
}
// End of a chunk, executed by every wavefront: store the prefetched slab, barrier, and after a
// slice end the combine of the slice into P.  Returns false once the beamlet is finished.
// DMA: this wavefront may have loads into LDS in flight (cccs_tab_kernel's next chunk header or
// record block), complete before the barrier (vector memory operations complete in issue order).
__device__ __forceinline__ bool chunk_end(const CcBlock& K, const CcLds& S, int done, int sn, float pf0,
                                          float pf1, int& s, int& si, int& written, bool dma = false) {
    if (done && sn < K.L && K.prefetch) {
        const int i0 = K.tid, i1 = K.tid + K.nthreads;
        if (i0 < 2 * kNpix) store_slab(K, S, si + 1, i0, pf0);
        if (K.two && i1 < 2 * kNpix) store_slab(K, S, si + 1, i1, pf1);
    }
    if (dma) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (!done) return true;
    // slice s: rho sum over azimuths (in azimuth order) of the dose in units of geometric length /
    // (rho cumulative length of the last azimuth + 1e-4); the quotient from the hardware reciprocal
    for (int i = K.tid; i < kNpix; i += K.nthreads) {
        const float* d = S.sD + (si & 1) * K.n_phi * kNpix + i;
        float sum = 0.0f;
        int f = 0;
        for (; f + 4 <= K.n_phi; f += 4) {  // four reads in flight, summed in azimuth order
            const float v0 = d[f * kNpix], v1 = d[(f + 1) * kNpix], v2 = d[(f + 2) * kNpix], v3 = d[(f + 3) * kNpix];
            sum = (((sum + v0) + v1) + v2) + v3;
        }
        for (; f < K.n_phi; f++) sum += d[f * kNpix];
        const float rho = S.rho[(si % 3) * kNpix + i];
        K.P[(uint32_t)s * kNpix + i] = sum * rho * __builtin_amdgcn_rcpf(fmaf(S.sC[(si & 1) * kNpix + i], rho, kEps));
    }
    for (int z = s + 1; z < sn && z < K.L; z++)  // slices the phi-0 march stepped over
        for (int i = K.tid; i < kNpix; i += K.nthreads) K.P[(uint32_t)z * kNpix + i] = 0.0f;
    written = sn < K.L ? sn : K.L;
    if (sn >= K.L) return false;
    if (!K.prefetch) {  // small workgroups load the next slab here
        __syncthreads();
        for (int i = K.tid; i < 2 * kNpix; i += K.nthreads)
            store_slab(K, S, si + 1, i, i < kNpix ? K.Tg[(uint32_t)sn * kNpix + i] : K.Rg[(uint32_t)sn * kNpix + i - kNpix]);
        __syncthreads();
    }
    s = sn;
    si++;
    return true;
}

__device__ __forceinline__ void zero_tail(const CcBlock& K, int written) {
    for (int z = written; z < K.L; z++)
        for (int i = K.tid; i < kNpix; i += K.nthreads) K.P[(uint32_t)z * kNpix + i] = 0.0f;
}

// March wavefront, in two phases per chunk.  Phase 1: lane f < n_phi marches azimuth f with the
// FP32 operation order of the CPU oracle (d_nextPoint, DoseComp.cu:76-111) and only records the
// lattice points it reaches.  Phase 2: every lane takes one (azimuth, step) pair and derives the
// step's record from its two end points (and the one before, for the previous midpoint voxel), so
// the decision and segment arithmetic runs 64 steps per instruction instead of n_phi.
struct MarchLane {
    float dx, dy, dz, ix_, iy_, iz_, sgx, sgy, sgz;
    float bxp, byp, bzp;
    int nsteps;
    bool act;
};

struct MarchConst {
    float bsx, bsy, ls, lim_min, inv_sad;
    int lane, kper, lf, lk;  // phase 2: lane -> azimuth lf, step lk of each group of kper steps
};

// Where phase 2 puts a chunk's records: the march kernel's LDS records (TABLE false), or a march
// table's per-azimuth streams in global memory (TABLE true; geometry terms only, see TabOut).
struct TabOut {
    float4* rec;       // azimuth 0's stream
    uint32_t cap;      // records per azimuth
    int* ofs;          // LDS: records of each azimuth written before this chunk; set to the chunk's start
};

// Returns false (TABLE only) when a table stream cannot take the chunk; nothing is written then.
template <bool TABLE>
__device__ __forceinline__ bool march_chunk(MarchLane& M, const MarchConst& C, int n_phi, const CcLds& S,
                                            float4* rec, int* hdr, int s_cur, bool new_slice,
                                            const TabOut& T = TabOut{}) {
    const int lane = C.lane;
    if (new_slice) {
        M.act = lane < n_phi;
        M.nsteps = 0;
    }
    // phase 1: lattice points; pos[f][0] is where the chunk starts
    float4* mypos = S.pos + (lane < n_phi ? lane : 0) * (kRecCap + 1);
    if (lane < n_phi) mypos[0] = make_float4(M.bxp, M.byp, M.bzp, 0.0f);
    int cnt = 0, k = 0;
    for (; k < kRecCap; k++) {
        if (__ballot(M.act) == 0) break;
        if (M.act) {
            const float sx = axis_step(M.bxp, M.sgx, M.ix_);
            const float sy = axis_step(M.byp, M.sgy, M.iy_);
            const float sz = axis_step(M.bzp, M.sgz, M.iz_);
            const float st = fminf(fminf(sx, sy), sz);
            float nx = fmaf(st, M.dx, M.bxp), ny = fmaf(st, M.dy, M.byp), nz = fmaf(st, M.dz, M.bzp);
            // rint == round here: a tie (.5) is never within 1e-4 of an integer
            const bool flag = fabsf(nz - rintf(nz)) < kEps;
            if (flag) {
                nx = fmaf(sz, M.dx, M.bxp);
                ny = fmaf(sz, M.dy, M.byp);
                nz = rintf(fmaf(sz, M.dz, M.bzp));
            }
            mypos[k + 1] = make_float4(nx, ny, nz, 0.0f);
            M.bxp = nx;
            M.byp = ny;
            M.bzp = nz;
            cnt = k + 1;
            M.nsteps++;
            if (flag || M.nsteps >= kMaxStepsPerSlice) M.act = false;
        }
    }
    const bool done = __ballot(M.act) == 0;
    if (lane < n_phi) hdr[lane] = cnt;
    if (lane == 0) {
        hdr[kMaxPhi] = done ? 1 : 0;
        hdr[kMaxPhi + 1] = max((int)rintf(M.bzp), s_cur + 1);  // the phi-0 march decides the next slice
    }
    if (TABLE) {
        // a chunk's records never straddle a block of kRecCap records of the stream: one that does
        // not fit in the rest of the current block starts the next one (the compute wavefronts load
        // and finish the stream block by block)
        bool over = false;
        if (lane < n_phi) {
            uint32_t o = (uint32_t)T.ofs[lane];
            if ((o & (kRecCap - 1)) + (uint32_t)cnt > (uint32_t)kRecCap)
                o = (o + kRecCap - 1) & ~(uint32_t)(kRecCap - 1);
            over = o + (uint32_t)cnt > T.cap;
            T.ofs[lane] = (int)o;
        }
        if (__ballot(over) != 0) return false;
    }
    // phase 2: records; k is the largest step count of the chunk
    const bool mine = C.lf < n_phi;
    const int f = mine ? C.lf : 0;
    const int c = hdr[f];
    const float4* P = S.pos + f * (kRecCap + 1);
    float4* out = rec + f * kRecStride;
    for (int k0 = 0; k0 < k; k0 += C.kper) {
        const int j = k0 + C.lk;
        if (!mine || j >= c) continue;
        const float4 p0 = P[j], p1 = P[j + 1];
        const int mx = (int)floorf((p0.x + p1.x) * 0.5f);
        const int my = (int)floorf((p0.y + p1.y) * 0.5f);
        int mxp, myp;  // midpoint voxel of the previous step
        if (j > 0) {
            const float4 pm = P[j - 1];
            mxp = (int)floorf((pm.x + p0.x) * 0.5f);
            myp = (int)floorf((pm.y + p0.y) * 0.5f);
        } else {
            const int2 cr = S.carry[f];
            mxp = cr.x;
            myp = cr.y;
        }
        const int ddx = ((mx - mxp + 8) & 15) - 8;
        const int ddy = ((my - myp + 8) & 15) - 8;
        // segment length terms shared by all rays: physical lateral offset of a ray at lattice
        // coordinate c is (c - 8) bs f(z), f(z) = (z ls + lim_min) / sad (unwrapped)
        const float dzs = (p1.z - p0.z) * C.ls;
        const float fe = fmaf(p1.z, C.ls, C.lim_min) * C.inv_sad;
        const float df = dzs * C.inv_sad;
        const float gx = C.bsx * df, gy = C.bsy * df;
        const float axc = fmaf(gx, p0.x - 8.0f, C.bsx * (p1.x - p0.x) * fe);
        const float ayc = fmaf(gy, p0.y - 8.0f, C.bsy * (p1.y - p0.y) * fe);
        const int fbx = (int)floorf(p0.x), fby = (int)floorf(p0.y);
        const int fnx = (int)floorf(p1.x), fny = (int)floorf(p1.y);
        uint32_t w = (uint32_t)(ddx & 31) | ((uint32_t)(ddy & 31) << 5) | ((uint32_t)(mx & 15) << 10) |
                     ((uint32_t)(my & 15) << 14) | (ddx == 1 || ddx == -1 ? kUnitX : 0u) |
                     (ddy == 1 || ddy == -1 ? kUnitY : 0u);
        // rays that cross the lateral BEV edge restart with X = 0 (DoseComp.cu:372-377); a step
        // of at most one cell per axis crosses with one ray column and one ray row at most
        if (abs(fnx - fbx) <= 1 && abs(fny - fby) <= 1) {
            const int cxo = crossing_offset(fbx, fnx), cyo = crossing_offset(fby, fny);
            if (cxo >= 0) {
                // The column's X is zeroed after this step.  When the next step shifts the state by
                // one voxel and moves exactly this column across the row end, that shift writes the
                // zero instead (the state is not read in between), so the reset is left to it.
                const int col = (mx + cxo) & 15;
                bool fuse = false;
                if (j + 1 < c) {
                    const float4 p2 = P[j + 2];
                    const int dn = (((int)floorf((p1.x + p2.x) * 0.5f) - mx + 8) & 15) - 8;
                    fuse = (dn == 1 && col == 15) || (dn == -1 && col == 0);
                }
                if (!fuse) w |= kColReset | ((uint32_t)col << 19);
            }
            if (cyo >= 0) w |= kRowReset | ((uint32_t)((my + cyo) & 15) << 24);
        } else {
            w |= kGeneralReset;
        }
        if (j > 0 && (ddx == 1 || ddx == -1)) {
            // the same decision for the previous step, made by the lane that wrote its record
            const float4 pm = P[j - 1];
            const int pbx = (int)floorf(pm.x), pby = (int)floorf(pm.y);
            if (abs(fbx - pbx) <= 1 && abs(fby - pby) <= 1) {
                const int pxo = crossing_offset(pbx, fbx);
                if (pxo >= 0) {
                    const int pcol = (mxp + pxo) & 15;
                    if ((ddx == 1 && pcol == 15) || (ddx == -1 && pcol == 0)) w |= kZeroFillX;
                }
            }
        }
        const int gxw = (fbx & 15) | ((fnx - fbx) << 4);
        const int gyw = (fby & 15) | ((fny - fby) << 4);
        if (TABLE) {
            // the beamlet terms (lim_min, sad) are applied by the compute wavefronts: the lengths
            // and the x, y offsets of the step as above, before they meet lim_min and 1 / sad
            const uint32_t idx = (uint32_t)T.ofs[f] + (uint32_t)j;  // < T.cap (checked above)
            float4* o = T.rec + 2 * ((uint64_t)f * T.cap + idx);
            o[0] = make_float4(dzs, p1.z, p0.x - 8.0f, p0.y - 8.0f);
            o[1] = make_float4(C.bsx * (p1.x - p0.x), C.bsy * (p1.y - p0.y), __uint_as_float(w),
                               __int_as_float((int)(((uint32_t)gxw & 0xFFFFu) | ((uint32_t)gyw << 16))));
        } else {
            float hc;
            out[2 * j] = heron_record(gx, gy, axc, ayc, dzs * dzs, hc);
            out[2 * j + 1] = make_float4(hc, __uint_as_float(w), __int_as_float(gxw), __int_as_float(gyw));
        }
        if (j == c - 1) S.carry[f] = make_int2(mx, my);  // after every read of this chunk's carry
    }
    return true;
}

// March state of this lane (azimuth lane, if below n_phi) at the start of the walk of cone t, for a
// beamlet with BEV pitch (bs0, bs1) and longitudinal spacing ls; phase-2 lane assignment.
__device__ __forceinline__ void march_init(MarchLane& M, MarchConst& C, const ConeTable* __restrict__ cones, int t,
                                           int n_phi, float bs0, float bs1, float ls) {
    const int lane = threadIdx.x & 63;
    const int f = lane < n_phi ? lane : 0;
    {
        // grid direction of (theta, phi_f): unit vector scaled to voxel units, renormalised
        const float ux = cones->ux[t][f], uy = cones->uy[t][f], uz = cones->uz[t][f];
        const float vxd = ux / bs0, vyd = uy / bs1, vzd = uz / ls;
        const float rn = 1.0f / sqrtf(vxd * vxd + vyd * vyd + vzd * vzd);
        M.dx = vxd * rn;
        M.dy = vyd * rn;
        M.dz = vzd * rn;
        M.ix_ = 1.0f / M.dx;
        M.iy_ = 1.0f / M.dy;
        M.iz_ = 1.0f / M.dz;
        M.sgx = M.dx > 0.0f ? 1.0f : -1.0f;
        M.sgy = M.dy > 0.0f ? 1.0f : -1.0f;
        M.sgz = M.dz > 0.0f ? 1.0f : -1.0f;
    }
    M.bxp = 0.5f;
    M.byp = 0.5f;
    M.bzp = 0.0f;
    M.nsteps = 0;
    M.act = false;
    C.bsx = uniform(bs0);
    C.bsy = uniform(bs1);
    C.ls = uniform(ls);
    C.lane = lane;
    C.kper = 64 / n_phi;
    C.lk = lane / n_phi;
    C.lf = C.lk < C.kper ? lane - C.lk * n_phi : n_phi;  // n_phi: idle in phase 2
}

__device__ __forceinline__ void cccs_march(const BeamletDev& B, const ConeTable* __restrict__ cones, int t,
                                           const CcBlock& K, const CcLds& S) {
    const int lane = threadIdx.x & 63;
    MarchLane M;
    MarchConst C;
    march_init(M, C, cones, t, K.n_phi, B.bs[0], B.bs[1], B.ls);
    C.inv_sad = uniform(1.0f / B.sad);
    C.lim_min = uniform(B.lim_min);
    if (lane < K.n_phi) S.carry[lane] = make_int2(0, 0);  // the midpoint voxel starts at (0, 0)
    const int n_phi = K.n_phi;

    march_chunk<false>(M, C, K.n_phi, S, S.rec, S.hdr, 0, true);
    __syncthreads();
    int s = 0, buf = 0, si = 0, written = 0;  // si: ordinal of the slice being processed
    for (;;) {
        const int done = S.hdr[buf * kHdr + kMaxPhi], sn = S.hdr[buf * kHdr + kMaxPhi + 1];
        float pf0 = 0.0f, pf1 = 0.0f;
        if (done && sn < K.L && K.prefetch) slab_regs(K, sn, pf0, pf1);
        if (!(done && sn >= K.L))
            march_chunk<false>(M, C, n_phi, S, S.rec + (buf ^ 1) * n_phi * kRecStride, S.hdr + (buf ^ 1) * kHdr,
                               done ? sn : s, done != 0);
        if (!chunk_end(K, S, done, sn, pf0, pf1, s, si, written)) break;
        buf ^= 1;
    }
    zero_tail(K, written);
}

// March tables: one wavefront per (march geometry, forward cone) walks the lattice of every azimuth
// over the geometry's longest beamlet, chunk by chunk exactly as the march wavefront of
// cccs_kernel does, and writes the chunk headers and geometry records that the compute wavefronts
// of every beamlet of that geometry read (cccs_tab_kernel).  A stream or header array too small
// for the walk sets *overflow and ends the table early; the caller reports it.
__host__ __device__ constexpr size_t table_lds_bytes(int n_phi) {
    return (size_t)n_phi * (kRecCap + 1) * sizeof(float4) + (size_t)n_phi * sizeof(int2) +
           (size_t)(kHdr + kWave) * sizeof(int);
}

__global__ __launch_bounds__(64) void cccs_table_kernel(const TabEntry* __restrict__ tab,
                                                        const ConeTable* __restrict__ cones, int n_fwd, int n_phi,
                                                        float4* __restrict__ rec, int* __restrict__ hdr_g,
                                                        uint32_t* __restrict__ overflow) {
    extern __shared__ float4 lds4[];
    const int t = blockIdx.x % n_fwd;
    const TabEntry E = tab[blockIdx.x];
    CcLds S{};
    S.pos = lds4;
    S.carry = reinterpret_cast<int2*>(S.pos + n_phi * (kRecCap + 1));
    int* hdr = reinterpret_cast<int*>(S.carry + n_phi);
    int* ofs = hdr + kHdr;  // records of each azimuth written so far
    const int lane = threadIdx.x;
    MarchLane M;
    MarchConst C;
    march_init(M, C, cones, t, n_phi, E.bs0, E.bs1, E.ls);
    if (lane < n_phi) S.carry[lane] = make_int2(0, 0);
    ofs[lane] = 0;
    int* H = hdr_g + E.hdr_base;
    const TabOut T{rec + E.rec_base, E.rec_cap, ofs};
    uint32_t c = 0;
    int s = 0;
    bool first = true, ok = true;
    for (;;) {
        // the chunk sequence of cccs_march: the next slice is decided by the phi-0 march
        int s_cur = 0;
        bool new_slice = true;
        if (!first) {
            const int done = hdr[kMaxPhi], sn = hdr[kMaxPhi + 1];
            if (done && sn >= E.Lmax) break;
            if (done) s = sn;
            s_cur = s;
            new_slice = done != 0;
        }
        first = false;
        if (c + 1 >= E.chunk_cap || !march_chunk<true>(M, C, n_phi, S, nullptr, hdr, s_cur, new_slice, T)) {
            ok = false;
            break;
        }
        int* h = H + (uint64_t)c * kTabHdr;
        if (lane < n_phi) {
            h[lane] = hdr[lane];
            ofs[lane] += hdr[lane];
        }
        if (lane == 0) {
            h[kTabHdr - 2] = hdr[kMaxPhi];
            h[kTabHdr - 1] = hdr[kMaxPhi + 1];
        }
        c++;
    }
    // the terminating chunk: no steps, slice done, next slice beyond every beamlet
    int* h = H + (uint64_t)(c < E.chunk_cap ? c : E.chunk_cap - 1) * kTabHdr;
    if (lane < kTabHdr) h[lane] = lane == kTabHdr - 2 ? 1 : (lane == kTabHdr - 1 ? (1 << 30) : 0);
    if (!ok && lane == 0) overflow[0] = 1u;
}

// What a compute wavefront of cccs_tab_kernel reads its chunks from: the chunk headers of its
// (geometry, cone) table, its azimuth's record stream, and the beamlet terms the geometry records
// are finished with.
struct TabRead {
    const int* hdr;      // chunk 0's header
    const float4* rec;   // this azimuth's stream
    uint32_t cap;        // records in it
    uint32_t chunks;     // chunk headers (the walk ends at the terminating one before them)
    float lim_min, inv_sad, bsx, bsy, ls;
};

// 16 bytes per lane from global address base + off (base wave-uniform, off < 2^32 per lane) into
// LDS at lds + 16 lane (global_load_lds_dwordx4; lds wave-uniform).  Written as asm so that the
// compiler, which does not see the LDS write, adds no wait for it before the wavefront's other LDS
// reads; the caller waits (vmcnt) before the barrier after which the data is read.  The
// instruction has no register result.  m0 (the LDS address) is set in the same asm statement; no
// instruction the compiler emits in these kernels reads m0 (no LDS DMA, s_movrel or message of its
// own), so the clobber of a register the compiler reserves is safe here.
__device__ __forceinline__ void load_to_lds16(const void* base, uint32_t off, void* lds) {
    const uint32_t a = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void*)lds;
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(off), "s"(base), "s"(a) : "m0");
}

// Header of chunk `chunk` of a march table into LDS at `hdr`, by wavefront 0 (4 lanes x 16 bytes).
__device__ __forceinline__ void tab_load_header(const TabRead& TR, uint32_t chunk, int* hdr, int wave) {
    const uint32_t lane = threadIdx.x & 63;
    if (wave == 0 && lane < 4)
        load_to_lds16(TR.hdr + (uint64_t)min(chunk, TR.chunks - 1) * kTabHdr, 16 * lane, hdr);
}
// Block `blk` of this azimuth's stream (records kRecCap blk ... kRecCap (blk + 1) - 1; 64 lanes x
// 16 bytes, indices clamped into the stream: what lies past its last record is never read) into
// LDS at `rec`.
__device__ __forceinline__ void tab_load_block(const TabRead& TR, uint32_t blk, float4* rec) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t i = min(2 * kRecCap * blk + lane, 2 * TR.cap - 1);
    load_to_lds16(TR.rec, 16 * i, rec);
}
static_assert(2 * kRecCap == kWave, "a table block is one 16-byte load per lane");

// Compute wavefront of azimuth `wave`: lane l owns voxels (x = l & 15, y = 4 (l >> 4) + r).  The
// X states are kept pre-scaled by A/a and B/b, and the dose and cumulative length are summed in
// units of the geometric length s (p = s rho), multiplied by rho in the combine.  TAB: the chunks
// come from a march table (cccs_tab_kernel), else from the march wavefront's LDS records.
template <bool EXACT_G, bool CUMU, bool TAB = false>
__device__ __forceinline__ void cccs_compute(const CcBlock& K, const CcLds& S, int wave,
                                             const TabRead& TR = TabRead{}) {
    const int lane = threadIdx.x & 63;
    const int n_phi = K.n_phi;

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
    // ray offsets of my voxels relative to the midpoint voxel (mx, my), starting at (0, 0):
    // fix = (x - mx) & 15, fiy[r] = (4 g + r - my) & 15.  An offset belongs to the ray, so it moves
    // with the X state when the midpoint voxel moves (a shift by one voxel subtracts 1 mod 16)
    float fix = (float)vx;
    float fiy[kSlots];
#pragma unroll
    for (int r = 0; r < kSlots; r++) fiy[r] = (float)(4 * g + r);

    if (TAB) {
        tab_load_header(TR, 0, S.hdr, wave);
        tab_load_block(TR, 0, S.rec + wave * kTabRec);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __syncthreads();  // chunk 0's header, block 0 and slab 0 are in LDS
    int s = 0, buf = 0, si = 0, written = 0;  // si: ordinal of the slice being processed
    // TAB: chunk, stream position after the previous chunk, block finished last (-1: none)
    uint32_t tc = 0, tofs = 0;
    int blk = -1;
    bool fresh = true;
    float TA[kSlots], TB[kSlots], nRa[kSlots], nRb[kSlots], D[kSlots], Cs[kSlots];
    for (;;) {
        int done, sn, cnt;
        const float4* rp;
        if (TAB) {
            // this chunk's header is in LDS (loaded during the previous chunk).  My azimuth's records
            // of the chunk lie in one block of the stream (the table's placement rule, repeated
            // here); a chunk that starts a new block finishes the whole block in place with the
            // beamlet terms (the arithmetic of the march wavefront's phase 2, same operation order;
            // the block arrived during an earlier chunk) and starts loading the next block into the
            // other buffer, which the previous block no longer needs.
            const int* h = S.hdr + (tc & 1) * kTabHdr;
            done = h[kTabHdr - 2];
            sn = h[kTabHdr - 1];
            cnt = min(max(__builtin_amdgcn_readfirstlane(h[wave]), 0), kRecCap);
            uint32_t o = tofs;
            if ((o & (kRecCap - 1)) + (uint32_t)cnt > (uint32_t)kRecCap)
                o = (o + kRecCap - 1) & ~(uint32_t)(kRecCap - 1);
            const int ob = (int)(o / kRecCap);
            float4* my = S.rec + ((ob & 1) * n_phi + wave) * kTabRec;
            if (cnt > 0 && ob != blk) {
                if (lane < kRecCap) {
                    const float4 g0 = my[2 * lane], g1 = my[2 * lane + 1];
                    // g0 = (dzs, z1, x0 - 8, y0 - 8), g1 = (bsx dx, bsy dy, w, gx | gy << 16)
                    const float fe = fmaf(g0.y, TR.ls, TR.lim_min) * TR.inv_sad;
                    const float df = g0.x * TR.inv_sad;
                    const float gx = TR.bsx * df, gy = TR.bsy * df;
                    const int gxy = __float_as_int(g1.w);
                    float hc;
                    my[2 * lane] = heron_record(gx, gy, fmaf(gx, g0.z, g1.x * fe), fmaf(gy, g0.w, g1.y * fe),
                                                g0.x * g0.x, hc);
                    my[2 * lane + 1] = make_float4(hc, g1.z, __int_as_float((int)(short)(gxy & 0xFFFF)),
                                                   __int_as_float(gxy >> 16));
                }
                blk = ob;
                tab_load_block(TR, (uint32_t)ob + 1, S.rec + (((ob + 1) & 1) * n_phi + wave) * kTabRec);
            }
            rp = my + 2 * (o & (kRecCap - 1));
            tofs = o + (uint32_t)cnt;
            tc++;
            tab_load_header(TR, tc, S.hdr + (tc & 1) * kTabHdr, wave);
        } else {
            done = S.hdr[buf * kHdr + kMaxPhi];
            sn = S.hdr[buf * kHdr + kMaxPhi + 1];
            cnt = __builtin_amdgcn_readfirstlane(S.hdr[buf * kHdr + wave]);
            rp = S.rec + (buf * n_phi + wave) * kRecStride;
        }
        float pf0 = 0.0f, pf1 = 0.0f;
        if (done && sn < K.L && K.prefetch) slab_regs(K, sn, pf0, pf1);
        if (fresh) {
            const float* slab = S.slab + (si & 1) * kSlab + vbase;
#pragma unroll
            for (int r = 0; r < kSlots; r++) {
                TA[r] = slab[16 * r];
                TB[r] = slab[kNpix + 16 * r];
                nRa[r] = slab[2 * kNpix + 16 * r];
                nRb[r] = slab[3 * kNpix + 16 * r];
                D[r] = 0.0f;
                Cs[r] = 0.0f;
            }
        }
        const float2* rp2 = reinterpret_cast<const float2*>(rp);  // (c, w) of record k at 4 k + 2
        // one step from its record (ra, rb)
        auto step = [&](const float4& ra, const float2& rb, int k) {
            const uint32_t w = __builtin_amdgcn_readfirstlane(__float_as_uint(rb.y));
            const int ddx = ((int)(w << 27)) >> 27;
            const int ddy = ((int)(w << 22)) >> 27;
            // steps with no shift of the midpoint voxel (and with no reset below) skip the decision
            // trees with one scalar test each
            if (w & kShiftBits) {
                // rotate the X state (and my rays' x offsets) to follow the midpoint voxel
                if (w & kUnitX) {
                    if (w & 0x10u) {  // ddx = -1
                        if (w & kZeroFillX) {  // with the previous step's reset of column 0
#pragma unroll
                            for (int r = 0; r < kSlots; r++) {
                                XA[r] = row_from_right_zero(XA[r]);
                                XB[r] = row_from_right_zero(XB[r]);
                            }
                        } else {
#pragma unroll
                            for (int r = 0; r < kSlots; r++) {
                                XA[r] = row_from_right(XA[r]);
                                XB[r] = row_from_right(XB[r]);
                            }
                        }
                        fix = row_from_right(fix);
                    } else {  // ddx = +1
                        if (w & kZeroFillX) {  // with the previous step's reset of column 15
#pragma unroll
                            for (int r = 0; r < kSlots; r++) {
                                XA[r] = row_from_left_zero(XA[r]);
                                XB[r] = row_from_left_zero(XB[r]);
                            }
                        } else {
#pragma unroll
                            for (int r = 0; r < kSlots; r++) {
                                XA[r] = row_from_left(XA[r]);
                                XB[r] = row_from_left(XB[r]);
                            }
                        }
                        fix = row_from_left(fix);
                    }
                } else if (w & 0x1Fu) {
                    shift_x(XA, XB, ddx);
                    fix = (float)((vx - (int)((w >> 10) & 15)) & 15);
                }
                if (w & kUnitY) {
                    if (w & 0x200u) {  // ddy = -1
                        const float ta = wave_from(XA[0], down_lane);
                        const float tb = wave_from(XB[0], down_lane);
                        const float tz = wave_from(fiy[0], down_lane);
                        slots_up(XA, ta);
                        slots_up(XB, tb);
                        slots_up(fiy, tz);
                    } else {  // ddy = +1
                        const float ta = wave_from(XA[kSlots - 1], up_lane);
                        const float tb = wave_from(XB[kSlots - 1], up_lane);
                        const float tz = wave_from(fiy[kSlots - 1], up_lane);
                        slots_down(XA, ta);
                        slots_down(XB, tb);
                        slots_down(fiy, tz);
                    }
                } else if (w & 0x3E0u) {
                    shift_y3(XA, XB, fiy, ddy, up_lane, down_lane);
                }
            }
            // geometric segment length of each ray by the Heron step of heron_record (the record's
            // ray terms are scaled by k and rb.x is its constant c)
            const float ddxl = fmaf(ra.x, fix, ra.z);
            const float dxz2 = fmaf(ddxl, ddxl, rb.x);
#pragma unroll
            for (int r = 0; r < kSlots; r++) {
                const float ddyl = fmaf(ra.y, fiy[r], ra.w);
                const float sl = fmaf(ddyl, ddyl, dxz2);
                // e^{-a p}, e^{-b p}, p = rho s
                const float ea = __builtin_amdgcn_exp2f(nRa[r] * sl);
                const float eb = __builtin_amdgcn_exp2f(nRb[r] * sl);
                if (EXACT_G) {
                    // a p = a rho s = -ln(2) (-a log2(e) rho s)
                    const float ap = -kLn2 * (nRa[r] * sl), bp = -kLn2 * (nRb[r] * sl);
                    const float ga = ap > 0.0f ? (1.0f - ea) * __builtin_amdgcn_rcpf(ap) : 1.0f;
                    const float gb = bp > 0.0f ? (1.0f - eb) * __builtin_amdgcn_rcpf(bp) : 1.0f;
                    D[r] = fmaf(fmaf(ga, XA[r] - TA[r], TA[r]) + fmaf(gb, XB[r] - TB[r], TB[r]), sl, D[r]);
                } else {
                    D[r] = fmaf(XA[r] + XB[r], sl, D[r]);
                }
                if (CUMU) Cs[r] += sl;
                // X <- T (1 - e) + e X = T - e (T - X)
                XA[r] = fmaf(-ea, TA[r] - XA[r], TA[r]);
                XB[r] = fmaf(-eb, TB[r] - XB[r], TB[r]);
            }
            // rays that crossed the lateral BEV edge restart with X = 0 (DoseComp.cu:372-377)
            if (w & kResetBits) {
                if (w & kColReset) {
                    const bool m = vx == (int)((w >> 19) & 15);
#pragma unroll
                    for (int r = 0; r < kSlots; r++) {
                        XA[r] = m ? 0.0f : XA[r];
                        XB[r] = m ? 0.0f : XB[r];
                    }
                }
                if (w & kRowReset) {
                    const int yrow = (int)((w >> 24) & 15);
                    const bool m = g == (yrow >> 2);
                    const int rr = yrow & 3;
#pragma unroll
                    for (int r = 0; r < kSlots; r++)
                        if (r == rr) {
                            XA[r] = m ? 0.0f : XA[r];
                            XB[r] = m ? 0.0f : XB[r];
                        }
                }
                if (w & kGeneralReset) {
                    const int mx = (int)((w >> 10) & 15), my = (int)((w >> 14) & 15);
                    const float2 gw = rp2[4 * k + 3];
                    const int gxw = __builtin_amdgcn_readfirstlane(__float_as_int(gw.x));
                    const int gyw = __builtin_amdgcn_readfirstlane(__float_as_int(gw.y));
                    const int fbx = gxw & 15, fnx = fbx + (gxw >> 4);
                    const int fby = gyw & 15, fny = fby + (gyw >> 4);
                    const int ix = (vx - mx) & 15;
                    const bool cx = ((fbx + ix) >> 4) != ((fnx + ix) >> 4);
#pragma unroll
                    for (int r = 0; r < kSlots; r++) {
                        const int iy = (4 * g + r - my) & 15;
                        const bool c = cx || (((fby + iy) >> 4) != ((fny + iy) >> 4));
                        XA[r] = c ? 0.0f : XA[r];
                        XB[r] = c ? 0.0f : XB[r];
                    }
                }
            }
        };
        // two record register sets in turn, the next record read while a step computes.  The read
        // one record past the chunk's last (record kRecCap at most: the azimuth's padding float4 and
        // the first float4 after it, both inside this workgroup's LDS) is never used.
        const float4* q = rp;
        float4 ra = q[0], qa;
        float2 rb = reinterpret_cast<const float2*>(q)[2], qb;
        for (int k = 0; k < cnt; k += 2) {
            qa = q[2];
            qb = reinterpret_cast<const float2*>(q + 2)[2];
            step(ra, rb, k);
            if (k + 1 >= cnt) break;
            ra = q[4];
            rb = reinterpret_cast<const float2*>(q + 4)[2];
            step(qa, qb, k + 1);
            q += 4;
        }
        if (done) {  // publish this azimuth's slice dose (and the last azimuth's cumulative length)
            float* myD = S.sD + ((si & 1) * n_phi + wave) * kNpix;
#pragma unroll
            for (int r = 0; r < kSlots; r++) {
                myD[vbase + 16 * r] = D[r];
                if (CUMU) S.sC[(si & 1) * kNpix + vbase + 16 * r] = Cs[r];
            }
        }
        if (!chunk_end(K, S, done, sn, pf0, pf1, s, si, written, TAB)) break;
        fresh = done != 0;
        buf ^= 1;
    }
    zero_tail(K, written);
}

// 72 VGPRs (7 wavefronts per SIMD) and 49.6 KiB of LDS fit three workgroups of 9 wavefronts on a CU
// (the few values the compiler then spills live outside the step loop); 80 VGPRs fit two.
template <bool EXACT_G>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7, 8))) void cccs_kernel(
    const BeamletDev* __restrict__ beams, int n, const ConeTable* __restrict__ cones, int n_fwd,
    const float* __restrict__ terma, const float* __restrict__ dens_bev, float* __restrict__ partial) {
    extern __shared__ float4 lds4[];
    const int t = blockIdx.x / n;  // forward-cone slot, widest cone first
    const int b = blockIdx.x - t * n;
    const BeamletDev& B = beams[b];
    CcBlock K;
    K.L = B.L;
    if (K.L <= 0) return;
    K.tid = threadIdx.x;
    K.nthreads = blockDim.x;
    K.n_phi = (blockDim.x >> 6) - 1;
    K.Tg = terma + B.bev_off;
    K.Rg = dens_bev + B.bev_off;
    K.P = cone_partial(partial, B, n_fwd, t);
    K.two = K.nthreads < 2 * kNpix;  // each thread prefetches two slab values
    K.prefetch = K.nthreads >= kNpix;  // smaller workgroups reload each slab
    const CcLds S = carve_lds(lds4, K.n_phi);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    K.A_a = cones->A_a[t];
    K.B_b = cones->B_b[t];
    K.na = cones->na[t];
    K.nb = cones->nb[t];
    // slab of slice 0 into buffer 0
    for (int i = K.tid; i < 2 * kNpix; i += K.nthreads) store_slab(K, S, 0, i, i < kNpix ? K.Tg[i] : K.Rg[i - kNpix]);
    if (wave == K.n_phi)
        cccs_march(B, cones, t, K, S);
    else if (wave == K.n_phi - 1)
        cccs_compute<EXACT_G, true>(K, S, wave);
    else
        cccs_compute<EXACT_G, false>(K, S, wave);
}

// cccs_kernel without the march wavefront: the workgroup's n_phi compute wavefronts read their
// chunks from the march table of the beamlet's geometry (cccs_table_kernel) and finish the
// records with the beamlet's lim_min and sad themselves.  Same arithmetic, same results.
template <bool EXACT_G>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6, 8))) void cccs_tab_kernel(
    const BeamletDev* __restrict__ beams, int n, const ConeTable* __restrict__ cones, int n_fwd,
    const float* __restrict__ terma, const float* __restrict__ dens_bev, float* __restrict__ partial,
    const TabEntry* __restrict__ tab, const float4* __restrict__ tab_rec,
    const int* __restrict__ tab_hdr) {
    extern __shared__ float4 lds4[];
    const int t = blockIdx.x / n;  // forward-cone slot, widest cone first
    const int b = blockIdx.x - t * n;
    const BeamletDev& B = beams[b];
    CcBlock K;
    K.L = B.L;
    if (K.L <= 0) return;
    K.tid = threadIdx.x;
    K.nthreads = blockDim.x;
    K.n_phi = blockDim.x >> 6;
    K.Tg = terma + B.bev_off;
    K.Rg = dens_bev + B.bev_off;
    K.P = cone_partial(partial, B, n_fwd, t);
    K.two = K.nthreads < 2 * kNpix;
    K.prefetch = K.nthreads >= kNpix;
    const CcLds S = carve_lds_tab(lds4, K.n_phi);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    K.A_a = cones->A_a[t];
    K.B_b = cones->B_b[t];
    K.na = cones->na[t];
    K.nb = cones->nb[t];
    const TabEntry& E = tab[(uint32_t)B.geom * (uint32_t)n_fwd + (uint32_t)t];
    TabRead TR;
    TR.hdr = tab_hdr + E.hdr_base;
    TR.rec = tab_rec + E.rec_base + 2 * (uint64_t)wave * E.rec_cap;
    TR.cap = E.rec_cap;
    TR.chunks = E.chunk_cap;
    TR.lim_min = B.lim_min;
    TR.inv_sad = 1.0f / B.sad;
    TR.bsx = B.bs[0];
    TR.bsy = B.bs[1];
    TR.ls = B.ls;
    for (int i = K.tid; i < 2 * kNpix; i += K.nthreads) store_slab(K, S, 0, i, i < kNpix ? K.Tg[i] : K.Rg[i - kNpix]);
    if (wave == K.n_phi - 1)
        cccs_compute<EXACT_G, true, true>(K, S, wave, TR);
    else
        cccs_compute<EXACT_G, false, true>(K, S, wave, TR);
}

// dose[b][s][p] = sum over forward cones of partial[b][t][s][p], cone slots in order.
__global__ __launch_bounds__(256) void cone_sum_kernel(const BeamletDev* __restrict__ beams, int n_fwd,
                                                       const float* __restrict__ partial,
                                                       float* __restrict__ dose) {
    const int b = blockIdx.y;
    const BeamletDev& B = beams[b];
    const uint64_t count = (uint64_t)B.L * kNpix;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float* src = partial + (uint64_t)B.bev_off * n_fwd + i;
    float acc = 0.0f;
    for (int t = 0; t < n_fwd; t++) acc += src[(uint64_t)t * count];
    dose[B.bev_off + i] = acc;
}

// The same, four values per thread, when the partial and dose buffers are 16-byte aligned and every
// beamlet's BEV offset is a multiple of four floats.
__global__ __launch_bounds__(256) void cone_sum4_kernel(const BeamletDev* __restrict__ beams, int n_fwd,
                                                       const float* __restrict__ partial,
                                                       float* __restrict__ dose) {
    const int b = blockIdx.y;
    const BeamletDev& B = beams[b];
    const uint64_t count4 = (uint64_t)B.L * (kNpix / 4);
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count4) return;
    const float4* src = reinterpret_cast<const float4*>(partial + (uint64_t)B.bev_off * n_fwd) + i;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int t = 0; t < n_fwd; t++) {
        const float4 v = src[(uint64_t)t * count4];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
    }
    reinterpret_cast<float4*>(dose + B.bev_off)[i] = acc;
}

}  // namespace

int cccs_table_max_phi() { return kMaxMarchPhi; }
int cccs_chunk_steps() { return kRecCap; }

hipError_t launch_cccs_tables(const CccsTables& tab, const ConeTable* cones_dev, int n_fwd, int n_phi,
                              hipStream_t stream) {
    if (n_phi < 1 || n_phi > kMaxMarchPhi) return hipErrorInvalidValue;
    hipError_t e = launch_fill_u32(tab.overflow, 1, 0u, stream);  // a kernel, also under graph capture
    if (e != hipSuccess || tab.n_gt <= 0) return e;
    hipLaunchKernelGGL(cccs_table_kernel, dim3((unsigned)tab.n_gt), dim3(kWave), table_lds_bytes(n_phi), stream,
                       tab.ent, cones_dev, n_fwd, n_phi, tab.rec, tab.hdr, tab.overflow);
    return hipGetLastError();
}

hipError_t launch_cccs(const BeamletDev* beams, int n, int max_L, const ConeTable* cones_dev, int n_fwd,
                       int n_phi, bool exact_g, const float* terma, const float* dens_bev, float* partial,
                       float* dose, bool vec4, hipStream_t stream, hipEvent_t* ev, const CccsTables* tab) {
    if (n <= 0) return hipSuccess;
    if (ev) (void)hipEventRecord(ev[0], stream);
    if (n_fwd > 0 && tab && n_phi <= kMaxMarchPhi) {
        // compute wavefronts only, reading their chunks from the tables
        const dim3 grid((unsigned)(n * n_fwd)), block(n_phi * kWave);
        const size_t shm = tab_lds_bytes(n_phi);
        if (exact_g)
            hipLaunchKernelGGL(cccs_tab_kernel<true>, grid, block, shm, stream, beams, n, cones_dev, n_fwd, terma,
                               dens_bev, partial, tab->ent, tab->rec, tab->hdr);
        else
            hipLaunchKernelGGL(cccs_tab_kernel<false>, grid, block, shm, stream, beams, n, cones_dev, n_fwd, terma,
                               dens_bev, partial, tab->ent, tab->rec, tab->hdr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    } else if (n_fwd > 0) {
        dim3 grid((unsigned)(n * n_fwd));
        // compute wavefronts + the march wavefront, within 1024 threads and 64 KiB of LDS
        if (n_phi <= kMaxMarchPhi) {
            const dim3 block((n_phi + 1) * kWave);
            const size_t shm = march_lds_bytes(n_phi);
            if (exact_g)
                hipLaunchKernelGGL(cccs_kernel<true>, grid, block, shm, stream, beams, n, cones_dev, n_fwd, terma,
                                   dens_bev, partial);
            else
                hipLaunchKernelGGL(cccs_kernel<false>, grid, block, shm, stream, beams, n, cones_dev, n_fwd, terma,
                                   dens_bev, partial);
        } else {
            // one azimuth per wavefront up to 16 azimuths (1024 threads), two beyond
            const int apw = n_phi <= kMaxPhiPerWave1 ? 1 : 2;
            const int waves = (n_phi + apw - 1) / apw;
            const dim3 block(waves * kWave);
            const size_t shm = (size_t)(6 + 2 * waves) * kNpix * sizeof(float);
            if (apw == 1) {
                if (exact_g)
                    hipLaunchKernelGGL((cccs_inline_kernel<true, 1>), grid, block, shm, stream, beams, n, cones_dev,
                                       n_fwd, terma, dens_bev, partial);
                else
                    hipLaunchKernelGGL((cccs_inline_kernel<false, 1>), grid, block, shm, stream, beams, n, cones_dev,
                                       n_fwd, terma, dens_bev, partial);
            } else {
                if (exact_g)
                    hipLaunchKernelGGL((cccs_inline_kernel<true, 2>), grid, block, shm, stream, beams, n, cones_dev,
                                       n_fwd, terma, dens_bev, partial);
                else
                    hipLaunchKernelGGL((cccs_inline_kernel<false, 2>), grid, block, shm, stream, beams, n, cones_dev,
                                       n_fwd, terma, dens_bev, partial);
            }
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], stream);
    if (vec4 && reinterpret_cast<uintptr_t>(dose) % 16 == 0 && reinterpret_cast<uintptr_t>(partial) % 16 == 0) {
        dim3 g2((unsigned)(((uint64_t)max_L * (kNpix / 4) + 255) / 256), (unsigned)n), b2(256);
        hipLaunchKernelGGL(cone_sum4_kernel, g2, b2, 0, stream, beams, n_fwd, partial, dose);
    } else {
        dim3 g2((unsigned)(((uint64_t)max_L * kNpix + 255) / 256), (unsigned)n), b2(256);
        hipLaunchKernelGGL(cone_sum_kernel, g2, b2, 0, stream, beams, n_fwd, partial, dose);
    }
    if (ev) (void)hipEventRecord(ev[2], stream);
    return hipGetLastError();
}

}  // namespace fdi
