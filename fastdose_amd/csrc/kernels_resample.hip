// kernels_resample.hip -- BEV -> PVCS resampling of beamlet doses straight into CSR rows.
//
// Semantics (IMRTOpt/src/IMRTgeom.cu:370-516): voxel v of the CT grid receives, for beamlet b,
// the mean over 5^3 supersamples q of the beamlet's BEV dose point-sampled at
//     u = x_bev kx / y_bev + 8,  w = z_bev kz / y_bev + 8,  d = (y_bev - lim_min) / ls,
// (x, y, z)_bev = Rf (q - source); samples with u, w outside [0, 16) or d outside [0, L) add 0
// and still count in the mean.  Only voxels that pass the reference's two culling tests are
// evaluated: the 8^3 super-voxel containing v has its centre within extent + half super-voxel
// diagonal of the beamlet axis, and v's centre lies within extent + half voxel diagonal.
// Exact zeros are dropped (cuSPARSE DenseToSparse semantics, IMRTSparseMatEns.cu:66-165), so
// the dense nBeamlets x nVoxels intermediate of the reference never exists.
//
// Work decomposition.  One workgroup per (z-plane, beamlet).  For each row y of the plane the
// workgroup bounds the x-range of voxels whose supersamples can reach the beamlet's frustum (six
// linear inequalities solved in FP64 with a margin); the concatenation of those ranges, in (y, x)
// order, is the candidate list.  Stage 1 counts candidates, stage 2 evaluates them and compacts
// nonzeros in voxel order, stage 3 copies them into the CSR arrays.  All decisions (culling
// tests, which BEV voxel a supersample hits) use FP32 with the operation order of the CPU oracle.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "xcd.h"

namespace fdi {

namespace {

constexpr int kRowThreads = 256;

struct PlaneGeom {
    double b0[3];  // BEV coordinates of the centre of voxel (0, y, z) minus source
    double g[3];   // BEV step per voxel along x
};

__device__ __forceinline__ void plane_row(const BeamletDev& B, int y, int z, float vsx, float vsy, float vsz,
                                          PlaneGeom& pg) {
    const double px = 0.5 * vsx - (double)B.src[0];
    const double py = (y + 0.5) * vsy - (double)B.src[1];
    const double pz = (z + 0.5) * vsz - (double)B.src[2];
    for (int r = 0; r < 3; r++) {
        pg.b0[r] = (double)B.Rf[3 * r] * px + (double)B.Rf[3 * r + 1] * py + (double)B.Rf[3 * r + 2] * pz;
        pg.g[r] = (double)B.Rf[3 * r] * vsx;
    }
}

// Candidate x-range [x0, x1) of row (y, z): the voxels whose supersamples can fall inside the
// beamlet's BEV frustum (depth in [lim_min, lim_min + L ls], |x|, |z| <= half the 16-pixel field
// at that depth).  rho bounds the distance of any supersample from its voxel centre plus a
// rounding margin.
__device__ __forceinline__ void row_interval(const BeamletDev& B, int y, int z, int X, float vsx, float vsy,
                                             float vsz, double rho, int& x0, int& x1) {
    PlaneGeom pg;
    plane_row(B, y, z, vsx, vsy, vsz, pg);
    const double cx = 0.5 * kFmap * (double)B.bs[0] / (double)B.sad;
    const double cz = 0.5 * kFmap * (double)B.bs[1] / (double)B.sad;
    const double rx = rho * sqrt(1.0 + cx * cx), rz = rho * sqrt(1.0 + cz * cz);
    const double ylo = (double)B.lim_min - rho;
    const double yhi = (double)B.lim_min + (double)B.L * (double)B.ls + rho;
    // alpha + beta * x <= gamma
    const double al[6] = {-pg.b0[1], pg.b0[1], pg.b0[0] - cx * pg.b0[1], -pg.b0[0] - cx * pg.b0[1],
                          pg.b0[2] - cz * pg.b0[1], -pg.b0[2] - cz * pg.b0[1]};
    const double be[6] = {-pg.g[1], pg.g[1], pg.g[0] - cx * pg.g[1], -pg.g[0] - cx * pg.g[1],
                          pg.g[2] - cz * pg.g[1], -pg.g[2] - cz * pg.g[1]};
    const double ga[6] = {-ylo, yhi, rx, rx, rz, rz};
    double lo = -1.0, hi = (double)X;
    bool empty = false;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const double rhs = ga[k] - al[k];
        if (be[k] > 0.0) {
            hi = fmin(hi, rhs / be[k]);
        } else if (be[k] < 0.0) {
            lo = fmax(lo, rhs / be[k]);
        } else if (rhs < 0.0) {
            empty = true;
        }
    }
    if (empty || !(lo <= hi)) {
        x0 = 0;
        x1 = 0;
        return;
    }
    // the voxels whose centre lies in [lo, hi] (the margin is in rho)
    int a = (int)ceil(lo), c = (int)floor(hi) + 1;
    a = a < 0 ? 0 : a;
    c = c > X ? X : c;
    x0 = a;
    x1 = c > a ? c : a;
}

// Largest distance of a supersample from its voxel centre, (1/2 - 1/(2 n)) |voxel diagonal| for
// n samples per axis, plus 0.1 mm for the rounding of the FP32 sample decisions.
__device__ __forceinline__ double rho_of(float vsx, float vsy, float vsz) {
    return (0.5 - 0.5 / kResampleSS) * sqrt((double)vsx * vsx + (double)vsy * vsy + (double)vsz * vsz) + 1e-2;
}

// block-wide exclusive scan of one u32 per thread (blockDim.x <= 1024, multiple of 64)
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < nw; w++) {
            uint32_t t = s_wave[w];
            s_wave[w] = acc;
            acc += t;
        }
        s_wave[nw] = acc;
    }
    __syncthreads();
    const uint32_t res = s_wave[wave] + x - v;
    total = s_wave[nw];
    __syncthreads();
    return res;
}

// Candidates per (beamlet, z-plane): a workgroup takes kCountPlanes planes of one beamlet, its
// threads spread over all (plane, row) pairs; integer sums, so the order does not matter.
constexpr int kCountPlanes = 8;

__global__ __launch_bounds__(kRowThreads) void resample_count_kernel(const BeamletDev* __restrict__ beams, int X,
                                                                     int Y, int Z, float vsx, float vsy, float vsz,
                                                                     uint32_t* __restrict__ cand_count,
                                                                     uint2* __restrict__ cand_rows) {
    __shared__ uint32_t s_cnt[kCountPlanes], s_ylo[kCountPlanes], s_yhi[kCountPlanes];
    const int z0 = blockIdx.x * kCountPlanes, b = blockIdx.y;
    const BeamletDev& B = beams[b];
    const double rho = rho_of(vsx, vsy, vsz);
    // Planes (rows) whose voxel centres lie outside the z-range (y-range) of the widened frustum
    // (the extremes of its eight vertices, mapped back with the transpose of the rotation, plus
    // 10 um for rounding) hold no candidate: only the rows [ya, yb) of the planes in [za, zb) are
    // walked.
    double zmin = 1e300, zmax = -1e300, ymin = 1e300, ymax = -1e300;
    {
        const double cx = 0.5 * kFmap * (double)B.bs[0] / (double)B.sad;
        const double cz = 0.5 * kFmap * (double)B.bs[1] / (double)B.sad;
        const double rx = rho * sqrt(1.0 + cx * cx), rz = rho * sqrt(1.0 + cz * cz);
        const double ylo = (double)B.lim_min - rho;
        const double yhi = (double)B.lim_min + (double)B.L * (double)B.ls + rho;
        for (int v = 0; v < 8; v++) {
            const double by = (v & 1) ? yhi : ylo;
            const double bx = ((v & 2) ? 1.0 : -1.0) * fabs(cx * by + rx);
            const double bz = ((v & 4) ? 1.0 : -1.0) * fabs(cz * by + rz);
            const double qz = (double)B.src[2] + (double)B.Rf[2] * bx + (double)B.Rf[5] * by + (double)B.Rf[8] * bz;
            const double qy = (double)B.src[1] + (double)B.Rf[1] * bx + (double)B.Rf[4] * by + (double)B.Rf[7] * bz;
            zmin = fmin(zmin, qz);
            zmax = fmax(zmax, qz);
            ymin = fmin(ymin, qy);
            ymax = fmax(ymax, qy);
        }
    }
    const int za = max(z0, (int)floor((zmin - 1e-3) / (double)vsz - 0.5));
    const int zb = min(min(z0 + kCountPlanes, Z), (int)ceil((zmax + 1e-3) / (double)vsz - 0.5) + 1);
    const int ya = max(0, (int)floor((ymin - 1e-3) / (double)vsy - 0.5));
    const int yb = min(Y, (int)ceil((ymax + 1e-3) / (double)vsy - 0.5) + 1);
    if (threadIdx.x < kCountPlanes) {
        s_cnt[threadIdx.x] = 0u;
        s_ylo[threadIdx.x] = 0xFFFFFFFFu;
        s_yhi[threadIdx.x] = 0u;
    }
    __syncthreads();
    const int np = max(zb - za, 0), ny = max(yb - ya, 0);
    for (int i = threadIdx.x; i < np * ny; i += blockDim.x) {
        const int p = i / ny, y = ya + (i - p * ny);
        int x0, x1;
        row_interval(B, y, za + p, X, vsx, vsy, vsz, rho, x0, x1);
        if (x1 > x0) {
            atomicAdd(&s_cnt[za - z0 + p], (uint32_t)(x1 - x0));
            atomicMin(&s_ylo[za - z0 + p], (uint32_t)y);
            atomicMax(&s_yhi[za - z0 + p], (uint32_t)y + 1u);
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < min(kCountPlanes, Z - z0)) {
        const size_t seg = (size_t)b * Z + z0 + threadIdx.x;
        const uint32_t cnt = s_cnt[threadIdx.x];
        cand_count[seg] = cnt;
        // rows [y_lo, y_hi) of the plane that hold candidates (the values pass walks only those)
        cand_rows[seg] = cnt ? make_uint2(s_ylo[threadIdx.x], s_yhi[threadIdx.x]) : make_uint2(0u, 0u);
    }
}

// floor(v) as an integer (v_cvt_flr_i32_f32: out-of-range values saturate)
__device__ __forceinline__ int floor_int(float v) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// value of voxel (x, y, z) for beamlet B; decisions in FP32 with the oracle's order.  RCP_EXACT:
// the device's reciprocal followed by one Newton step equals the IEEE quotient for every input
// (checked on the device when a context is created); otherwise the one significand where the step
// can tie is corrected.
// The 5^3 samples are summed as five partial sums, one per x-sample plane ii (25 samples each, in
// (jj, kk) order), added in ii order: p0 + p1 + p2 + p3 + p4.  k lanes (1, 2, 3 or 5, wave-uniform)
// share a voxel: in round r lane sub takes plane ii = r k + sub, and the group's first lane adds
// the round's partial sums in ii order, so the sum has the same rounding for every k.  Returns
// the value on the group's first lane, 0 on the others.
template <bool RCP_EXACT>
__device__ float voxel_value(const BeamletDev& B, const float* __restrict__ dose, int x, int y, int z, float vsx,
                             float vsy, float vsz, float extent, float half_diag, float half_sdiag, int k, int sub) {
    const float* M = B.Rf;
    const float s0 = B.src[0], s1 = B.src[1], s2 = B.src[2];
    // super-voxel test (d_superVoxelInterp, IMRTgeom.cu:471-516)
    {
        const float cx = (8 * vsx) * ((x >> 3) + 0.5f) - s0;
        const float cy = (8 * vsy) * ((y >> 3) + 0.5f) - s1;
        const float cz = (8 * vsz) * ((z >> 3) + 0.5f) - s2;
        const float bx = fmaf(M[2], cz, fmaf(M[1], cy, M[0] * cx));
        const float bz = fmaf(M[8], cz, fmaf(M[7], cy, M[6] * cx));
        const float dist = sqrtf(bx * bx + bz * bz);
        if (!(dist < extent + half_sdiag)) return 0.0f;
    }
    // voxel-centre test (IMRTgeom.cu:410-421)
    {
        const float cx = (x + 0.5f) * vsx - s0;
        const float cy = (y + 0.5f) * vsy - s1;
        const float cz = (z + 0.5f) * vsz - s2;
        const float bx = fmaf(M[2], cz, fmaf(M[1], cy, M[0] * cx));
        const float bz = fmaf(M[8], cz, fmaf(M[7], cy, M[6] * cx));
        const float dist = sqrtf(bx * bx + bz * bz);
        if (dist > half_diag + extent) return 0.0f;
    }
    const float inv_ss = 1.0f / kResampleSS;
    const float half = 0.5f * kFmap;
    const float* __restrict__ D = dose + B.bev_off;
    // per-axis sample coordinates minus source
    float qy[kResampleSS], qz[kResampleSS];
#pragma unroll
    for (int i = 0; i < kResampleSS; i++) {
        const float o = (i + 0.5f) * inv_ss;
        qy[i] = ((float)y + o) * vsy - s1;
        qz[i] = ((float)z + o) * vsz - s2;
    }
    // Projection u = x kx / y + 8 (w alike) with the oracle's rounding: u = fma(x kx, RN(1 / y), 8).
    // RN(1 / y) is the hardware reciprocal r (within one ulp) refined by one Newton step with fused
    // multiply-adds, r + r (1 - y r): correctly rounded whenever r is one of the two floats around
    // 1 / y, except for the all-ones significand, where the step can tie.  The rotation's partial
    // products are hoisted out of the sample loops (same expression, same rounding).
    // The range tests are the oracle's float compares, taken on the floors: for finite values
    // floor(v) in [0, n) iff 0 <= v < n, and out-of-range floors saturate to values the unsigned
    // compares reject; dd, u and w are finite whenever dd passes, since then y >= lim_min > 0.
    // The BEV dose is read through a buffer resource of the beamlet's L x 256 values: a sample that
    // fails reads beyond the range and gets 0 from the hardware, so the five samples of a column
    // issue their loads together and add exactly what the oracle adds.
    const uint32_t Lu = (uint32_t)B.L;
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void*)D, (short)0, (int)(Lu * kNpix * sizeof(float)), 0x00020000);
    constexpr uint32_t kOutside = 0x80000000u;  // byte offset beyond any beamlet's range
    const int rounds = (kResampleSS + k - 1) / k;
    const int lane = threadIdx.x & 63;
    float sum = 0.0f;
    for (int r = 0; r < rounds; r++) {
        const int ii = r * k + sub;
        float acc = 0.0f;
        if (ii < kResampleSS) {
            const float qx = ((float)x + ((float)ii + 0.5f) * inv_ss) * vsx - s0;
            const float px = M[0] * qx, py = M[3] * qx, pz = M[6] * qx;
            for (int jj = 0; jj < kResampleSS; jj++) {
                const float ax = fmaf(M[1], qy[jj], px), ay = fmaf(M[4], qy[jj], py), az = fmaf(M[7], qy[jj], pz);
                float val[kResampleSS];
#pragma unroll
                for (int kk = 0; kk < kResampleSS; kk++) {
                    const float by = fmaf(M[5], qz[kk], ay);
                    const float dd = (by - B.lim_min) * B.inv_ls;
                    const float bx = fmaf(M[2], qz[kk], ax);
                    const float bz = fmaf(M[8], qz[kk], az);
                    const float tx = bx * B.kx, tz = bz * B.kz;
                    const float r0 = __builtin_amdgcn_rcpf(by);
                    float rinv = fmaf(r0, fmaf(-by, r0, 1.0f), r0);
                    if (!RCP_EXACT) {
                        // all-ones significand: the quotient's significand is 0x000001, and the step
                        // gives it or (from below) the tie 0x000000
                        const uint32_t t = (__float_as_uint(by) & 0x7FFFFFu) + 1u;  // 0x800000 iff all ones
                        rinv = __uint_as_float(__float_as_uint(rinv) | (t >> 23));
                    }
                    const float u = fmaf(tx, rinv, half);
                    const float w = fmaf(tz, rinv, half);
                    const int fd = floor_int(dd), fu = floor_int(u), fw = floor_int(w);
                    const bool ok = ((uint32_t)fd < Lu) & ((uint32_t)(fu | fw) < (uint32_t)kFmap);
                    const uint32_t boff =
                        ok ? (((uint32_t)fd * kNpix + (uint32_t)fw * kFmap + (uint32_t)fu) << 2) : kOutside;
                    val[kk] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)boff, 0, 0));
                }
#pragma unroll
                for (int kk = 0; kk < kResampleSS; kk++) acc += val[kk];
            }
        }
        // the round's partial sums, in plane order, onto the group's first lane (the lanes of a
        // group share the voxel, so they all reach this point together)
        sum += acc;
        for (int d = 1; d < k; d++) {
            const float o = __shfl(acc, lane + d, 64);
            if (r * k + d < kResampleSS) sum += o;
        }
    }
    return sub == 0 ? sum * (1.0f / (kResampleSS * kResampleSS * kResampleSS)) : 0.0f;
}

// At most 64 VGPRs (8 wavefronts per SIMD): the sample gathers need the latency hiding; with the
// compiler's own choice (126 VGPRs, 4 wavefronts) the kernel runs 25 % slower.
// A workgroup takes one (z-plane, beamlet) segment: the rows [y_lo, y_hi) the count pass found to
// hold candidates, their x-ranges and prefix, then the candidates 256 at a time, one lane each; a
// wavefront left with at most 32 candidates (the segment's last) gives each voxel 2, 3 or 5 lanes,
// which split its x-sample planes (voxel_value).
template <bool RCP_EXACT>
__global__ __launch_bounds__(kRowThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void resample_values_kernel(
    const BeamletDev* __restrict__ beams, int X, int Y, int Z, float vsx, float vsy, float vsz, float extent,
    const float* __restrict__ dose, const uint64_t* __restrict__ cand_off, const uint2* __restrict__ cand_rows,
    uint64_t cap, uint32_t* __restrict__ overflow, uint64_t* __restrict__ tmp_cols, float* __restrict__ tmp_vals,
    uint32_t* __restrict__ nnz_count, uint32_t* __restrict__ rowmax_bits) {
    extern __shared__ uint32_t s_rows[];  // x0[Y], off[Y + 1] of rows y_lo + i
    __shared__ uint32_t s_wave[kRowThreads / 64 + 1];
    uint32_t* s_x0 = s_rows;
    uint32_t* s_off = s_rows + Y;
    // (z, b) from the XCD-grouped workgroup order: a beamlet's planes share one L2 for its BEV dose
    const uint32_t wg = xcd_logical_id(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int z = (int)(wg % (uint32_t)Z), b = (int)(wg / (uint32_t)Z);
    const size_t seg = (size_t)b * Z + z;
    if (cand_off[seg + 1] == cand_off[seg]) {  // no candidate in this plane (most planes of a beamlet)
        if (threadIdx.x == 0) nnz_count[seg] = 0u;
        return;
    }
    const BeamletDev& B = beams[b];
    const double rho = rho_of(vsx, vsy, vsz);
    const uint2 yr = cand_rows[seg];
    const int y_lo = (int)yr.x, n_rows = (int)(yr.y - yr.x);

    // row ranges and their prefix (rows handled in chunks of blockDim)
    uint32_t carry = 0;
    for (int r0 = 0; r0 < n_rows; r0 += blockDim.x) {
        const int r = r0 + threadIdx.x;
        uint32_t len = 0;
        if (r < n_rows) {
            int x0, x1;
            row_interval(B, y_lo + r, z, X, vsx, vsy, vsz, rho, x0, x1);
            s_x0[r] = (uint32_t)x0;
            len = (uint32_t)(x1 - x0);
        }
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(len, s_wave, tot);
        if (r < n_rows) s_off[r] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) s_off[n_rows] = carry;
    __syncthreads();
    const uint32_t total = carry;
    const uint64_t base = cand_off[seg];
    if (base + total > cap) {  // staging too small: report, write nothing (the caller re-runs)
        if (threadIdx.x == 0) {
            overflow[0] = 1u;
            nnz_count[seg] = 0u;
        }
        return;
    }

    const float half_diag = sqrtf(vsx * vsx + vsy * vsy + vsz * vsz) * 0.5f;
    const float svx = 8 * vsx, svy = 8 * vsy, svz = 8 * vsz;
    const float half_sdiag = sqrtf(svx * svx + svy * svy + svz * svz) * 0.5f;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;

    uint32_t written = 0;
    float vmax = 0.0f;
    for (uint32_t c0 = 0; c0 < total; c0 += blockDim.x) {
        // this wavefront's candidates of the pass, and the lanes each of them gets
        const uint32_t wbase = c0 + 64u * wave;
        const uint32_t m = total > wbase ? min(total - wbase, 64u) : 0u;
        const int k = m > 32u ? 1 : (m > 21u ? 2 : (m > 12u ? 3 : 5));
        const uint32_t grp = lane / (uint32_t)k, sub = lane - grp * (uint32_t)k;
        const bool mine = grp < m;
        const uint32_t c = wbase + grp;
        float v = 0.0f;
        int x = 0, y = 0;
        if (mine) {
            // row containing candidate c: largest r with s_off[r] <= c
            int lo = 0, hi = n_rows - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= c) lo = mid; else hi = mid - 1;
            }
            y = y_lo + lo;
            x = (int)(s_x0[lo] + (c - s_off[lo]));
            v = voxel_value<RCP_EXACT>(B, dose, x, y, z, vsx, vsy, vsz, extent, half_diag, half_sdiag, k, (int)sub);
        }
        const uint32_t nz = (mine && sub == 0u && v != 0.0f) ? 1u : 0u;
        vmax = fmaxf(vmax, v);
        uint32_t tot;
        const uint32_t pos = block_exclusive_scan(nz, s_wave, tot);
        if (nz) {
            const uint64_t o = base + written + pos;
            tmp_cols[o] = (uint64_t)x + (uint64_t)X * ((uint64_t)y + (uint64_t)Y * (uint64_t)z);
            tmp_vals[o] = v;
        }
        written += tot;
    }
    if (threadIdx.x == 0) nnz_count[seg] = written;
    if (rowmax_bits) {
        // beamlet maximum for the relative threshold (values are >= 0, so their bit patterns order
        // like the floats)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
        if ((threadIdx.x & 63) == 0 && vmax > 0.0f) atomicMax(&rowmax_bits[b], __float_as_uint(vmax));
    }
}

// Entries of segment (z, b) at or above thr * (beamlet maximum).
__global__ __launch_bounds__(256) void resample_threshold_count_kernel(int Z, const uint64_t* __restrict__ cand_off,
                                                                       const uint32_t* __restrict__ nnz_count,
                                                                       const float* __restrict__ tmp_vals,
                                                                       const uint32_t* __restrict__ rowmax_bits,
                                                                       float thr, uint32_t* __restrict__ keep_count) {
    __shared__ uint32_t s_wave[256 / 64 + 1];
    const size_t seg = (size_t)blockIdx.y * Z + blockIdx.x;
    const float cut = thr * __uint_as_float(rowmax_bits[blockIdx.y]);
    const uint32_t cnt = nnz_count[seg];
    const uint64_t src = cand_off[seg];
    uint32_t k = 0;
    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) k += tmp_vals[src + i] >= cut ? 1u : 0u;
    uint32_t total;
    block_exclusive_scan(k, s_wave, total);
    if (threadIdx.x == 0) keep_count[seg] = total;
}

// The arrays the compacted entries go to: the first destination that holds all `total` of them;
// false (write nothing) when neither does.  The choice is the same in every workgroup.
__device__ __forceinline__ bool pick_dst(uint64_t total, const CsrDst& d0, const CsrDst& d1, uint64_t*& columns,
                                         float*& values) {
    if (total <= d0.capacity) {
        columns = d0.columns;
        values = d0.values;
        return true;
    }
    if (total <= d1.capacity) {
        columns = d1.columns;
        values = d1.values;
        return true;
    }
    return false;
}

// Ordered copy of the entries that pass the threshold into the CSR arrays.
__global__ __launch_bounds__(256) void resample_compact_threshold_kernel(
    int Z, const uint64_t* __restrict__ cand_off, const uint64_t* __restrict__ nnz_off,
    const uint32_t* __restrict__ nnz_count, const uint64_t* __restrict__ tmp_cols, const float* __restrict__ tmp_vals,
    const uint32_t* __restrict__ rowmax_bits, float thr, const uint64_t* __restrict__ total, CsrDst dst0,
    CsrDst dst1) {
    __shared__ uint32_t s_wave[256 / 64 + 1];
    uint64_t* columns;
    float* values;
    if (!pick_dst(total[0], dst0, dst1, columns, values)) return;
    const size_t seg = (size_t)blockIdx.y * Z + blockIdx.x;
    const float cut = thr * __uint_as_float(rowmax_bits[blockIdx.y]);
    const uint32_t cnt = nnz_count[seg];
    const uint64_t src = cand_off[seg], dst = nnz_off[seg];
    uint32_t written = 0;
    for (uint32_t c0 = 0; c0 < cnt; c0 += blockDim.x) {
        const uint32_t i = c0 + threadIdx.x;
        const float v = i < cnt ? tmp_vals[src + i] : 0.0f;
        const uint32_t keep = (i < cnt && v >= cut) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pos = block_exclusive_scan(keep, s_wave, tot);
        if (keep) {
            columns[dst + written + pos] = tmp_cols[src + i];
            values[dst + written + pos] = v;
        }
        written += tot;
    }
}

// Ordered copy of every segment's entries into the CSR arrays.  With PLANES > 1 a workgroup takes
// that many planes of one beamlet, one wavefront per plane at a time (on large batches: a segment
// holds a few hundred entries, so a workgroup per segment would spend more on its launch than on
// its copy); with PLANES == 1 the workgroup copies one segment (small batches keep the parallelism).
constexpr int kCompactPlanes = 16;
constexpr uint64_t kCompactGroupSegments = 65536;  // segments of a batch from which planes are grouped
template <int PLANES>
__global__ __launch_bounds__(256) void resample_compact_kernel(int Z, const uint64_t* __restrict__ cand_off,
                                                               const uint64_t* __restrict__ nnz_off,
                                                               const uint32_t* __restrict__ nnz_count,
                                                               const uint64_t* __restrict__ tmp_cols,
                                                               const float* __restrict__ tmp_vals,
                                                               const uint64_t* __restrict__ total, CsrDst dst0,
                                                               CsrDst dst1) {
    uint64_t* columns;
    float* values;
    if (!pick_dst(total[0], dst0, dst1, columns, values)) return;
    constexpr bool kGrouped = PLANES > 1;
    const int z0 = (int)blockIdx.x * PLANES, z1 = min(z0 + PLANES, Z);
    const int first = kGrouped ? z0 + (int)(threadIdx.x >> 6) : z0;
    const int zstep = kGrouped ? (int)(blockDim.x >> 6) : 1;
    const uint32_t t0 = kGrouped ? (threadIdx.x & 63u) : threadIdx.x, tstep = kGrouped ? 64u : blockDim.x;
    for (int z = first; z < z1; z += zstep) {
        const size_t seg = (size_t)blockIdx.y * Z + z;
        const uint32_t cnt = nnz_count[seg];
        const uint64_t src = cand_off[seg], dst = nnz_off[seg];
        for (uint32_t i = t0; i < cnt; i += tstep) {
            columns[dst + i] = tmp_cols[src + i];
            values[dst + i] = tmp_vals[src + i];
        }
    }
}

// One Newton step from the hardware reciprocal against the IEEE quotient, for every significand
// and the exponents [exp0, exp0 + gridDim.y); counts the inputs where they differ.
__global__ __launch_bounds__(256) void rcp_check_kernel(int exp0, uint32_t* __restrict__ mismatches) {
    const uint32_t m = blockIdx.x * 256u + threadIdx.x;  // the grid covers exactly 2^23 significands
    const int ex = exp0 + (int)blockIdx.y;
    const float b = __uint_as_float(((uint32_t)(ex + 127) << 23) | m);
    const float r0 = __builtin_amdgcn_rcpf(b);
    const float r = fmaf(r0, fmaf(-b, r0, 1.0f), r0);
    if (r != 1.0f / b) atomicAdd(mismatches, 1u);
}

__global__ void row_offsets_kernel(int n, int Z, const uint64_t* __restrict__ nnz_off,
                                   const uint64_t* __restrict__ total, uint64_t* __restrict__ offsets) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) offsets[b] = nnz_off[(size_t)b * Z];
    if (b == n) offsets[n] = total[0];
}

}  // namespace

hipError_t launch_resample_count(const BeamletDev* beams, int n, const fd_volume& vol, uint32_t* cand_count,
                                 uint2* cand_rows, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    dim3 grid((vol.dim[2] + kCountPlanes - 1) / kCountPlanes, n), block(kRowThreads);
    hipLaunchKernelGGL(resample_count_kernel, grid, block, 0, stream, beams, (int)vol.dim[0], (int)vol.dim[1],
                       (int)vol.dim[2], vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2], cand_count,
                       cand_rows);
    return hipGetLastError();
}

hipError_t launch_resample_values(const BeamletDev* beams, int n, const fd_volume& vol, float extent,
                                  const float* dose_bev, const uint64_t* cand_off, const uint2* cand_rows,
                                  uint64_t cap, uint32_t* overflow, uint64_t* tmp_cols, float* tmp_vals,
                                  uint32_t* nnz_count, uint32_t* rowmax_bits, bool rcp_exact, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    dim3 grid(vol.dim[2], n), block(kRowThreads);
    const size_t shm = (2 * (size_t)vol.dim[1] + 1) * sizeof(uint32_t);
    if (rcp_exact)
        hipLaunchKernelGGL(resample_values_kernel<true>, grid, block, shm, stream, beams, (int)vol.dim[0],
                           (int)vol.dim[1], (int)vol.dim[2], vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2],
                           extent, dose_bev, cand_off, cand_rows, cap, overflow, tmp_cols, tmp_vals, nnz_count, rowmax_bits);
    else
        hipLaunchKernelGGL(resample_values_kernel<false>, grid, block, shm, stream, beams, (int)vol.dim[0],
                           (int)vol.dim[1], (int)vol.dim[2], vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2],
                           extent, dose_bev, cand_off, cand_rows, cap, overflow, tmp_cols, tmp_vals, nnz_count, rowmax_bits);
    return hipGetLastError();
}

hipError_t launch_rcp_check(int exp0, int n_exp, uint32_t* mismatches, hipStream_t stream) {
    hipLaunchKernelGGL(rcp_check_kernel, dim3((1u << 23) / 256, n_exp), dim3(256), 0, stream, exp0, mismatches);
    return hipGetLastError();
}

hipError_t launch_resample_threshold_count(int n, int Z, const uint64_t* cand_off, const uint32_t* nnz_count,
                                           const float* tmp_vals, const uint32_t* rowmax_bits, float thr,
                                           uint32_t* keep_count, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(resample_threshold_count_kernel, dim3(Z, n), dim3(256), 0, stream, Z, cand_off, nnz_count,
                       tmp_vals, rowmax_bits, thr, keep_count);
    return hipGetLastError();
}

hipError_t launch_resample_compact(int n, int Z, const uint64_t* cand_off, const uint64_t* nnz_off,
                                   const uint32_t* nnz_count, const uint64_t* tmp_cols, const float* tmp_vals,
                                   const uint32_t* rowmax_bits, float thr, uint64_t* offsets, CsrDst dst0,
                                   CsrDst dst1, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    dim3 grid(Z, n), block(256);
    const uint64_t* total = nnz_off + (size_t)n * Z;
    if (!dst0.columns || !dst0.values) dst0.capacity = 0;
    if (!dst1.columns || !dst1.values) dst1.capacity = 0;
    if (rowmax_bits && thr > 0.0f) {
        hipLaunchKernelGGL(resample_compact_threshold_kernel, grid, block, 0, stream, Z, cand_off, nnz_off, nnz_count,
                           tmp_cols, tmp_vals, rowmax_bits, thr, total, dst0, dst1);
    } else if ((uint64_t)n * (uint64_t)Z >= kCompactGroupSegments) {
        hipLaunchKernelGGL(resample_compact_kernel<kCompactPlanes>, dim3((Z + kCompactPlanes - 1) / kCompactPlanes, n),
                           block, 0, stream, Z, cand_off, nnz_off, nnz_count, tmp_cols, tmp_vals, total, dst0, dst1);
    } else {
        hipLaunchKernelGGL(resample_compact_kernel<1>, grid, block, 0, stream, Z, cand_off, nnz_off, nnz_count,
                           tmp_cols, tmp_vals, total, dst0, dst1);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // nnz_off has n*Z entries followed by the total at index n*Z
    hipLaunchKernelGGL(row_offsets_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, stream, n, Z, nnz_off,
                       nnz_off + (size_t)n * Z, offsets);
    return hipGetLastError();
}

}  // namespace fdi
