// fd_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of FastDose's CCCS beamlet-dose path (xuqifan897/FastDose), used as the
// correctness checker for the HIP library and as the timed CPU baseline.  Nothing in the
// product (fastdose_amd/, include/) links, loads or calls this file; only tests/, the smoke
// check and the CPU-baseline leg of bench.py do.
//
// Precision policy.
//   * Physics values (Terma, radiological path, CCCS recursion, resampled dose) are FP64.
//   * Every integer-valued decision of the path (beam range / long_dim, beamlet canonical
//     form, ray-lattice voxel indices and slice ends in the CCCS march, which BEV voxel a
//     supersample hits, which PVCS voxels are candidates, beamlet/PTV flags) is made in FP32
//     exactly as the reference's own arithmetic defines it, with a fixed operation order
//     (no contraction: built with -ffp-contract=off, fused steps spelled out with fmaf).
//     The HIP path makes the same decisions with the same FP32 operations, so index sets
//     compare bit-for-bit.
//
// Reference files followed (paths relative to the reference repository):
//   FastDoseLib/src/geom.cu:11-56            rotations (Rodrigues), beam <-> PVCS
//   FastDoseLib/src/beam.cu:134-221          BEAM_h::calc_range
//   FastDoseLib/src/spectrum.cu:14-60        spectrum reader
//   FastDoseLib/src/kernel.cu:13-72          CCCS kernel-table reader
//   FastDoseLib/src/TermaComp.cu:179-253     Terma ray trace
//   FastDoseLib/src/DoseComp.cu:62-175,217-392  CCCS recursion and its device helpers
//   IMRTOpt/src/IMRTBeamBundle.cu:15-127,269-316,341-569  bundle init, PTV flags, beamletInit
//   IMRTOpt/src/IMRTgeom.cu:370-516          BEV -> PVCS resampling with super-voxel culling
//   IMRTOpt/src/IMRTSparseMatEns.cu:66-239   dense -> CSR and the on-disk layout
//   KernelGen/derivation.md:94-291           CCCS maths (exact g_i option)
//
// Quirk decisions (reference behaviour kept unless stated):
//   Q1  g == 1 (integer-division Taylor series, DoseComp.cu:167-175); flag ORC_EXACT_G gives
//       g = (1 - e^-x)/x of derivation.md instead.
//   Q2  DoseBEV starts at zero for every beamlet (reference never clears it between beams).
//   Q3  dose-matrix resampling is point sampling (IMRTgeom.cu:294-300).
//   Q6  supersamples with u or w == 16, or beyond the beamlet's own long_dim, read 0.
//   Q7  lateral periodic wrap + X reset at BEV edge crossings; forward cones only.
//   Q8  race-free intent: every lane steps from the same pre-step base; resets precede adds.  A
//       ray is exactly the shared lattice point plus its integer lane offset (no per-lane FP32
//       rounding of the offset sum), so the lane -> voxel map of a segment is a wrapped shift.
//   Q9  per slice, dose summed over all phi is normalised by the LAST phi's cumulative p.
//   Q14 the wrapped lateral index is (floor(shared midpoint) + lane offset) mod 16 on the integer
//       (the reference's per-lane fmodf then +16 in FP32 can round a value just below an integer
//       up to it, or to 16 itself).
//   Q15 slice loop driven by phi 0's position (sliceIdx = round(base[0].z), DoseComp.cu:277);
//       a march that skips a slice boundary (ceil(z + 1e-3) rule) is kept as is.
//   Q16 segment length uses unwrapped lateral coordinates in the divergence factor.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cfloat>
#include <vector>
#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct {
    float voxel_size[3];
    uint32_t dim[3];
    uint32_t bbox_start[3];
    uint32_t bbox_dim[3];
} orc_volume;

typedef struct {
    float isocenter[3];
    float beamlet_size[2];
    uint32_t fmap_size[2];
    float sad;
    float angles[3];
    float long_spacing;
    float lim_min;
    float lim_max;
    uint32_t long_dim;
    float source[3];
} orc_beam;

typedef struct {
    int32_t n;
    float energy[24];
    float fluence[24];
    float mu[24];
    float mu_en[24];
} orc_spectrum;

typedef struct {
    int32_t n_theta;
    int32_t n_phi;
    float theta_begin[24];
    float theta_end[24];
    float theta_mid[24];
    float A[24];
    float a[24];
    float B[24];
    float b[24];
    float phi[24];
} orc_kernel;

typedef struct {
    float isocenter[3];
    float beamlet_size[2];
    int32_t fluence_dim[2];
    float sad;
    float angles[3];
    float long_spacing;
    int32_t sub_dim[2];
    int32_t sub_on[2];
} orc_bundle;

enum { ORC_EXACT_G = 1 };

}  // extern "C"

namespace {

const float kEps = 1e-4f;        // eps_fastdose (macros.h.in:26)
const float kLargerEps = 1e-3f;  // larger_eps_fastdose (macros.h.in:27)
const float kPiF = 3.141592654f; // CUDART_PI_F
const double kPiO2 = 1.57079632679489661923;  // CUDART_PIO2

struct f3 { float x, y, z; };
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline f3 dvs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len(f3 a) { return sqrtf(dot(a, a)); }
inline f3 nrm(f3 a) { float inv = 1.0f / sqrtf(dot(a, a)); return scl(a, inv); }
inline float comp(const f3& a, int j) { return j == 0 ? a.x : (j == 1 ? a.y : a.z); }

// geom.cu:11-34 -- Rodrigues rotation of p about unit axis r by angle t (FP32 host form)
f3 rot_axis(f3 p, f3 r, float t) {
    float s, c;
    sincosf(t, &s, &c);
    float omc = 1 - c;
    float pdr = p.x * r.x + p.y * r.y + p.z * r.z;
    float ftc = omc * pdr;
    return mk(ftc * r.x + c * p.x + s * (r.y * p.z - r.z * p.y),
              ftc * r.y + c * p.y + s * (r.z * p.x - r.x * p.z),
              ftc * r.z + c * p.z + s * (r.x * p.y - r.y * p.x));
}

// geom.cu:36-44 -- PVCS -> BEV
f3 rot_beam(f3 v, float th, float ph, float coll) {
    float s, c;
    sincosf(-ph, &s, &c);
    f3 axis = mk(s, 0.0f, c);
    f3 tmp = rot_axis(v, axis, -th);
    return rot_axis(tmp, mk(0.f, 1.f, 0.f), ph + coll);
}

// geom.cu:47-56 -- BEV -> PVCS
f3 inv_rot_beam(f3 v, float th, float ph, float coll) {
    f3 tmp = rot_axis(v, mk(0.f, 1.f, 0.f), -(ph + coll));
    float s, c;
    sincosf(-ph, &s, &c);
    f3 axis = mk(s, 0.0f, c);
    return rot_axis(tmp, axis, th);
}

// FP64 rotation matrices of the same composition, rounded once to FP32.  These are the
// per-beamlet transforms the HIP kernels use (fixed FP32 op order below).
void rodrigues64(const double r[3], double t, double M[9]) {
    double s = std::sin(t), c = std::cos(t), o = 1.0 - c;
    M[0] = c + o * r[0] * r[0];        M[1] = o * r[0] * r[1] - s * r[2]; M[2] = o * r[0] * r[2] + s * r[1];
    M[3] = o * r[1] * r[0] + s * r[2]; M[4] = c + o * r[1] * r[1];        M[5] = o * r[1] * r[2] - s * r[0];
    M[6] = o * r[2] * r[0] - s * r[1]; M[7] = o * r[2] * r[1] + s * r[0]; M[8] = c + o * r[2] * r[2];
}
void matmul3(const double A[9], const double B[9], double C[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C[3 * i + j] = A[3 * i + 0] * B[0 + j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
// forward (PVCS -> BEV) = Ry(ph + coll) * Raxis(couch(-ph), -th)
void beam_matrix64(double th, double ph, double coll, double M[9], bool inverse) {
    double ay[3] = {0.0, 1.0, 0.0};
    double ac[3] = {std::sin(-ph), 0.0, std::cos(-ph)};
    double R1[9], R2[9];
    if (!inverse) {
        rodrigues64(ac, -th, R1);
        rodrigues64(ay, ph + coll, R2);
        matmul3(R2, R1, M);
    } else {
        rodrigues64(ay, -(ph + coll), R1);
        rodrigues64(ac, th, R2);
        matmul3(R2, R1, M);
    }
}
void beam_matrix32(const orc_beam& b, float M[9], bool inverse) {
    double D[9];
    beam_matrix64((double)b.angles[0], (double)b.angles[1], (double)b.angles[2], D, inverse);
    for (int i = 0; i < 9; i++) M[i] = (float)D[i];
}
// b = M (x) d with the fixed FP32 order shared with the HIP kernels
inline f3 mat_apply32(const float M[9], f3 d) {
    return mk(fmaf(M[2], d.z, fmaf(M[1], d.y, M[0] * d.x)),
              fmaf(M[5], d.z, fmaf(M[4], d.y, M[3] * d.x)),
              fmaf(M[8], d.z, fmaf(M[7], d.y, M[6] * d.x)));
}

// beam.cu:134-221
int calc_range(orc_beam& bm, const orc_volume& vol) {
    float hx = bm.fmap_size[0] * bm.beamlet_size[0] / 2;
    float hy = bm.fmap_size[1] * bm.beamlet_size[1] / 2;
    f3 bnd[5] = {mk(hx, bm.sad, hy), mk(hx, bm.sad, -hy), mk(-hx, bm.sad, hx),  // Q5: [2].z uses x
                 mk(-hx, bm.sad, -hy), mk(0, bm.sad, 0)};
    for (int i = 0; i < 5; i++) bnd[i] = inv_rot_beam(bnd[i], bm.angles[0], bm.angles[1], bm.angles[2]);
    f3 src = inv_rot_beam(mk(0, -bm.sad, 0), bm.angles[0], bm.angles[1], bm.angles[2]);
    src = add(src, mk(bm.isocenter[0], bm.isocenter[1], bm.isocenter[2]));
    bm.source[0] = src.x; bm.source[1] = src.y; bm.source[2] = src.z;
    f3 vs = mk(vol.voxel_size[0], vol.voxel_size[1], vol.voxel_size[2]);
    f3 b0 = mk((float)vol.bbox_start[0] * vs.x, (float)vol.bbox_start[1] * vs.y, (float)vol.bbox_start[2] * vs.z);
    f3 b1 = mk((float)(vol.bbox_start[0] + vol.bbox_dim[0]) * vs.x,
               (float)(vol.bbox_start[1] + vol.bbox_dim[1]) * vs.y,
               (float)(vol.bbox_start[2] + vol.bbox_dim[2]) * vs.z);
    float mn = FLT_MAX, mx = FLT_MIN;
    bool hit = false;
    for (int i = 0; i < 5; i++) {
        const f3& bd = bnd[i];
        f3 inter[6];
        for (int j = 0; j < 3; j++) {
            double den = comp(bd, j);
            den = std::signbit(den) ? den - kEps : den + kEps;
            double al1 = (comp(b0, j) - comp(src, j)) / den;
            double al2 = (comp(b1, j) - comp(src, j)) / den;
            inter[2 * j] = add(src, scl(bd, (float)al1));
            inter[2 * j + 1] = add(src, scl(bd, (float)al2));
        }
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 2; k++) {
                const f3& p = inter[2 * j + k];
                int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                if (comp(p, j1) >= comp(b0, j1) && comp(p, j1) <= comp(b1, j1) &&
                    comp(p, j2) >= comp(b0, j2) && comp(p, j2) <= comp(b1, j2)) {
                    float l = len(sub(p, src)) * bm.sad / len(bd);
                    mn = std::min(l, mn);
                    mx = std::max(l, mx);
                    hit = true;
                }
            }
    }
    if (!hit) return 1;
    bm.lim_min = mn;
    bm.lim_max = mx;
    bm.long_dim = (uint32_t)std::ceil((bm.lim_max - bm.lim_min) / bm.long_spacing);
    return 0;
}

// IMRTBeamBundle.cu:341-569
int beamlet_init(const orc_bundle& bb, int ix, int iy, const orc_volume& vol, orc_beam& out, float* fluence) {
    float hfx = bb.fluence_dim[0] * bb.beamlet_size[0] * 0.5f;
    float hfy = bb.fluence_dim[1] * bb.beamlet_size[1] * 0.5f;
    f3 v0 = mk(ix * bb.beamlet_size[0] - hfx, bb.sad, iy * bb.beamlet_size[1] - hfy);
    f3 v1 = mk((ix + 1) * bb.beamlet_size[0] - hfx, bb.sad, iy * bb.beamlet_size[1] - hfy);
    f3 v2 = mk((ix + 1) * bb.beamlet_size[0] - hfx, bb.sad, (iy + 1) * bb.beamlet_size[1] - hfy);
    f3 v3 = mk(ix * bb.beamlet_size[0] - hfx, bb.sad, (iy + 1) * bb.beamlet_size[1] - hfy);
    v0 = nrm(v0); v1 = nrm(v1); v2 = nrm(v2); v3 = nrm(v3);

    f3 c1 = nrm(add(v0, v2));
    f3 v0c1 = dvs(v0, dot(v0, c1)), v1c1 = dvs(v1, dot(v1, c1));
    f3 v2c1 = dvs(v2, dot(v2, c1)), v3c1 = dvs(v3, dot(v3, c1));
    float d0c1 = len(sub(v0c1, c1)), d1c1 = len(sub(v1c1, c1));
    float d2c1 = len(sub(v2c1, c1)), d3c1 = len(sub(v3c1, c1));
    if (std::fabs(d0c1 - d2c1) > kEps) return 1;
    float dc1 = std::min(std::min(d0c1, d1c1), d3c1);

    f3 c2 = nrm(add(v1, v3));
    f3 v0c2 = dvs(v0, dot(v0, c2)), v1c2 = dvs(v1, dot(v1, c2));
    f3 v2c2 = dvs(v2, dot(v2, c2)), v3c2 = dvs(v3, dot(v3, c2));
    float d0c2 = len(sub(v0c2, c2)), d1c2 = len(sub(v1c2, c2));
    float d2c2 = len(sub(v2c2, c2)), d3c2 = len(sub(v3c2, c2));
    if (std::fabs(d1c2 - d3c2) > kEps) return 1;
    float dc2 = std::min(std::min(d0c2, d1c2), d2c2);

    f3 dir, codir;
    float fh, fw, sadr;
    if (dc1 > dc2) {
        dir = c1;
        f3 p0 = nrm(sub(v0c1, c1));
        f3 p1 = nrm(sub(v1c1, c1));
        codir = nrm(add(p0, p0));  // Q4: reference adds v0 to itself here
        sadr = bb.sad / c1.y;
        fh = len(scl(scl(add(p0, p1), dc1), 0.5f)) * sadr * 2;
        fw = len(scl(scl(sub(p0, p1), dc1), 0.5f)) * sadr * 2;
    } else {
        dir = c2;
        f3 p0 = nrm(sub(v0c2, c2));
        f3 p1 = nrm(sub(v1c2, c2));
        codir = nrm(add(p0, p1));
        sadr = bb.sad / c2.y;
        fh = len(scl(scl(add(p0, p1), dc2), 0.5f)) * sadr * 2;
        fw = len(scl(scl(sub(p0, p1), dc2), 0.5f)) * sadr * 2;
    }

    f3 dpv = inv_rot_beam(dir, bb.angles[0], bb.angles[1], bb.angles[2]);
    f3 sad_axis = mk(0.f, 1.f, 0.f);
    float theta = std::acos(dot(dpv, sad_axis));
    float smp = dpv.z * (1.0f / sqrtf(dpv.z * dpv.z + dpv.x * dpv.x + kEps * kEps));
    float phi = std::asin(smp);
    float thr = theta, phr = phi;
    f3 ird = inv_rot_beam(sad_axis, thr, phr, 0.f);
    float al = dot(ird, dpv);
    if (al < 1 - kEps) {
        thr = theta; phr = kPiF - phi;
        ird = inv_rot_beam(sad_axis, thr, phr, 0.f); al = dot(ird, dpv);
        if (al < 1 - kEps) {
            thr = -theta; phr = phi;
            ird = inv_rot_beam(sad_axis, thr, phr, 0.f); al = dot(ird, dpv);
            if (al < 1 - kEps) {
                thr = -theta; phr = kPiF - phi;
                ird = inv_rot_beam(sad_axis, thr, phr, 0.f); al = dot(ird, dpv);
                if (al < 1 - kEps) return 1;
            }
        }
    }
    f3 cdpv = inv_rot_beam(codir, bb.angles[0], bb.angles[1], bb.angles[2]);
    f3 cd_bev = mk(0.f, 0.f, -1.f);
    f3 cdir = inv_rot_beam(cd_bev, thr, phr, 0.f);
    float cc = dot(cdpv, cdir);
    cc = (cc < -1.0f) ? -1.0f : ((cc > 1.0f) ? 1.0f : cc);
    float coll = std::acos(cc);
    float collr = coll;
    cdir = inv_rot_beam(cd_bev, thr, phr, collr);
    float cal = dot(cdir, cdpv);
    if (cal < 1 - kEps * 10) {
        collr = -coll;
        cdir = inv_rot_beam(cd_bev, thr, phr, collr);
        float cal2 = dot(cdir, cdpv);
        if (cal2 < 1 - kEps * 10) return 1;
    }
    out.angles[0] = thr; out.angles[1] = phr; out.angles[2] = collr;
    f3 disp = mk((ix + 0.5f - 0.5f * bb.fluence_dim[0]) * bb.beamlet_size[0], 0,
                 (iy + 0.5f - 0.5f * bb.fluence_dim[1]) * bb.beamlet_size[1]);
    f3 dispp = inv_rot_beam(disp, bb.angles[0], bb.angles[1], bb.angles[2]);
    out.isocenter[0] = bb.isocenter[0] + dispp.x;
    out.isocenter[1] = bb.isocenter[1] + dispp.y;
    out.isocenter[2] = bb.isocenter[2] + dispp.z;
    out.beamlet_size[0] = fw / bb.sub_on[0];
    out.beamlet_size[1] = fh / bb.sub_on[1];
    out.fmap_size[0] = (uint32_t)bb.sub_dim[0];
    out.fmap_size[1] = (uint32_t)bb.sub_dim[1];
    out.sad = sadr;
    out.long_spacing = bb.long_spacing;
    if (fluence) {
        int n = bb.sub_dim[0] * bb.sub_dim[1];
        for (int i = 0; i < n; i++) fluence[i] = 0.0f;
        int lx = (int)((out.fmap_size[0] - bb.sub_on[0]) * 0.5f);
        int ly = (int)((out.fmap_size[1] - bb.sub_on[1]) * 0.5f);
        for (int j = ly; j < ly + bb.sub_on[1]; j++)
            for (int i = lx; i < lx + bb.sub_on[0]; i++) fluence[i + j * out.fmap_size[0]] = 1.0f;
    }
    return calc_range(out, vol);
}

// FP32 "is the linear texture of a binary mask > 0 here" with exact weights.
// Texel i covers [i, i+1); linear filtering weights the two texels around x - 0.5.
bool mask_linear_positive(const uint8_t* mask, const orc_volume& vol, f3 c) {
    float cb[3] = {c.x - 0.5f, c.y - 0.5f, c.z - 0.5f};
    int i0[3];
    bool frac[3];
    for (int d = 0; d < 3; d++) {
        float fl = floorf(cb[d]);
        if (!(fl > -2.0f && fl < (float)vol.dim[d] + 1.0f)) return false;
        i0[d] = (int)fl;
        frac[d] = (cb[d] - fl) > 0.0f;
    }
    for (int k = 0; k < 2; k++) {
        if (k == 1 && !frac[2]) break;
        int z = i0[2] + k;
        if (z < 0 || z >= (int)vol.dim[2]) continue;
        for (int j = 0; j < 2; j++) {
            if (j == 1 && !frac[1]) break;
            int y = i0[1] + j;
            if (y < 0 || y >= (int)vol.dim[1]) continue;
            for (int i = 0; i < 2; i++) {
                if (i == 1 && !frac[0]) break;
                int x = i0[0] + i;
                if (x < 0 || x >= (int)vol.dim[0]) continue;
                if (mask[(size_t)x + vol.dim[0] * ((size_t)y + (size_t)vol.dim[1] * z)]) return true;
            }
        }
    }
    return false;
}

// Trilinear sample (texture semantics: border 0, weights at x - 0.5), FP64.
double trilinear64(const float* vol_data, const orc_volume& vol, double cx, double cy, double cz) {
    double c[3] = {cx - 0.5, cy - 0.5, cz - 0.5};
    int i0[3];
    double w1[3];
    for (int d = 0; d < 3; d++) {
        double fl = std::floor(c[d]);
        if (!(fl > -2.0 && fl < (double)vol.dim[d] + 1.0)) return 0.0;
        i0[d] = (int)fl;
        w1[d] = c[d] - fl;
    }
    double acc = 0.0;
    for (int k = 0; k < 2; k++) {
        int z = i0[2] + k;
        if (z < 0 || z >= (int)vol.dim[2]) continue;
        double wz = k ? w1[2] : 1.0 - w1[2];
        for (int j = 0; j < 2; j++) {
            int y = i0[1] + j;
            if (y < 0 || y >= (int)vol.dim[1]) continue;
            double wy = j ? w1[1] : 1.0 - w1[1];
            for (int i = 0; i < 2; i++) {
                int x = i0[0] + i;
                if (x < 0 || x >= (int)vol.dim[0]) continue;
                double wx = i ? w1[0] : 1.0 - w1[0];
                acc += wx * wy * wz * (double)vol_data[(size_t)x + vol.dim[0] * ((size_t)y + (size_t)vol.dim[1] * z)];
            }
        }
    }
    return acc;
}

// ---- CCCS lattice (FP32, DoseComp.cu:76-164) ---------------------------------------------

// d_nextPoint (DoseComp.cu:76-111): larger eps for the ceil/floor, z snapped on a slice end.
// Step = distance * (1/dir) with a host-rounded reciprocal; position advanced with fmaf.
inline f3 next_point(f3 cur, f3 d, f3 inv, bool* flag) {
    float sx = d.x > 0 ? (ceilf(cur.x + kLargerEps) - cur.x) * inv.x
                       : (d.x < 0 ? (floorf(cur.x - kLargerEps) - cur.x) * inv.x : INFINITY);
    float sy = d.y > 0 ? (ceilf(cur.y + kLargerEps) - cur.y) * inv.y
                       : (d.y < 0 ? (floorf(cur.y - kLargerEps) - cur.y) * inv.y : INFINITY);
    float sz = d.z > 0 ? (ceilf(cur.z + kLargerEps) - cur.z) * inv.z
                       : (d.z < 0 ? (floorf(cur.z - kLargerEps) - cur.z) * inv.z : INFINITY);
    float st = fminf(fminf(sx, sy), sz);
    f3 r = mk(fmaf(st, d.x, cur.x), fmaf(st, d.y, cur.y), fmaf(st, d.z, cur.z));
    *flag = fabsf(r.z - roundf(r.z)) < kEps;
    if (*flag) {
        r = mk(fmaf(sz, d.x, cur.x), fmaf(sz, d.y, cur.y), fmaf(sz, d.z, cur.z));
        r.z = roundf(r.z);
    }
    return r;
}

// d_calcCoords (DoseComp.cu:147-164) for one lateral axis: periodic wrap of the FP32 midpoint,
// taken as floor(m) mod n on the integer (Q14: the reference's fmodf + n can round a value just
// below an integer up to it, or to n itself).
inline int wrap_index(float m, float n) {
    long long i = (long long)floorf(m);
    long long k = (long long)n;
    long long r = i % k;
    return (int)(r < 0 ? r + k : r);
}

inline long long mod_ll(long long a, long long n) {
    long long r = a % n;
    return r < 0 ? r + n : r;
}
inline long long floordiv_ll(long long a, long long n) { return (a - mod_ll(a, n)) / n; }

// d_calcLineSeg (DoseComp.cu:114-144) in FP64: physical distance of two lattice points
// (lateral offsets scaled by the divergence factor (z ls + lim_min) / sad, unwrapped -- Q16).
double line_seg64(const double* s, const double* e, const orc_beam* bm) {
    double hx = bm->fmap_size[0] * 0.5, hy = bm->fmap_size[1] * 0.5;
    double bsx = bm->beamlet_size[0], bsy = bm->beamlet_size[1], ls = bm->long_spacing;
    double lm = bm->lim_min, sad = bm->sad;
    double sx = (s[0] - hx) * bsx, sy = (s[1] - hy) * bsy, sz = s[2] * ls;
    double f = (sz + lm) / sad;
    sx *= f;
    sy *= f;
    double ex = (e[0] - hx) * bsx, ey = (e[1] - hy) * bsy, ez = e[2] * ls;
    f = (ez + lm) / sad;
    ex *= f;
    ey *= f;
    double dx = ex - sx, dy = ey - sy, dz = ez - sz;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// Per-beamlet, per-cone grid direction (DoseComp.cu:250-259): unit vector of (theta, phi)
// scaled to voxel units and renormalised, FP32.  The unit vector itself is computed in FP64
// and rounded once.
void cone_direction(float th, float ph, const orc_beam& bm, f3* d, f3* inv) {
    double st = std::sin((double)th), ct = std::cos((double)th);
    double sp = std::sin((double)ph), cp = std::cos((double)ph);
    f3 u = mk((float)(st * cp), (float)(st * sp), (float)ct);
    f3 v = mk(u.x / bm.beamlet_size[0], u.y / bm.beamlet_size[1], u.z / bm.long_spacing);
    f3 n = nrm(v);
    *d = n;
    *inv = mk(1.0f / n.x, 1.0f / n.y, 1.0f / n.z);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- geometry entry points
void orc_rotate_axis(const float* p, const float* r, float t, float* out) {
    f3 o = rot_axis(mk(p[0], p[1], p[2]), mk(r[0], r[1], r[2]), t);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
void orc_rotate_beam(const float* v, float th, float ph, float coll, float* out) {
    f3 o = rot_beam(mk(v[0], v[1], v[2]), th, ph, coll);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
void orc_inverse_rotate_beam(const float* v, float th, float ph, float coll, float* out) {
    f3 o = inv_rot_beam(mk(v[0], v[1], v[2]), th, ph, coll);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
void orc_beam_matrix(const orc_beam* b, int inverse, float* M) { beam_matrix32(*b, M, inverse != 0); }
void orc_beam_matrix64(double th, double ph, double coll, int inverse, double* M) {
    beam_matrix64(th, ph, coll, M, inverse != 0);
}
int orc_calc_range(orc_beam* b, const orc_volume* vol) { return calc_range(*b, *vol); }
int orc_beamlet_init(const orc_bundle* bb, int ix, int iy, const orc_volume* vol, orc_beam* out, float* fluence) {
    return beamlet_init(*bb, ix, iy, *vol, *out, fluence);
}
void orc_cone_direction(float th, float ph, const orc_beam* bm, float* d, float* inv) {
    f3 a, b;
    cone_direction(th, ph, *bm, &a, &b);
    d[0] = a.x; d[1] = a.y; d[2] = a.z; inv[0] = b.x; inv[1] = b.y; inv[2] = b.z;
}
int orc_next_point(const float* cur, const float* d, const float* inv, float* out) {
    bool flag;
    f3 r = next_point(mk(cur[0], cur[1], cur[2]), mk(d[0], d[1], d[2]), mk(inv[0], inv[1], inv[2]), &flag);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return flag ? 1 : 0;
}
int orc_wrap_index(float m, float n) { return wrap_index(m, n); }
// d_calcLineSeg (DoseComp.cu:114-144) in FP64 from FP32 endpoints
double orc_line_seg(const float* s, const float* e, const orc_beam* bm) {
    const double sd[3] = {s[0], s[1], s[2]}, ed[3] = {e[0], e[1], e[2]};
    return line_seg64(sd, ed, bm);
}

// IMRTBeamBundle.cu:36-65 -- PTV centroid (double) in voxel units then cm, returned as FP32
int orc_ptv_isocenter(const uint8_t* mask, const orc_volume* vol, float* iso) {
    double cx = 0, cy = 0, cz = 0;
    size_t n = 0;
    for (uint32_t k = 0; k < vol->dim[2]; k++)
        for (uint32_t j = 0; j < vol->dim[1]; j++)
            for (uint32_t i = 0; i < vol->dim[0]; i++) {
                size_t idx = i + (size_t)vol->dim[0] * (j + (size_t)vol->dim[1] * k);
                if (mask[idx] > 0) { cx += i + 0.5; cy += j + 0.5; cz += k + 0.5; n++; }
            }
    if (n == 0) return 1;
    cx /= n; cy /= n; cz /= n;
    cx *= vol->voxel_size[0]; cy *= vol->voxel_size[1]; cz *= vol->voxel_size[2];
    iso[0] = (float)cx; iso[1] = (float)cy; iso[2] = (float)cz;
    return 0;
}

// IMRTBeamBundle.cu:130-316 -- beamlet/PTV intersection flags of one bundle (d_beamletFlagInit
// semantics: 3x3 lateral samples per depth step of a beam spanning the bundle's fluence map,
// depth SXD = lim_min + i * long_spacing for i < long_dim, linear mask texture > 0).
int orc_beamlet_flags(const orc_bundle* bb, const uint8_t* mask, const orc_volume* vol, uint8_t* flags) {
    orc_beam bm;
    std::memset(&bm, 0, sizeof(bm));
    for (int i = 0; i < 3; i++) { bm.isocenter[i] = bb->isocenter[i]; bm.angles[i] = bb->angles[i]; }
    bm.beamlet_size[0] = bb->beamlet_size[0]; bm.beamlet_size[1] = bb->beamlet_size[1];
    bm.fmap_size[0] = (uint32_t)bb->fluence_dim[0]; bm.fmap_size[1] = (uint32_t)bb->fluence_dim[1];
    bm.sad = bb->sad;
    bm.long_spacing = bb->long_spacing;
    int miss = calc_range(bm, *vol);
    int fx = bb->fluence_dim[0], fy = bb->fluence_dim[1];
    if (miss) {
        for (int i = 0; i < fx * fy; i++) flags[i] = 0;
        return 1;
    }
    float Mi[9];
    beam_matrix32(bm, Mi, true);
    f3 src = mk(bm.source[0], bm.source[1], bm.source[2]);
    f3 vs = mk(vol->voxel_size[0], vol->voxel_size[1], vol->voxel_size[2]);
    const int ss = 3;
    for (int b = 0; b < fx * fy; b++) {
        int bx = b % fx, by = b / fx;
        uint8_t on = 0;
        for (uint32_t i = 0; i < bm.long_dim && !on; i++) {
            float sxd = bm.lim_min + i * bm.long_spacing;
            float s = sxd / bm.sad;
            for (int j = 0; j < ss && !on; j++)
                for (int k = 0; k < ss && !on; k++) {
                    f3 disp = mk((bx + (j + 0.5f) / ss - fx * 0.5f) * bm.beamlet_size[0], bm.sad,
                                 (by + (k + 0.5f) / ss - fy * 0.5f) * bm.beamlet_size[1]);
                    f3 pd = mat_apply32(Mi, disp);
                    f3 c = mk(fmaf(pd.x, s, src.x) / vs.x, fmaf(pd.y, s, src.y) / vs.y, fmaf(pd.z, s, src.z) / vs.z);
                    if (mask_linear_positive(mask, *vol, c)) on = 1;
                }
        }
        flags[b] = on;
    }
    return 0;
}

// ---------------------------------------------------------------- Terma (TermaComp.cu:179-253)
// FP64: sample k = 4i+j+1 sits at depth fraction t_k = lim_min/sad + k*ls/(4 sad) along the
// pixel ray; radiological path accumulates rho*|step|; T = mean over the 4 substeps of
// sum_e fluence_e E_e mu_e exp(-mu_e rpl) / t_k^2 (times the pixel fluence); rho_bar likewise.
void orc_terma(const orc_beam* bm, const float* fluence, const orc_volume* vol, const float* density,
               const orc_spectrum* spec, double* terma, double* dens) {
    double M[9];
    beam_matrix64(bm->angles[0], bm->angles[1], bm->angles[2], M, true);
    const int fx = bm->fmap_size[0], fy = bm->fmap_size[1], npix = fx * fy;
    const int L = bm->long_dim;
    const double sad = bm->sad, ls = bm->long_spacing, lm = bm->lim_min;
    const double vs[3] = {vol->voxel_size[0], vol->voxel_size[1], vol->voxel_size[2]};
    const double src[3] = {bm->source[0], bm->source[1], bm->source[2]};
    const int SS = 4;
    for (int p = 0; p < npix; p++) {
        int ix = p % fx, iy = p / fx;
        double pb[3] = {(ix + 0.5 - fx * 0.5) * bm->beamlet_size[0], sad, (iy + 0.5 - fy * 0.5) * bm->beamlet_size[1]};
        double pv[3];
        for (int r = 0; r < 3; r++) pv[r] = M[3 * r] * pb[0] + M[3 * r + 1] * pb[1] + M[3 * r + 2] * pb[2];
        double dt = ls / (sad * SS);
        double stepn = std::sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]) * dt;
        double rpl = 0.0;
        double fl = fluence[p];
        for (int i = 0; i < L; i++) {
            double tav = 0.0, dav = 0.0;
            for (int j = 0; j < SS; j++) {
                double t = lm / sad + (double)(SS * i + j + 1) * dt;
                double c[3];
                for (int r = 0; r < 3; r++) c[r] = (src[r] + pv[r] * t) / vs[r];
                double rho = trilinear64(density, *vol, c[0], c[1], c[2]);
                rpl += rho * stepn;
                dav += rho;
                double tl = 0.0;
                for (int e = 0; e < spec->n; e++) {
                    double mu = spec->mu[e];
                    tl += (double)spec->fluence[e] * fl * (double)spec->energy[e] * mu * std::exp(-mu * rpl);
                }
                tav += tl / (t * t);
            }
            terma[(size_t)i * npix + p] = tav / SS;
            dens[(size_t)i * npix + p] = dav / SS;
        }
    }
}

// ---------------------------------------------------------------- CCCS (DoseComp.cu:217-392)
// dose[L*npix] is zeroed here (Q2) and accumulated over the forward cones.
// Returns the number of (lane, segment) steps taken (N_seg).
int64_t orc_cccs(const orc_beam* bm, const double* terma, const double* dens, const orc_kernel* kern,
                 int flags, double* dose) {
    const int fx = bm->fmap_size[0], fy = bm->fmap_size[1], npix = fx * fy;
    const int L = bm->long_dim;
    std::fill(dose, dose + (size_t)L * npix, 0.0);
    const bool exact_g = (flags & ORC_EXACT_G) != 0;
    const double eps_dose = (double)kEps;
    const int nphi = kern->n_phi;
    std::vector<double> XA((size_t)nphi * npix), XB((size_t)nphi * npix), D(npix), C(npix);
    std::vector<f3> dir(nphi), inv(nphi), base(nphi);
    int64_t nseg = 0;
    auto gfun = [&](double x) -> double {
        if (!exact_g) return 1.0;
        if (x < 1e-4) return 1.0 - x / 2 + x * x / 6 - x * x * x / 24;
        return (1.0 - std::exp(-x)) / x;
    };
    for (int t = 0; t < kern->n_theta; t++) {
        float th = kern->theta_mid[t];
        if ((double)th > kPiO2) continue;
        double A = kern->A[t], a = kern->a[t], B = kern->B[t], b = kern->b[t];
        std::fill(XA.begin(), XA.end(), 0.0);
        std::fill(XB.begin(), XB.end(), 0.0);
        for (int f = 0; f < nphi; f++) {
            cone_direction(th, kern->phi[f], *bm, &dir[f], &inv[f]);
            base[f] = mk(0.5f, 0.5f, 0.0f);
        }
        while (true) {
            int s = (int)roundf(base[0].z);
            if (s >= L) break;
            const double* T = terma + (size_t)s * npix;
            const double* R = dens + (size_t)s * npix;
            std::fill(D.begin(), D.end(), 0.0);
            for (int f = 0; f < nphi; f++) {
                std::fill(C.begin(), C.end(), 0.0);
                double* xa = XA.data() + (size_t)f * npix;
                double* xb = XB.data() + (size_t)f * npix;
                while (true) {
                    bool flag;
                    f3 np = next_point(base[f], dir[f], inv[f], &flag);
                    // Q8: every ray sits exactly at the shared lattice point plus its integer lane
                    // offset, so its midpoint voxel is the shared midpoint voxel shifted by the
                    // offset and wrapped (Q14), and a BEV-edge crossing is a change of the 16-block
                    // of the integer floor along the segment.
                    const long long Mx = (long long)floorf((base[f].x + np.x) * 0.5f);
                    const long long My = (long long)floorf((base[f].y + np.y) * 0.5f);
                    const long long Fbx = (long long)floorf(base[f].x), Fnx = (long long)floorf(np.x);
                    const long long Fby = (long long)floorf(base[f].y), Fny = (long long)floorf(np.y);
                    for (int p = 0; p < npix; p++) {
                        int ix = p % fx, iy = p / fx;
                        int m = (int)mod_ll(Mx + ix, fx) + fx * (int)mod_ll(My + iy, fy);
                        double Tm = T[m], Rm = R[m];
                        const double bt[3] = {(double)base[f].x + ix, (double)base[f].y + iy, (double)base[f].z};
                        const double nt[3] = {(double)np.x + ix, (double)np.y + iy, (double)np.z};
                        double p_i = line_seg64(bt, nt, bm) * Rm;
                        double xav = xa[p], xbv = xb[p];
                        double ga = gfun(a * p_i), gb = gfun(b * p_i);
                        double segd = A / a * ((1 - ga) * Tm + ga * xav) + B / b * ((1 - gb) * Tm + gb * xbv);
                        D[m] += segd * p_i;
                        C[m] += p_i;
                        double ea = std::exp(-a * p_i), eb = std::exp(-b * p_i);
                        double nxa = Tm * (1 - ea) + ea * xav;
                        double nxb = Tm * (1 - eb) + eb * xbv;
                        if (floordiv_ll(Fbx + ix, fx) != floordiv_ll(Fnx + ix, fx) ||
                            floordiv_ll(Fby + iy, fy) != floordiv_ll(Fny + iy, fy)) {
                            nxa = 0.0;
                            nxb = 0.0;
                        }
                        xa[p] = nxa;
                        xb[p] = nxb;
                    }
                    nseg += npix;
                    base[f] = np;
                    if (flag) break;
                }
            }
            double* out = dose + (size_t)s * npix;
            for (int p = 0; p < npix; p++) out[p] += D[p] / (C[p] + eps_dose);
        }
    }
    return nseg;
}

// Count of (lane, segment) steps for one beamlet without the physics (same lattice).
int64_t orc_cccs_count(const orc_beam* bm, const orc_kernel* kern) {
    const int npix = bm->fmap_size[0] * bm->fmap_size[1];
    const int L = bm->long_dim;
    const int nphi = kern->n_phi;
    std::vector<f3> dir(nphi), inv(nphi), base(nphi);
    int64_t nseg = 0;
    for (int t = 0; t < kern->n_theta; t++) {
        float th = kern->theta_mid[t];
        if ((double)th > kPiO2) continue;
        for (int f = 0; f < nphi; f++) {
            cone_direction(th, kern->phi[f], *bm, &dir[f], &inv[f]);
            base[f] = mk(0.5f, 0.5f, 0.0f);
        }
        while (true) {
            int s = (int)roundf(base[0].z);
            if (s >= L) break;
            for (int f = 0; f < nphi; f++) {
                while (true) {
                    bool flag;
                    f3 np = next_point(base[f], dir[f], inv[f], &flag);
                    nseg += npix;
                    base[f] = np;
                    if (flag) break;
                }
            }
        }
    }
    return nseg;
}

// ---------------------------------------------------------------- resampler (IMRTgeom.cu:370-516)
// Point-sampled 5^3 supersampling of one beamlet's BEV dose onto the PVCS grid, restricted to
// the reference's two culling tests (8^3 super-voxel axis distance, voxel-centre axis
// distance).  Decisions in FP32 (order shared with the HIP kernel), sums in FP64.
// Writes up to `cap` nonzero (linear voxel index, value) pairs in ascending index order and
// returns the nonzero count (which may exceed cap).
int64_t orc_resample(const orc_beam* bm, const double* dose_bev, const orc_volume* vol, float extent,
                     int64_t cap, int64_t* cols, double* vals, int64_t* n_samples_in) {
    float M[9];
    beam_matrix32(*bm, M, false);
    const f3 src = mk(bm->source[0], bm->source[1], bm->source[2]);
    const f3 vs = mk(vol->voxel_size[0], vol->voxel_size[1], vol->voxel_size[2]);
    const int fx = bm->fmap_size[0], fy = bm->fmap_size[1], npix = fx * fy;
    const float nfx = (float)fx, nfy = (float)fy;
    const float hfx = 0.5f * nfx, hfy = 0.5f * nfy;
    const float kx = bm->sad / bm->beamlet_size[0];
    const float kz = bm->sad / bm->beamlet_size[1];
    const float inv_ls = 1.0f / bm->long_spacing;
    const float Lf = (float)bm->long_dim;
    const float half_diag = len(vs) * 0.5f;
    const f3 svs = mk(8 * vs.x, 8 * vs.y, 8 * vs.z);
    const float half_sdiag = len(svs) * 0.5f;
    const float inv_ss = 1.0f / 5;
    const uint32_t X = vol->dim[0], Y = vol->dim[1], Z = vol->dim[2];
    const uint32_t SX = (X + 7) / 8, SY = (Y + 7) / 8, SZ = (Z + 7) / 8;
    std::vector<std::pair<int64_t, double>> out;
    int64_t nin = 0;
    for (uint32_t sz = 0; sz < SZ; sz++)
        for (uint32_t sy = 0; sy < SY; sy++)
            for (uint32_t sx = 0; sx < SX; sx++) {
                f3 c = mk(svs.x * (sx + 0.5f), svs.y * (sy + 0.5f), svs.z * (sz + 0.5f));
                f3 b = mat_apply32(M, sub(c, src));
                float dist = sqrtf(b.x * b.x + b.z * b.z);
                if (!(dist < extent + half_sdiag)) continue;
                for (uint32_t z = sz * 8; z < std::min(Z, sz * 8 + 8); z++)
                    for (uint32_t y = sy * 8; y < std::min(Y, sy * 8 + 8); y++)
                        for (uint32_t x = sx * 8; x < std::min(X, sx * 8 + 8); x++) {
                            f3 vc = mk((x + 0.5f) * vs.x, (y + 0.5f) * vs.y, (z + 0.5f) * vs.z);
                            f3 bc = mat_apply32(M, sub(vc, src));
                            float dv = sqrtf(bc.x * bc.x + bc.z * bc.z);
                            if (dv > half_diag + extent) continue;
                            double acc = 0.0;
                            for (int kk = 0; kk < 5; kk++) {
                                float ok = (kk + 0.5f) * inv_ss;
                                for (int jj = 0; jj < 5; jj++) {
                                    float oj = (jj + 0.5f) * inv_ss;
                                    for (int ii = 0; ii < 5; ii++) {
                                        float oi = (ii + 0.5f) * inv_ss;
                                        f3 q = mk(((float)x + oi) * vs.x, ((float)y + oj) * vs.y, ((float)z + ok) * vs.z);
                                        f3 bq = mat_apply32(M, sub(q, src));
                                        float dd = (bq.y - bm->lim_min) * inv_ls;
                                        if (!(dd >= 0.0f && dd < Lf)) continue;
                                        float rinv = 1.0f / bq.y;
                                        float u = fmaf(bq.x * kx, rinv, hfx);
                                        float w = fmaf(bq.z * kz, rinv, hfy);
                                        if (!(u >= 0.0f && u < nfx && w >= 0.0f && w < nfy)) continue;
                                        int s = (int)dd, pu = (int)u, pw = (int)w;
                                        acc += dose_bev[(size_t)s * npix + pw * fx + pu];
                                        nin++;
                                    }
                                }
                            }
                            double val = acc / 125.0;
                            if (val != 0.0) out.emplace_back((int64_t)x + (int64_t)X * ((int64_t)y + (int64_t)Y * z), val);
                        }
            }
    std::sort(out.begin(), out.end(), [](const std::pair<int64_t, double>& l, const std::pair<int64_t, double>& r) {
        return l.first < r.first;
    });
    int64_t n = (int64_t)out.size();
    for (int64_t i = 0; i < n && i < cap; i++) { cols[i] = out[i].first; vals[i] = out[i].second; }
    if (n_samples_in) *n_samples_in = nin;
    return n;
}

// ---------------------------------------------------------------- dense -> CSR
// cusparseDenseToSparse semantics used by MatCSREnsemble::addMat (IMRTSparseMatEns.cu:66-165):
// row-major input, exact zeros dropped, ascending columns, 0-based, 64-bit offsets.
int64_t orc_dense_to_csr(const float* dense, int64_t rows, int64_t cols, uint64_t* offsets,
                         uint64_t* columns, float* values) {
    int64_t nnz = 0;
    offsets[0] = 0;
    for (int64_t r = 0; r < rows; r++) {
        for (int64_t c = 0; c < cols; c++) {
            float v = dense[r * cols + c];
            if (v != 0.0f) {
                if (columns) columns[nnz] = (uint64_t)c;
                if (values) values[nnz] = v;
                nnz++;
            }
        }
        offsets[r + 1] = (uint64_t)nnz;
    }
    return nnz;
}

// ---------------------------------------------------------------- text readers
// spectrum.cu:14-60: rows "energy fluence mu mu_en name"; fluence normalised to sum 1.
// Blank lines are skipped (Q12).
int orc_read_spectrum(const char* path, orc_spectrum* out) {
    std::ifstream f(path);
    if (!f) return 1;
    std::string row;
    int n = 0;
    float sum = 0.f;
    while (std::getline(f, row)) {
        if (row.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        if (n >= 24) return 2;
        std::istringstream iss(row);
        std::string name;
        iss >> out->energy[n] >> out->fluence[n] >> out->mu[n] >> out->mu_en[n] >> name;
        sum += out->fluence[n];
        n++;
    }
    out->n = n;
    for (int i = 0; i < n; i++) out->fluence[i] /= sum;
    return 0;
}

// kernel.cu:13-72: header line, rows "thetaBegin thetaEnd A a B b"; theta_mid = (b+e)/2;
// phi_i = (i + 0.5) * (2 pi / nPhi).
int orc_read_kernel(const char* path, int nphi, orc_kernel* out) {
    std::ifstream f(path);
    if (!f) return 1;
    std::string row;
    std::getline(f, row);
    int n = 0;
    while (std::getline(f, row)) {
        if (row.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        if (n >= 24) return 2;
        std::istringstream iss(row);
        iss >> out->theta_begin[n] >> out->theta_end[n] >> out->A[n] >> out->a[n] >> out->B[n] >> out->b[n];
        out->theta_mid[n] = (out->theta_begin[n] + out->theta_end[n]) / 2;
        n++;
    }
    if (nphi > 24) return 2;
    out->n_theta = n;
    out->n_phi = nphi;
    float itv = 2 * kPiF / nphi;
    for (int i = 0; i < nphi; i++) out->phi[i] = (float)((i + 0.5) * itv);
    return 0;
}

// ---------------------------------------------------------------- whole-path CPU baseline
// For n beamlets sharing one fluence map: Terma -> CCCS -> resample, parallel over beamlets.
// nnz[i] receives each beamlet's nonzero count, checksum[i] the sum of its values.
int orc_dose_beamlets(int n, const orc_beam* beams, const float* fluence, const orc_volume* vol,
                      const float* density, const orc_spectrum* spec, const orc_kernel* kern, float extent,
                      int flags, int threads, int64_t* nnz, double* checksum) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; i++) {
        const orc_beam& bm = beams[i];
        size_t sz = (size_t)bm.long_dim * bm.fmap_size[0] * bm.fmap_size[1];
        std::vector<double> T(sz), R(sz), D(sz);
        orc_terma(&bm, fluence, vol, density, spec, T.data(), R.data());
        orc_cccs(&bm, T.data(), R.data(), kern, flags, D.data());
        int64_t cnt = orc_resample(&bm, D.data(), vol, extent, 0, nullptr, nullptr, nullptr);
        std::vector<int64_t> c(cnt);
        std::vector<double> v(cnt);
        orc_resample(&bm, D.data(), vol, extent, cnt, c.data(), v.data(), nullptr);
        double s = 0;
        for (double x : v) s += x;
        nnz[i] = cnt;
        checksum[i] = s;
    }
    return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
