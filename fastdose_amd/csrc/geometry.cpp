// geometry.cpp -- host-side beam and beamlet geometry of libfastdose_hip.
//
// These routines decide integers (long_dim, beamlet flags) and the canonical beamlet form, so
// they follow the reference's FP32 host arithmetic operation for operation (built without
// contraction).  Sources: FastDoseLib/src/geom.cu:11-56 (rotations), FastDoseLib/src/beam.cu:
// 134-221 (calc_range), IMRTOpt/src/IMRTBeamBundle.cu:15-127, 130-316, 319-569 (bundles).
#include <cmath>
#include <cfloat>
#include <cstring>
#include <algorithm>
#include <vector>

#include "fd_internal.h"
#include "geometry.h"

namespace fdi {

namespace {

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length3(V3 a) { return sqrtf(dot3(a, a)); }
inline V3 normalize3(V3 a) {
    float r = 1.0f / sqrtf(dot3(a, a));
    return a * r;
}
inline float at(const V3& a, int j) { return j == 0 ? a.x : (j == 1 ? a.y : a.z); }

constexpr float kPi = 3.141592654f;  // CUDART_PI_F

// Rodrigues rotation of p about the unit axis r by t: (1-cos t)(p.r) r + p cos t + (r x p) sin t.
V3 rotate_about(V3 p, V3 r, float t) {
    float s, c;
    sincosf(t, &s, &c);
    float k = (1 - c) * (p.x * r.x + p.y * r.y + p.z * r.z);
    return v3(k * r.x + c * p.x + s * (r.y * p.z - r.z * p.y),
              k * r.y + c * p.y + s * (r.z * p.x - r.x * p.z),
              k * r.z + c * p.z + s * (r.x * p.y - r.y * p.x));
}

// BEV -> PVCS: undo the collimator(+couch correction) turn about y, then the gantry turn about
// the couch-rotated axis.
V3 bev_to_pvcs(V3 v, float gantry, float couch, float coll) {
    V3 t = rotate_about(v, v3(0.f, 1.f, 0.f), -(couch + coll));
    float s, c;
    sincosf(-couch, &s, &c);
    return rotate_about(t, v3(s, 0.0f, c), gantry);
}

}  // namespace

void beam_matrix(double gantry, double couch, double coll, bool to_pvcs, float out[9]) {
    auto rod = [](const double r[3], double t, double M[9]) {
        double s = std::sin(t), c = std::cos(t), o = 1.0 - c;
        M[0] = c + o * r[0] * r[0];        M[1] = o * r[0] * r[1] - s * r[2]; M[2] = o * r[0] * r[2] + s * r[1];
        M[3] = o * r[1] * r[0] + s * r[2]; M[4] = c + o * r[1] * r[1];        M[5] = o * r[1] * r[2] - s * r[0];
        M[6] = o * r[2] * r[0] - s * r[1]; M[7] = o * r[2] * r[1] + s * r[0]; M[8] = c + o * r[2] * r[2];
    };
    const double y_axis[3] = {0.0, 1.0, 0.0};
    const double couch_axis[3] = {std::sin(-couch), 0.0, std::cos(-couch)};
    double first[9], second[9], M[9];
    if (!to_pvcs) {   // PVCS -> BEV: gantry turn by -gantry, then y turn by couch + coll
        rod(couch_axis, -gantry, first);
        rod(y_axis, couch + coll, second);
    } else {          // BEV -> PVCS
        rod(y_axis, -(couch + coll), first);
        rod(couch_axis, gantry, second);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            M[3 * i + j] = second[3 * i] * first[j] + second[3 * i + 1] * first[3 + j] + second[3 * i + 2] * first[6 + j];
    for (int i = 0; i < 9; i++) out[i] = (float)M[i];
}

int calc_range(fd_beam& beam, const fd_volume& vol) {
    const float hx = beam.fmap_size[0] * beam.beamlet_size[0] / 2;
    const float hy = beam.fmap_size[1] * beam.beamlet_size[1] / 2;
    // four field corners and the central axis, BEV (the third corner's z uses the x half
    // width exactly as beam.cu:143 does)
    V3 edge[5] = {v3(hx, beam.sad, hy), v3(hx, beam.sad, -hy), v3(-hx, beam.sad, hx), v3(-hx, beam.sad, -hy),
                  v3(0, beam.sad, 0)};
    for (auto& e : edge) e = bev_to_pvcs(e, beam.angles[0], beam.angles[1], beam.angles[2]);
    V3 src = bev_to_pvcs(v3(0, -beam.sad, 0), beam.angles[0], beam.angles[1], beam.angles[2]) +
             v3(beam.isocenter[0], beam.isocenter[1], beam.isocenter[2]);
    beam.source[0] = src.x;
    beam.source[1] = src.y;
    beam.source[2] = src.z;

    const V3 lo = v3((float)vol.bbox_start[0] * vol.voxel_size[0], (float)vol.bbox_start[1] * vol.voxel_size[1],
                     (float)vol.bbox_start[2] * vol.voxel_size[2]);
    const V3 hi = v3((float)(vol.bbox_start[0] + vol.bbox_dim[0]) * vol.voxel_size[0],
                     (float)(vol.bbox_start[1] + vol.bbox_dim[1]) * vol.voxel_size[1],
                     (float)(vol.bbox_start[2] + vol.bbox_dim[2]) * vol.voxel_size[2]);
    float near = FLT_MAX, far = FLT_MIN;
    bool any = false;
    for (const V3& e : edge) {
        for (int j = 0; j < 3; j++) {
            // ray/plane parameters in double with the reference's sign-aware epsilon
            double den = at(e, j);
            den += std::signbit(den) ? -(double)kEps : (double)kEps;
            for (int side = 0; side < 2; side++) {
                double alpha = ((side ? at(hi, j) : at(lo, j)) - at(src, j)) / den;
                V3 p = src + e * (float)alpha;
                int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                bool inside = at(p, j1) >= at(lo, j1) && at(p, j1) <= at(hi, j1) && at(p, j2) >= at(lo, j2) &&
                              at(p, j2) <= at(hi, j2);
                if (!inside) continue;
                float d = length3(p - src) * beam.sad / length3(e);
                near = std::min(d, near);
                far = std::max(d, far);
                any = true;
            }
        }
    }
    if (!any) {
        set_error("the beam doesn't intersect with the bounding box");
        return FD_ERR_GEOM;
    }
    if (!(near >= 0x1p-60f)) {  // the BEV depth axis starts at the source; the resampler needs depth > 0
        set_error("the beam source lies on the bounding box");
        return FD_ERR_GEOM;
    }
    beam.lim_min = near;
    beam.lim_max = far;
    beam.long_dim = (uint32_t)std::ceil((beam.lim_max - beam.lim_min) / beam.long_spacing);
    return FD_OK;
}

int beamlet_init(const fd_bundle& bb, int ix, int iy, const fd_volume& vol, fd_beam& out, float* fluence) {
    // Largest inscribed canonical pyramid of the beamlet (IMRTBeamBundle.cu:341-569).
    const float half_w = bb.fluence_dim[0] * bb.beamlet_size[0] * 0.5f;
    const float half_h = bb.fluence_dim[1] * bb.beamlet_size[1] * 0.5f;
    const float x0 = ix * bb.beamlet_size[0] - half_w, x1 = (ix + 1) * bb.beamlet_size[0] - half_w;
    const float z0 = iy * bb.beamlet_size[1] - half_h, z1 = (iy + 1) * bb.beamlet_size[1] - half_h;
    V3 c[4] = {normalize3(v3(x0, bb.sad, z0)), normalize3(v3(x1, bb.sad, z0)), normalize3(v3(x1, bb.sad, z1)),
               normalize3(v3(x0, bb.sad, z1))};

    struct Candidate {
        V3 axis;
        V3 proj[4];
        float dist[4];
        float radius;
    };
    auto make = [&](int a, int b) {
        Candidate k;
        k.axis = normalize3(c[a] + c[b]);
        for (int i = 0; i < 4; i++) {
            k.proj[i] = c[i] / dot3(c[i], k.axis);
            k.dist[i] = length3(k.proj[i] - k.axis);
        }
        return k;
    };
    Candidate d1 = make(0, 2);   // axis bisecting corners 0 and 2
    if (std::fabs(d1.dist[0] - d1.dist[2]) > kEps) {
        set_error("beamlet canonical form: unequal diagonal (case 1)");
        return FD_ERR_GEOM;
    }
    d1.radius = std::min(std::min(d1.dist[0], d1.dist[1]), d1.dist[3]);
    Candidate d2 = make(1, 3);   // axis bisecting corners 1 and 3
    if (std::fabs(d2.dist[1] - d2.dist[3]) > kEps) {
        set_error("beamlet canonical form: unequal diagonal (case 2)");
        return FD_ERR_GEOM;
    }
    d2.radius = std::min(std::min(d2.dist[0], d2.dist[1]), d2.dist[2]);

    const bool first = d1.radius > d2.radius;
    const Candidate& k = first ? d1 : d2;
    V3 e0 = normalize3(k.proj[0] - k.axis);
    V3 e1 = normalize3(k.proj[1] - k.axis);
    // the first case builds its reference direction from corner 0 twice (IMRTBeamBundle.cu:424)
    V3 co_dir = normalize3(first ? e0 + e0 : e0 + e1);
    const float sad_r = bb.sad / k.axis.y;
    const float height = length3((e0 + e1) * k.radius * 0.5f) * sad_r * 2;
    const float width = length3((e0 - e1) * k.radius * 0.5f) * sad_r * 2;

    // gantry/couch of the canonical axis: four sign cases
    const V3 axis_pvcs = bev_to_pvcs(k.axis, bb.angles[0], bb.angles[1], bb.angles[2]);
    const V3 y_axis = v3(0.f, 1.f, 0.f);
    const float theta = std::acos(dot3(axis_pvcs, y_axis));
    const float sin_mphi =
        axis_pvcs.z * (1.0f / sqrtf(axis_pvcs.z * axis_pvcs.z + axis_pvcs.x * axis_pvcs.x + kEps * kEps));
    const float phi = std::asin(sin_mphi);
    const float cand_t[4] = {theta, theta, -theta, -theta};
    const float cand_p[4] = {phi, kPi - phi, phi, kPi - phi};
    int pick = -1;
    for (int i = 0; i < 4; i++) {
        V3 r = bev_to_pvcs(y_axis, cand_t[i], cand_p[i], 0.f);
        if (!(dot3(r, axis_pvcs) < 1 - kEps)) {
            pick = i;
            break;
        }
    }
    if (pick < 0) {
        set_error("beamlet canonical form: gantry/couch search failed");
        return FD_ERR_GEOM;
    }
    const float g = cand_t[pick], cpl = cand_p[pick];

    // collimator from the reference direction: +acos, else -acos
    const V3 co_pvcs = bev_to_pvcs(co_dir, bb.angles[0], bb.angles[1], bb.angles[2]);
    const V3 co_bev = v3(0.f, 0.f, -1.f);
    float cc = dot3(co_pvcs, bev_to_pvcs(co_bev, g, cpl, 0.f));
    cc = (cc < -1.0f) ? -1.0f : ((cc > 1.0f) ? 1.0f : cc);
    const float coll = std::acos(cc);
    float coll_r = coll;
    if (dot3(bev_to_pvcs(co_bev, g, cpl, coll_r), co_pvcs) < 1 - kEps * 10) {
        coll_r = -coll;
        if (dot3(bev_to_pvcs(co_bev, g, cpl, coll_r), co_pvcs) < 1 - kEps * 10) {
            set_error("beamlet canonical form: collimator search failed");
            return FD_ERR_GEOM;
        }
    }

    out.angles[0] = g;
    out.angles[1] = cpl;
    out.angles[2] = coll_r;
    const V3 shift = bev_to_pvcs(v3((ix + 0.5f - 0.5f * bb.fluence_dim[0]) * bb.beamlet_size[0], 0,
                                    (iy + 0.5f - 0.5f * bb.fluence_dim[1]) * bb.beamlet_size[1]),
                                 bb.angles[0], bb.angles[1], bb.angles[2]);
    out.isocenter[0] = bb.isocenter[0] + shift.x;
    out.isocenter[1] = bb.isocenter[1] + shift.y;
    out.isocenter[2] = bb.isocenter[2] + shift.z;
    out.beamlet_size[0] = width / bb.sub_on[0];
    out.beamlet_size[1] = height / bb.sub_on[1];
    out.fmap_size[0] = (uint32_t)bb.sub_dim[0];
    out.fmap_size[1] = (uint32_t)bb.sub_dim[1];
    out.sad = sad_r;
    out.long_spacing = bb.long_spacing;
    if (fluence) {
        const int nx = bb.sub_dim[0], ny = bb.sub_dim[1];
        std::fill(fluence, fluence + (size_t)nx * ny, 0.0f);
        const int lead_x = (int)((out.fmap_size[0] - bb.sub_on[0]) * 0.5f);
        const int lead_y = (int)((out.fmap_size[1] - bb.sub_on[1]) * 0.5f);
        for (int j = lead_y; j < lead_y + bb.sub_on[1]; j++)
            for (int i = lead_x; i < lead_x + bb.sub_on[0]; i++) fluence[i + j * nx] = 1.0f;
    }
    return calc_range(out, vol);
}

namespace {

// Linear-filtered binary mask > 0 at unnormalised texture coordinate c (texel i spans [i, i+1),
// border 0): true iff some texel with a nonzero weight is set.
bool mask_hit(const uint8_t* mask, const fd_volume& vol, const float c[3]) {
    int base[3];
    int reach[3];
    for (int d = 0; d < 3; d++) {
        float t = c[d] - 0.5f;
        float f = floorf(t);
        if (!(f > -2.0f && f < (float)vol.dim[d] + 1.0f)) return false;
        base[d] = (int)f;
        reach[d] = (t - f) > 0.0f ? 2 : 1;
    }
    for (int k = 0; k < reach[2]; k++) {
        const int z = base[2] + k;
        if (z < 0 || z >= (int)vol.dim[2]) continue;
        for (int j = 0; j < reach[1]; j++) {
            const int y = base[1] + j;
            if (y < 0 || y >= (int)vol.dim[1]) continue;
            for (int i = 0; i < reach[0]; i++) {
                const int x = base[0] + i;
                if (x < 0 || x >= (int)vol.dim[0]) continue;
                if (mask[(size_t)x + vol.dim[0] * ((size_t)y + (size_t)vol.dim[1] * z)]) return true;
            }
        }
    }
    return false;
}

}  // namespace

int flag_beam_setup(const fd_bundle& bb, const fd_volume& vol, FlagBeamDev& d) {
    fd_beam beam;
    std::memset(&beam, 0, sizeof(beam));
    for (int i = 0; i < 3; i++) {
        beam.isocenter[i] = bb.isocenter[i];
        beam.angles[i] = bb.angles[i];
    }
    beam.beamlet_size[0] = bb.beamlet_size[0];
    beam.beamlet_size[1] = bb.beamlet_size[1];
    beam.fmap_size[0] = (uint32_t)bb.fluence_dim[0];
    beam.fmap_size[1] = (uint32_t)bb.fluence_dim[1];
    beam.sad = bb.sad;
    beam.long_spacing = bb.long_spacing;
    std::memset(&d, 0, sizeof(d));
    d.fx = bb.fluence_dim[0];
    d.fy = bb.fluence_dim[1];
    const int r = calc_range(beam, vol);
    beam_matrix(beam.angles[0], beam.angles[1], beam.angles[2], true, d.Ri);
    for (int i = 0; i < 3; i++) d.src[i] = beam.source[i];
    d.bs[0] = beam.beamlet_size[0];
    d.bs[1] = beam.beamlet_size[1];
    d.sad = beam.sad;
    d.lim_min = beam.lim_min;
    d.ls = beam.long_spacing;
    d.L = r == FD_OK ? (int32_t)beam.long_dim : 0;
    return r;
}

int beamlet_flags(const fd_bundle& bb, const uint8_t* mask, const fd_volume& vol, uint8_t* flags) {
    FlagBeamDev d;
    const int fx = bb.fluence_dim[0], fy = bb.fluence_dim[1];
    if (flag_beam_setup(bb, vol, d) != FD_OK) {
        std::fill(flags, flags + (size_t)fx * fy, (uint8_t)0);
        return FD_ERR_GEOM;
    }
    const int ss = 3;  // lateral supersampling of d_beamletFlagInit (IMRTBeamBundle.cu:238)
    for (int idx = 0; idx < fx * fy; idx++) {
        const int bx = idx % fx, by = idx / fx;
        bool on = false;
        for (uint32_t i = 0; i < (uint32_t)d.L && !on; i++) {
            const float depth = d.lim_min + i * d.ls;
            const float scale = depth / d.sad;
            for (int j = 0; j < ss && !on; j++) {
                for (int k = 0; k < ss && !on; k++) {
                    const float dx = (bx + (j + 0.5f) / ss - fx * 0.5f) * d.bs[0];
                    const float dy = d.sad;
                    const float dz = (by + (k + 0.5f) / ss - fy * 0.5f) * d.bs[1];
                    const float px = fmaf(d.Ri[2], dz, fmaf(d.Ri[1], dy, d.Ri[0] * dx));
                    const float py = fmaf(d.Ri[5], dz, fmaf(d.Ri[4], dy, d.Ri[3] * dx));
                    const float pz = fmaf(d.Ri[8], dz, fmaf(d.Ri[7], dy, d.Ri[6] * dx));
                    const float cpos[3] = {fmaf(px, scale, d.src[0]) / vol.voxel_size[0],
                                           fmaf(py, scale, d.src[1]) / vol.voxel_size[1],
                                           fmaf(pz, scale, d.src[2]) / vol.voxel_size[2]};
                    on = mask_hit(mask, vol, cpos);
                }
            }
        }
        flags[idx] = on ? 1 : 0;
    }
    return FD_OK;
}

int ptv_isocenter(const uint8_t* mask, const fd_volume& vol, float iso[3]) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    size_t count = 0;
    for (uint32_t k = 0; k < vol.dim[2]; k++)
        for (uint32_t j = 0; j < vol.dim[1]; j++) {
            const uint8_t* row = mask + (size_t)vol.dim[0] * (j + (size_t)vol.dim[1] * k);
            for (uint32_t i = 0; i < vol.dim[0]; i++)
                if (row[i] > 0) {
                    sx += i + 0.5;
                    sy += j + 0.5;
                    sz += k + 0.5;
                    count++;
                }
        }
    if (count == 0) {
        set_error("empty PTV mask");
        return FD_ERR_ARG;
    }
    iso[0] = (float)(sx / count * vol.voxel_size[0]);
    iso[1] = (float)(sy / count * vol.voxel_size[1]);
    iso[2] = (float)(sz / count * vol.voxel_size[2]);
    return FD_OK;
}

void fill_beamlet_dev(const fd_beam& b, uint64_t bev_off, BeamletDev& d) {
    beam_matrix(b.angles[0], b.angles[1], b.angles[2], false, d.Rf);
    beam_matrix(b.angles[0], b.angles[1], b.angles[2], true, d.Ri);
    fill_beamlet_dev_no_rotation(b, bev_off, d);
}

void fill_beamlet_dev_no_rotation(const fd_beam& b, uint64_t bev_off, BeamletDev& d) {
    for (int i = 0; i < 3; i++) d.src[i] = b.source[i];
    d.bs[0] = b.beamlet_size[0];
    d.bs[1] = b.beamlet_size[1];
    d.sad = b.sad;
    d.lim_min = b.lim_min;
    d.ls = b.long_spacing;
    d.kx = b.sad / b.beamlet_size[0];
    d.kz = b.sad / b.beamlet_size[1];
    d.inv_ls = 1.0f / b.long_spacing;
    d.L = (int32_t)b.long_dim;
    d.geom = 0;
    d.bev_off = bev_off;
}

}  // namespace fdi
