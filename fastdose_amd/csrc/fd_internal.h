// fd_internal.h -- types shared by the host code and the HIP kernels of libfastdose_hip.
#pragma once

#include <cstdint>
#include <cstddef>
#include <string>

#include "fastdose_hip.h"

namespace fdi {

// Kernel shape constants.  The dose-mode BEV grid is 16 x 16 pixels (subFluenceDim, default
// 16 in IMRTArgs.cpp:51); a wavefront carries all 256 rays of one cone as 4 rays per lane.
constexpr int kFmap = 16;
constexpr int kNpix = kFmap * kFmap;
constexpr int kWave = 64;
constexpr int kRaysPerLane = kNpix / kWave;
constexpr int kMaxPhi = 24;       // azimuths per cone (MAX_KERNEL_NUM, kernel.cuh:7)
constexpr int kSuperSampling = 4; // Terma substeps per slice (TermaComp.cu:65)
constexpr int kResampleSS = 5;    // BEV -> PVCS supersampling per axis (IMRTgeom.cu:352)

// Upper bound on lattice steps of one azimuth within one slice.  A forward cone needs at most
// 2 + 2 (ls / bs) tan(theta) (|cos phi| + |sin phi|) steps per slice; forward cones must have
// cos(theta_mid) >= kMinCosTheta (checked when a context is created), which keeps the count far
// below this cap, so the cap only guarantees that every wavefront's loop ends.
constexpr int kMaxStepsPerSlice = 1 << 16;
constexpr double kMinCosTheta = 1e-2;

constexpr float kLongRayVoxels = 256.0f;  // FD_OPT_RPL_FP64: FP64 path on rays longer than this
constexpr float kEps = 1e-4f;        // eps_fastdose (macros.h.in:26)
constexpr float kLargerEps = 1e-3f;  // larger_eps_fastdose (macros.h.in:27)

// Per-beamlet descriptor the kernels read (the role of d_BEAM_d, beam.cuh:84-133), with the
// host-rounded FP32 transforms every integer decision is made from.
struct BeamletDev {
    float Rf[9];       // PVCS -> BEV rotation, row-major
    float Ri[9];       // BEV -> PVCS rotation, row-major
    float src[3];      // source position [cm]
    float bs[2];       // beamlet_size (BEV pixel pitch at SAD) [cm]
    float sad;
    float lim_min;
    float ls;          // long_spacing [cm]
    float kx, kz;      // sad / bs.x, sad / bs.y
    float inv_ls;      // 1 / long_spacing
    int32_t L;         // long_dim
    int32_t geom;      // march geometry of the batch (CCCS march tables), else 0
    uint64_t bev_off;  // float offset of this beamlet's BEV volume
};
static_assert(sizeof(BeamletDev) % 8 == 0, "BeamletDev alignment");

// Per-bundle input of the beamlet-PTV intersection kernel (d_beamletFlagInit,
// IMRTBeamBundle.cu:130-316): the bundle's central beam after calc_range.
struct FlagBeamDev {
    float Ri[9];       // BEV -> PVCS rotation, row-major
    float src[3];
    float bs[2];       // beamlet_size of the bundle
    float sad;
    float lim_min;
    float ls;
    int32_t L;         // long_dim (0: the central beam misses the volume, all flags 0)
    int32_t fx, fy;    // fluence_dim
    uint32_t out_off;  // first flag of this bundle in the output
};

// Spectrum folded for the Terma loop: w_e = fluence_e * E_e * mu_e and mu_e.
// The Terma kernel evaluates sum_e w_e exp(-mu_e r) from a table over the radiological depth r:
// entry i holds (f(r_i), f(r_{i+1}) - f(r_i)) at r_i = i h, linear in between (h is chosen so the
// interpolation error stays below ~5e-6 of f).
struct SpectrumDev {
    int32_t n;
    float w[FD_MAX_TABLE];
    float mu[FD_MAX_TABLE];
    int32_t tab_n;      // table entries
    float tab_inv_h;    // 1 / h
    const float* tab;   // device table, (f, df) pairs
};
constexpr double kSpecTableStep = 0.02;  // h <= 0.02 g/cm^2

// Forward cones of the kernel table, heavy (widest) cone first.
struct ConeTable {
    int32_t n_fwd;        // forward cones (theta_mid <= pi/2)
    int32_t n_phi;
    int32_t theta_idx[FD_MAX_TABLE];  // original cone index of forward slot t
    float A_a[FD_MAX_TABLE];          // A/a
    float B_b[FD_MAX_TABLE];          // B/b
    float a[FD_MAX_TABLE];
    float b[FD_MAX_TABLE];
    float na[FD_MAX_TABLE];           // -a log2(e), -b log2(e): e^{-a p} = 2^{na p}
    float nb[FD_MAX_TABLE];
    float ux[FD_MAX_TABLE][kMaxPhi];  // unit direction of (theta, phi), FP64-rounded
    float uy[FD_MAX_TABLE][kMaxPhi];
    float uz[FD_MAX_TABLE][kMaxPhi];
};

// CCCS march table of one (march geometry, forward cone): the lattice walk of every azimuth over
// the longest beamlet of that geometry, chunk by chunk.  The walk and its decisions depend only on
// the BEV pixel pitch, the longitudinal spacing and the cone, so every beamlet sharing those three
// values reads the same table.  Records: per azimuth a stream of rec_cap 32-byte records (two
// float4 each) at float4 index rec_base + 2 rec_cap f; chunk headers: kTabHdr ints per chunk at
// int index hdr_base, [f] the step count of azimuth f, [kTabHdr - 2] slice done, [kTabHdr - 1]
// next slice; the chunk after the last holds no steps and ends the walk.
constexpr int kTabHdr = 16;
struct TabEntry {
    float bs0, bs1, ls;  // the geometry (bit patterns shared by its beamlets)
    int32_t Lmax;        // longest long_dim among them
    uint32_t rec_cap;    // records per azimuth
    uint32_t chunk_cap;  // chunk headers, the terminating one included
    uint64_t rec_base;   // float4 index of azimuth 0's stream
    uint64_t hdr_base;   // int index of chunk 0's header
    uint64_t pad;
};
static_assert(sizeof(TabEntry) == 48, "TabEntry layout");

void set_error(const std::string& msg);

}  // namespace fdi
