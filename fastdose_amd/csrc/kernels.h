// kernels.h -- launch wrappers of the libfastdose_hip kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "fd_internal.h"

namespace fdi {

// Destination of compacted CSR entries: columns, values and the entries they hold.
struct CsrDst {
    uint64_t* columns = nullptr;
    float* values = nullptr;
    uint64_t capacity = 0;
};

// Beamlet / PTV intersection flags of n bundles (d_beamletFlagInit, IMRTBeamBundle.cu:130-316);
// flags[beams[b].out_off + pixel], max_pixels = the largest fluence_dim^2.
hipError_t launch_beamlet_flags(const FlagBeamDev* beams, int n, int max_pixels, const uint8_t* mask,
                                const fd_volume& vol, uint8_t* flags, hipStream_t stream);

// Per-voxel trilinear corner records of the density (8 floats per base voxel of the zero-bordered
// volume, 4 x 4 x 4 bricks) for the Terma kernel; density_records_floats(vol) floats.
uint64_t density_records_floats(const fd_volume& vol);
hipError_t launch_expand_density(const float* density, const fd_volume& vol, float* records, hipStream_t stream);

// Terma + mean density along the 256 BEV rays of each beamlet (TermaComp.cu:179-253).
// density_pad: the corner records of launch_expand_density.
// scratch: 4 floats per BEV value (the layout of terma, 16-byte aligned), overwritten.
// segments: ray segments per ray (1, 2, 4, 8 or 16); a workgroup holds min(256, 1024 / segments)
// rays times `segments` threads, so with 8 or 16 a beamlet takes two or four workgroups.
// Beamlets with long_dim * long_spacing > rpl64_len [cm] accumulate the radiological path in FP64.
hipError_t launch_terma(const BeamletDev* beams, int n, const float* fluence, const float* density_pad,
                        const fd_volume& vol, const SpectrumDev& spec, float rpl64_len, float* terma,
                        float* dens_bev, float* scratch, int segments, hipStream_t stream);

// March tables of a batch (TabEntry): n_gt = (march geometries) x n_fwd entries, the record
// streams and chunk headers they index, and the flag a table that outgrows its capacity sets.
struct CccsTables {
    const TabEntry* ent;
    float4* rec;
    int* hdr;
    uint32_t* overflow;
    int n_gt;
};

// The march tables (one wavefront each); clears *overflow first.  n_phi <= cccs_table_max_phi().
hipError_t launch_cccs_tables(const CccsTables& tab, const ConeTable* cones_dev, int n_fwd, int n_phi,
                              hipStream_t stream);
int cccs_table_max_phi();
int cccs_chunk_steps();  // lattice steps of one azimuth per chunk at most

// CCCS over the forward cones (DoseComp.cu:217-392): per-cone partial BEV doses
// partial[n_fwd * bev_off_b + t * L_b * 256 + s * 256 + p] (at least n_fwd times the batch's BEV
// floats), then their sum into dose (four floats per thread when vec4: every bev_off_b a multiple
// of four).
// ev (nullable): events recorded before the CCCS kernel, between it and the cone sum, and after.
// tab (nullable, n_phi <= cccs_table_max_phi()): the beamlets' march tables, built on this stream
// beforehand (launch_cccs_tables), beams[b].geom indexing them; the CCCS reads its chunks from
// them, with the results of the march wavefront.
hipError_t launch_cccs(const BeamletDev* beams, int n, int max_L, const ConeTable* cones_dev, int n_fwd,
                       int n_phi, bool exact_g, const float* terma, const float* dens_bev, float* partial,
                       float* dose, bool vec4, hipStream_t stream, hipEvent_t* ev = nullptr,
                       const CccsTables* tab = nullptr);

// Resampling into CSR rows.  Stage 1 counts candidate voxels per (beamlet, z-plane) segment and
// records the rows [y_lo, y_hi) of the segment that hold them in cand_rows ((0, 0) when none).
hipError_t launch_resample_count(const BeamletDev* beams, int n, const fd_volume& vol, uint32_t* cand_count,
                                 uint2* cand_rows, hipStream_t stream);
// Stage 2: values of every candidate, nonzeros compacted in voxel order into tmp_cols/tmp_vals
// at cand_off[b * Z + z]; nonzero counts into nnz_count.  A segment whose candidates would end
// beyond `cap` entries writes nothing, gets a zero count and sets overflow[0] = 1.
// rowmax_bits (nullable, zeroed by the caller): per-beamlet maximum value, as float bits.
// rcp_exact: launch_rcp_check found no mismatch on this device (see there).
hipError_t launch_resample_values(const BeamletDev* beams, int n, const fd_volume& vol, float extent,
                                  const float* dose_bev, const uint64_t* cand_off, const uint2* cand_rows,
                                  uint64_t cap, uint32_t* overflow, uint64_t* tmp_cols, float* tmp_vals,
                                  uint32_t* nnz_count, uint32_t* rowmax_bits, bool rcp_exact, hipStream_t stream);
// Adds to *mismatches (zeroed by the caller) the number of floats b with exponent in
// [exp0, exp0 + n_exp) for which the resampler's reciprocal (hardware reciprocal + one Newton
// step) differs from the IEEE quotient 1 / b.
hipError_t launch_rcp_check(int exp0, int n_exp, uint32_t* mismatches, hipStream_t stream);
// Relative threshold: per segment, the entries >= thr * rowmax of its beamlet.
hipError_t launch_resample_threshold_count(int n, int Z, const uint64_t* cand_off, const uint32_t* nnz_count,
                                           const float* tmp_vals, const uint32_t* rowmax_bits, float thr,
                                           uint32_t* keep_count, hipStream_t stream);
// Stage 3: copy the compacted nonzeros into the CSR arrays (with rowmax_bits and thr > 0 only
// those >= thr * rowmax); row offsets from nnz_off.  The entries go to dst0 when its capacity
// holds the total (nnz_off[n Z]), else to dst1 when that one does, else nowhere (the columns and
// values are left untouched; the offsets are still written).  A destination with a null array
// counts as capacity 0.
hipError_t launch_resample_compact(int n, int Z, const uint64_t* cand_off, const uint64_t* nnz_off,
                                   const uint32_t* nnz_count, const uint64_t* tmp_cols, const float* tmp_vals,
                                   const uint32_t* rowmax_bits, float thr, uint64_t* offsets, CsrDst dst0,
                                   CsrDst dst1, hipStream_t stream);

// Exclusive scan of count[0..N) (u32) into off[0..N) (u64) plus total[0] = sum.  With scratch
// (scan_scratch_size(N) u64, nullable) large inputs take a three-pass multi-workgroup scan.
hipError_t launch_scan_u32(const uint32_t* count, uint64_t N, uint64_t* off, uint64_t* total, uint64_t* scratch,
                           hipStream_t stream);
uint64_t scan_scratch_size(uint64_t N);

// Dense (rows x cols, row-major) -> CSR: per-row nonzero counts, then fill.
hipError_t launch_dense_count(const float* dense, uint64_t rows, uint64_t cols, uint32_t* row_count,
                              hipStream_t stream);
hipError_t launch_dense_fill(const float* dense, uint64_t rows, uint64_t cols, const uint64_t* row_off,
                             uint64_t* columns, float* values, hipStream_t stream);
// offsets[r] = row_off[r] (r < rows), offsets[rows] = total[0].
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream);
hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t stream);
hipError_t launch_write_offsets(const uint64_t* row_off, const uint64_t* total, uint64_t rows, uint64_t* offsets,
                                hipStream_t stream);

// Packing and D^T X products (kernels_spmm.hip).
hipError_t launch_global_rows(int n_mat, const uint64_t* row_first, const uint64_t* off_off, const uint64_t* nnz_off,
                              const uint64_t* offsets, uint64_t R, uint64_t total_nnz, uint64_t* rowptr,
                              hipStream_t stream);
// scale[r] = row maximum; q (nullable) = value / scale in BF16 or FP16
hipError_t launch_pack_rows(uint64_t R, const uint64_t* rowptr, const float* values, int vtype, float* scale,
                            uint16_t* q, hipStream_t stream);
hipError_t launch_col_count(uint64_t nnz, const uint64_t* columns, uint32_t* count, hipStream_t stream);
hipError_t launch_col_scatter(uint64_t R, const uint64_t* rowptr, const uint64_t* columns, const void* vals,
                              int vbytes, const uint64_t* colptr, uint32_t* cursor, uint32_t* crow, void* cval,
                              hipStream_t stream);
hipError_t launch_col_sort(uint64_t C, const uint64_t* colptr, uint32_t* crow, void* cval, int vbytes,
                           hipStream_t stream);
hipError_t launch_spmm(int vtype, uint64_t C, const uint64_t* colptr, const uint32_t* crow, const void* cval,
                       const float* scale, const float* X, int k, float* Y, hipStream_t stream);

}  // namespace fdi
