// kernels.h -- launch wrappers of the libfastdose_hip kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "fd_internal.h"

namespace fdi {

// Terma + mean density along the 256 BEV rays of each beamlet (TermaComp.cu:179-253).
hipError_t launch_terma(const BeamletDev* beams, int n, const float* fluence, const float* density,
                        const fd_volume& vol, const SpectrumDev& spec, bool rpl_fp64, float* terma,
                        float* dens_bev, hipStream_t stream);

// CCCS over the forward cones (DoseComp.cu:217-392): per-cone partial BEV doses
// partial[(b * n_fwd + t) * stride + s * 256 + p], then their sum into dose.
// ev (nullable): events recorded before the CCCS kernel, between it and the cone sum, and after.
hipError_t launch_cccs(const BeamletDev* beams, int n, int max_L, const ConeTable* cones_dev, int n_fwd,
                       int n_phi, bool exact_g, const float* terma, const float* dens_bev, float* partial,
                       float* dose, uint64_t stride, hipStream_t stream, hipEvent_t* ev = nullptr);

// Resampling into CSR rows.  Stage 1 counts candidate voxels per (beamlet, z-plane).
hipError_t launch_resample_count(const BeamletDev* beams, int n, const fd_volume& vol, uint32_t* cand_count,
                                 hipStream_t stream);
// Stage 2: values of every candidate, nonzeros compacted in voxel order into tmp_cols/tmp_vals
// at cand_off[b * Z + z]; nonzero counts into nnz_count.
hipError_t launch_resample_values(const BeamletDev* beams, int n, const fd_volume& vol, float extent,
                                  const float* dose_bev, const uint64_t* cand_off, uint64_t* tmp_cols,
                                  float* tmp_vals, uint32_t* nnz_count, hipStream_t stream);
// Stage 3: copy the compacted nonzeros into the CSR arrays; row offsets from nnz_off.
hipError_t launch_resample_compact(int n, int Z, const uint64_t* cand_off, const uint64_t* nnz_off,
                                   const uint32_t* nnz_count, const uint64_t* tmp_cols, const float* tmp_vals,
                                   uint64_t* offsets, uint64_t* columns, float* values, hipStream_t stream);

// Exclusive scan of count[0..N) (u32) into off[0..N) (u64) plus off_total[0] = sum.
hipError_t launch_scan_u32(const uint32_t* count, uint64_t N, uint64_t* off, uint64_t* total,
                           hipStream_t stream);

// Dense (rows x cols, row-major) -> CSR: per-row nonzero counts, then fill.
hipError_t launch_dense_count(const float* dense, uint64_t rows, uint64_t cols, uint32_t* row_count,
                              hipStream_t stream);
hipError_t launch_dense_fill(const float* dense, uint64_t rows, uint64_t cols, const uint64_t* row_off,
                             uint64_t* columns, float* values, hipStream_t stream);
// offsets[r] = row_off[r] (r < rows), offsets[rows] = total[0].
hipError_t launch_write_offsets(const uint64_t* row_off, const uint64_t* total, uint64_t rows, uint64_t* offsets,
                                hipStream_t stream);

}  // namespace fdi
