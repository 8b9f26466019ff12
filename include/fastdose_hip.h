/* fastdose_hip.h -- C-ABI of the MI355X-native CCCS beamlet-dose library (libfastdose_hip.so).
 *
 * This boundary replaces the C++/CUDA entry points that FastDose's dose mode (`IMRT --mode 0`)
 * links against.  Each function below names the reference interface it stands in for
 * (paths relative to the reference repository).  Everything here is plain C: POD structs,
 * raw pointers and sizes, `void*` for a hipStream_t.  No torch or HIP types appear.
 *
 * Conventions.
 *   - Every function returns an int status: FD_OK (0) on success, a negative FD_ERR_* code on
 *     failure.  The library never calls exit().  fd_last_error() returns a message for the
 *     last failure on the calling thread.  (The reference returns bool with true = failure and
 *     exits on CUDA errors.)
 *   - "dev" pointers are device (HBM) pointers; everything else is host memory.
 *   - Beam/beamlet arrays (fd_beam) are host arrays; the library uploads what its kernels need.
 *   - A context (fd_ctx) owns the device-resident density and the kernel/spectrum tables, so
 *     several contexts may coexist in one process (the reference keeps these tables in
 *     process-global __constant__ memory).  Calls on one context must not overlap in time.
 *   - BEV volumes of n beamlets live in one buffer: beamlet i occupies floats
 *     [i*stride, i*stride + long_dim_i*npix), slice-major, pixel index ix + fmap_x*iy.
 */
#ifndef FASTDOSE_HIP_H
#define FASTDOSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FD_MAX_TABLE 24 /* MAX_KERNEL_NUM / MAX_THETA_ANGLES (spectrum.cuh:7, kernel.cuh:7) */

enum {
    FD_OK = 0,
    FD_ERR_ARG = -1,         /* invalid argument / size */
    FD_ERR_HIP = -2,         /* HIP runtime error */
    FD_ERR_GEOM = -3,        /* beam misses the bbox, canonical-form search failed */
    FD_ERR_IO = -4,          /* file could not be read or written */
    FD_ERR_CAPACITY = -5,    /* caller buffer too small; the needed size is reported */
    FD_ERR_UNSUPPORTED = -6  /* shape the kernels do not implement (e.g. fmap != 16x16) */
};

/* Option bits for fd_ctx_set_options. */
enum {
    FD_OPT_EXACT_G = 1,      /* use g = (1-e^-x)/x (KernelGen/derivation.md:248) instead of the
                                reference's g == 1 (DoseComp.cu:167-175) */
    FD_OPT_RPL_FP64 = 2,     /* accumulate the Terma radiological path in FP64 on rays longer
                                than 256 voxels (long_dim * long_spacing > 256 * min voxel size) */
    FD_OPT_TIMING = 4,       /* record HIP events around every kernel launch (fd_ctx_timings) */
    FD_OPT_RPL_FP64_ALL = 8, /* FP64 radiological path on every ray */
    FD_OPT_GRAPH = 16,       /* batched paths (fd_dosemat_add_beam(s), serving) replay each batch as one
                                captured HIP graph instead of separate launches */
    FD_OPT_IEEE_DIV = 32,    /* resampler: correct the refined hardware reciprocal on depths whose
                                significand is all ones even where the context's device check found
                                it exact there (results are identical either way; no counterpart
                                in the reference) */
    FD_OPT_TABLES = 64,      /* batched paths: always walk the CCCS lattice once per march geometry
                                (BEV pitch, longitudinal spacing) into tables the beamlets share;
                                by default only when a batch has 4 or more beamlets per geometry
                                (results are identical either way; FD_ERR_UNSUPPORTED beyond 10
                                azimuths; no counterpart in the reference) */
    FD_OPT_NO_TABLES = 128   /* batched paths: never use the tables (each CCCS workgroup marches) */
};

/* Value types of packed dose-loading matrices (fd_dosemat_pack, fd_spmm_create). */
enum {
    FD_VAL_F32 = 0,  /* the canonical FP32 values */
    FD_VAL_BF16 = 1, /* bfloat16 of value / row scale (row scale = the beamlet's maximum, FP32) */
    FD_VAL_FP16 = 2  /* IEEE half of value / row scale */
};

/* Kernel slots reported by fd_ctx_timings. */
enum {
    FD_T_TERMA = 0,
    FD_T_CCCS = 1,
    FD_T_CONE_SUM = 2,
    FD_T_RESAMPLE_COUNT = 3,
    FD_T_RESAMPLE_VALUES = 4,
    FD_T_RESAMPLE_COMPACT = 5,
    FD_T_CCCS_TABLES = 6,  /* the march tables of a batch (FD_OPT_TABLES) */
    FD_T_COUNT = 7
};

/* DENSITY_h geometry (density.cuh:9-28): voxel size [cm], volume dims (x fastest), bbox. */
typedef struct fd_volume {
    float voxel_size[3];
    uint32_t dim[3];
    uint32_t bbox_start[3];
    uint32_t bbox_dim[3];
} fd_volume;

/* BEAM_h (beam.cuh:10-34) without the fluence vector: angles = (gantry/theta, couch/phi,
 * collimator) in rad; lim_min/lim_max/long_dim/source are filled by fd_calc_range. */
typedef struct fd_beam {
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
} fd_beam;

/* SPECTRUM_h (spectrum.cuh:10-20): fluence normalised to sum 1. */
typedef struct fd_spectrum {
    int32_t n;
    float energy[FD_MAX_TABLE];
    float fluence[FD_MAX_TABLE];
    float mu[FD_MAX_TABLE];
    float mu_en[FD_MAX_TABLE];
} fd_spectrum;

/* KERNEL_h (kernel.cuh:10-25): per zenith cone A,a,B,b and theta bounds; phi grid. */
typedef struct fd_kernel {
    int32_t n_theta;
    int32_t n_phi;
    float theta_begin[FD_MAX_TABLE];
    float theta_end[FD_MAX_TABLE];
    float theta_mid[FD_MAX_TABLE];
    float A[FD_MAX_TABLE];
    float a[FD_MAX_TABLE];
    float B[FD_MAX_TABLE];
    float b[FD_MAX_TABLE];
    float phi[FD_MAX_TABLE];
} fd_kernel;

/* Scalar fields of IMRT::BeamBundle (IMRTBeamBundle.cuh:9-30). */
typedef struct fd_bundle {
    float isocenter[3];
    float beamlet_size[2];
    int32_t fluence_dim[2];
    float sad;
    float angles[3];
    float long_spacing;
    int32_t sub_dim[2];
    int32_t sub_on[2];
} fd_bundle;

/* A CSR matrix (rows = beamlets, columns = PVCS voxels x + X*(y + Y*z)), 0-based, 64-bit
 * indices, ascending columns, exact zeros dropped -- the layout cuSPARSE DenseToSparse
 * produces in MatCSREnsemble::addMat (IMRTSparseMatEns.cu:66-165). */
typedef struct fd_csr {
    uint64_t rows;
    uint64_t nnz;
    uint64_t capacity;   /* entries allocated in columns/values */
    uint64_t* offsets;   /* dev, rows + 1 */
    uint64_t* columns;   /* dev, capacity */
    float* values;       /* dev, capacity */
} fd_csr;

typedef struct fd_ctx fd_ctx;
typedef struct fd_dosemat fd_dosemat;

const char* fd_last_error(void);
const char* fd_version(void);

/* ------------------------------------------------------------------ host geometry (no GPU) */

/* BEAM_h::calc_range (beam.cu:134-221): source, lim_min, lim_max, long_dim. FD_ERR_GEOM if the
 * beam misses the bbox. */
int fd_calc_range(fd_beam* beam, const fd_volume* vol);

/* BeamBundle::beamletInit (IMRTBeamBundle.cu:341-569): canonical beamlet (ix, iy) of a bundle,
 * its sub-fluence (sub_dim[0]*sub_dim[1] floats, centre sub_on block = 1) and calc_range. */
int fd_beamlet_init(const fd_bundle* bundle, int ix, int iy, const fd_volume* vol, fd_beam* out,
                    float* fluence_out);

/* BeamBundle::beamletsInit (IMRTBeamBundle.cu:319-338): every flagged beamlet, row order
 * (iy outer, ix inner).  out/fluence_out hold room for all fluence_dim^2 beamlets. */
int fd_beamlets_init(const fd_bundle* bundle, const uint8_t* flags, const fd_volume* vol, fd_beam* out,
                     float* fluence_out, int* n_out);

/* BeamBundleInit's isocenter (IMRTBeamBundle.cu:36-65): PTV centroid [cm]. */
int fd_ptv_isocenter(const uint8_t* mask, const fd_volume* vol, float iso[3]);

/* beamletFlagInit / d_beamletFlagInit (IMRTBeamBundle.cu:130-316) on the host: flags of the
 * fluence_dim^2 beamlets that intersect the PTV mask. */
int fd_beamlet_flags_host(const fd_bundle* bundle, const uint8_t* ptv_mask, const fd_volume* vol,
                          uint8_t* flags);

/* beamletFlagInit / d_beamletFlagInit (IMRTBeamBundle.cu:130-316) on the device, for n_bundles
 * bundles at once: mask_dev is the u8 PTV mask of the context's volume in device memory (x
 * fastest); flags_host receives fluence_dim^2 flags per bundle, concatenated in bundle order.
 * Bit-identical to fd_beamlet_flags_host; a bundle whose central beam misses the volume gets all-zero
 * flags (fd_beamlet_flags_host reports FD_ERR_GEOM for it). */
int fd_beamlet_flags(fd_ctx* ctx, const uint8_t* mask_dev, const fd_bundle* bundles, int n_bundles,
                     uint8_t* flags_host, void* stream);

/* SPECTRUM_h::read_spectrum_file (spectrum.cu:14-60). */
int fd_read_spectrum(const char* path, fd_spectrum* out);
/* KERNEL_h::read_kernel_file (kernel.cu:13-72). */
int fd_read_kernel(const char* path, int n_phi, fd_kernel* out);

/* ------------------------------------------------------------------ device context */

/* density_h2d + bind_spectrum + bind_kernel (density.cu:23-58, TermaComp.cu:67-78,
 * DoseComp.cu:25-38): uploads the density volume and tables once.  The stage and batched calls on
 * one context share its scratch: issue them from one thread, on one stream at a time (a server,
 * fd_serve_start, gives each of its slots a stream and scratch of its own). */
int fd_ctx_create(int device, const fd_volume* vol, const float* density_host, const fd_spectrum* spectrum,
                  const fd_kernel* kernel, fd_ctx** out);
int fd_ctx_destroy(fd_ctx* ctx);
int fd_ctx_set_options(fd_ctx* ctx, int options);
/* Relative CSR threshold: the resampler keeps only entries >= rel * (the beamlet's maximum value).
 * 0 (default) drops exact zeros only, as cuSPARSE DenseToSparse in MatCSREnsemble::addMat
 * (IMRTSparseMatEns.cu:108-161) does; the reference has no threshold. */
int fd_ctx_set_threshold(fd_ctx* ctx, float rel);
/* Milliseconds each kernel slot took in the last call that launched it (FD_OPT_TIMING set),
 * measured with HIP events on the launch stream and summed over that call's launches (a batched
 * call launches each kernel once per chunk; graph replay records no events); -1 for slots not run.
 * Plays the role of the reference's cudaEvent timers (TermaComp.cu:310-333, DoseComp.cu:449-472). */
int fd_ctx_timings(fd_ctx* ctx, float* ms, int n);
/* How many launches the sums of fd_ctx_timings cover, per slot (0 for slots not run). */
int fd_ctx_timing_launches(fd_ctx* ctx, int* launches, int n);
/* Device pointer of the resident density volume (x fastest). */
int fd_ctx_density(fd_ctx* ctx, const float** density_dev);

/* TermaComputeCollective (TermaComp.cuh:7-17): Terma and mean density along the BEV rays of
 * n beamlets.  fluence_dev: n * npix floats. */
int fd_terma(fd_ctx* ctx, const fd_beam* beams, int n, const float* fluence_dev, float* terma_dev,
             float* density_bev_dev, uint64_t stride, void* stream);

/* DoseComputeCollective (DoseComp.cuh:9-19): CCCS recursion over the forward cones; writes
 * (does not accumulate) the BEV dose of each beamlet. */
int fd_cccs(fd_ctx* ctx, const fd_beam* beams, int n, const float* terma_dev, const float* density_bev_dev,
            float* dose_bev_dev, uint64_t stride, void* stream);

/* BEV2PVCSInterp + MatCSREnsemble::addMat (IMRTgeom.cuh:14-33, IMRTDoseMatEns.cuh:30): 5^3
 * point-sampled BEV -> PVCS resampling of n beamlets straight into CSR rows (no dense
 * intermediate).  out->offsets must hold n+1 entries; if out->capacity < nnz the call returns
 * FD_ERR_CAPACITY with out->nnz set (two-phase use; the offsets are written either way, the
 * columns and values only on success). */
int fd_resample_csr(fd_ctx* ctx, const fd_beam* beams, int n, const float* dose_bev_dev, uint64_t stride,
                    float extent, fd_csr* out, void* stream);

/* cusparseDenseToSparse as used by MatCSREnsemble::addMat (IMRTSparseMatEns.cu:78-161):
 * row-major dense (dev) -> CSR.  Same two-phase capacity rule as fd_resample_csr. */
int fd_dense_to_csr(const float* dense_dev, uint64_t rows, uint64_t cols, fd_csr* out, void* stream);

/* fd_dense_to_csr without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device or pinned memory (tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and two-phase capacity rule as fd_dense_to_csr; its per-row counts and
 * row offsets live in a library-held device array of about 12 MiB per device (held until the
 * process ends), so rows is limited to FD_NOALLOC_MAX_ROWS (FD_ERR_ARG beyond).  Every call on
 * one device shares that array: calls that overlap in time on one device race.  The call
 * synchronises `stream` before it returns. */
#define FD_NOALLOC_MAX_ROWS (1u << 20)
int fd_dense_to_csr_noalloc(const float* dense_dev, uint64_t rows, uint64_t cols, fd_csr* out, void* stream);

/* Number of (ray, segment) steps of the CCCS march of each beamlet (host count of the same FP32
 * lattice the kernel walks); used for the FLOP accounting and load balancing. */
int fd_cccs_segments(fd_ctx* ctx, const fd_beam* beams, int n, int64_t* nseg_out);

/* ------------------------------------------------------------------ dose-loading matrix */

/* MatCSREnsemble (IMRTDoseMatEns.cuh:12-60): one CSR matrix per beam, n_cols = PVCS voxels. */
int fd_dosemat_create(uint64_t n_cols, uint64_t est_nnz, fd_dosemat** out);
int fd_dosemat_destroy(fd_dosemat* m);
/* One iteration of DoseMatConstruction (IMRTDoseMat.cu:183-268): Terma -> CCCS -> resample of
 * one beam's n beamlets, appended as the next matrix.  fluence: n*npix host floats, as
 * fd_beamlet_init writes them. */
int fd_dosemat_add_beam(fd_ctx* ctx, fd_dosemat* m, const fd_beam* beamlets, int n, const float* fluence,
                        float extent, void* stream);
/* Batched DoseMatConstruction: n_beams beams whose beamlets are concatenated in `beamlets`
 * (counts[k] beamlets for beam k, fluence n_total*npix host floats) go through the kernels
 * together -- all beams of a chunk in one launch per stage, chunks sized to `max_chunk` beamlets
 * (0: library default) -- and are appended as n_beams matrices, in order.  Each beam's matrix is
 * identical to what fd_dosemat_add_beam gives for it alone. */
int fd_dosemat_add_beams(fd_ctx* ctx, fd_dosemat* m, const fd_beam* beamlets, const float* fluence,
                         const int* counts, int n_beams, int max_chunk, float extent, void* stream);

/* ---- Launch tuning (no counterpart in the reference, which launches one beam at a time with
 * fixed shapes): the parameters of the batched paths, chosen by fd_autotune. ---- */
typedef struct fd_tuning {
    int batch_beamlets;   /* beamlets per batch when fd_dosemat_add_beams / serving are given 0 */
    int graph;            /* 1: replay each batch as one captured HIP graph (FD_OPT_GRAPH) */
    int terma_segments;   /* Terma ray segments: 1, 2, 4 (wavefront groups of 256 rays), 8 or 16
                             (two workgroups of 128 or four of 64 rays per beamlet) */
    float beamlets_per_s; /* throughput of this setting in the tuning run (0: defaults, not tuned) */
} fd_tuning;

int fd_ctx_get_tuning(const fd_ctx* ctx, fd_tuning* out);
int fd_ctx_set_tuning(fd_ctx* ctx, const fd_tuning* t);

/* Times trial runs of fd_dosemat_add_beams over the given beams (whole beams in order; a few dozen
 * representative beams suffice): batch sizes of 1024 .. 16384 beamlets, then graph replay on and
 * off, then 4, 8 or 16 Terma segments, each setting run twice and its faster run kept.  The fastest
 * setting becomes the context's tuning and is returned in *out.  Synchronous; the trial matrices
 * are discarded. */
int fd_autotune(fd_ctx* ctx, const fd_beam* beamlets, const float* fluence, const int* counts, int n_beams,
                float extent, fd_tuning* out, void* stream);
/* Appends an existing device CSR (copied). */
int fd_dosemat_add_csr(fd_dosemat* m, const fd_csr* csr, void* stream);
int fd_dosemat_info(const fd_dosemat* m, uint64_t* n_matrices, uint64_t* total_rows, uint64_t* total_nnz);
/* Host copy of matrix k: offsets (rows+1), columns, values; sizes from fd_dosemat_matrix_size. */
int fd_dosemat_matrix_size(const fd_dosemat* m, uint64_t k, uint64_t* rows, uint64_t* nnz);
int fd_dosemat_copy_matrix(const fd_dosemat* m, uint64_t k, uint64_t* offsets, uint64_t* columns,
                           float* values);
/* MatCSREnsemble::tofile (IMRTSparseMatEns.cu:168-239): the doseMatFolder layout read by
 * scripts/readDose.m. */
int fd_dosemat_tofile(const fd_dosemat* m, const char* folder);
/* MatCSREnsemble::fromfile (IMRTSparseMatEns.cu:242-336).  The folder is read and checked on the
 * host first (table sizes, every matrix's offsets inside offsetsBuffer, starting at 0, non-decreasing
 * and ending at its nonzero count, every column below the ensemble's n_cols); FD_ERR_IO leaves the
 * ensemble as it was.  Replaces the ensemble's contents. */
int fd_dosemat_fromfile(fd_dosemat* m, const char* folder);

/* scripts/doseMerge.py:9-63 (doseMatMerge): concatenates the doseMatFolder file sets of n_folders
 * folders, in order, into out_folder -- NonZeroElements, numRowsPerMat, offsetsBuffer,
 * columnsBuffer, valuesBuffer and fluenceMap.bin byte for byte as that script writes them -- and
 * also writes CumuNonZeroElements, CumuNumRowsPerMat and OffsetBufferIdx of the merged set (which
 * the script leaves out), so fd_dosemat_fromfile reads the result.  Every input is checked as
 * fd_dosemat_fromfile checks it before anything is written.  Host only (no device). */
int fd_dosemat_merge_folders(const char* const* folders, int n_folders, const char* out_folder);

/* 16-bit copy of the ensemble's values: per row (beamlet, all matrices in order) an FP32 scale =
 * the row maximum, and q = value / scale rounded to nearest even in BF16 or FP16.  After it,
 * fd_dosemat_tofile also writes valuesBuffer.bf16.bin or valuesBuffer.fp16.bin (u16 per entry,
 * entry order of valuesBuffer.bin) and beamletScale.bin (f32 per row); the canonical files stay as
 * readDose.m reads them.  The reference stores FP32 only (IMRTSparseMatEns.cu:225-236). */
int fd_dosemat_pack(fd_dosemat* m, int value_type, void* stream);
/* Host copy of the packed data (rows_total scales, nnz_total u16 values). */
int fd_dosemat_copy_packed(const fd_dosemat* m, float* scale, uint16_t* values);

/* beamletFlagSave (IMRTBeamBundle.cu:572-601): u8 flags, n_beams * per_beam. */
int fd_write_fluence_map(const char* path, const uint8_t* flags, uint64_t n_beams, uint64_t per_beam);

/* ---- D^T X: fluence vectors through the dose-loading matrix ----------------------------------
 * The optimiser multiplies the matrix by fluence vectors (IMRTOptimize.cu:604-625, cuSPARSE on the
 * FP32 CSR; IMRTDebug.cu:211-315).  fd_spmm_create builds a voxel-major copy of an ensemble (all
 * matrices, rows in order) with FP32, BF16 or FP16 values (16-bit: per-row FP32 scale, as
 * fd_dosemat_pack); fd_spmm_run computes Y = D^T X for k right-hand sides with FP32 accumulation:
 * x_dev is rows_total x k (row-major), y_dev is n_cols x k (row-major).  Each column's entries are
 * summed in row order, so results are bitwise reproducible. */
typedef struct fd_spmm fd_spmm;
int fd_spmm_create(const fd_dosemat* m, int value_type, fd_spmm** out, void* stream);
int fd_spmm_info(const fd_spmm* p, uint64_t* rows, uint64_t* cols, uint64_t* nnz);
int fd_spmm_run(fd_spmm* p, const float* x_dev, int k, float* y_dev, void* stream);
int fd_spmm_destroy(fd_spmm* p);

/* ---- Serving: a resident request loop over one context ------------------------------------
 * The reference computes beams one at a time inside DoseMatConstruction (IMRTDoseMat.cu:183-268,
 * one beam per TermaComputeCollective / DoseComputeCollective / BEV2PVCSInterp round with a host
 * synchronisation after each); it has no request loop.  These entry points keep the context's
 * density, spectrum and cone tables resident and take dose requests (one or more beams, any
 * beamlet count) as they arrive.  Slot threads, each with its own stream and scratch, pull beams
 * from the queue in arrival order, pack them into batches of up to max_batch_beamlets, and admit
 * a batch only while the device scratch of all batches in flight stays within hbm_budget, so
 * requests that arrive while others run are admitted as slots and memory free up and their work
 * overlaps on the device.  Each beam's matrix is identical to fd_dosemat_add_beam's for it alone. */
typedef struct fd_server fd_server;

typedef struct fd_serve_opts {
    int n_slots;                 /* concurrent batches (streams); 0: 3 */
    int max_batch_beamlets;      /* beamlets per batch; 0: the context's tuning (fd_autotune), else 4096 */
    uint64_t hbm_budget;         /* bytes of batch scratch in flight; 0: half the free HBM (with the
                                    scratch of stopped servers the context keeps for reuse) */
    float extent;                /* resampler extent [cm]; 0: 2.0 */
    int options;                 /* FD_OPT_EXACT_G / FD_OPT_RPL_FP64(_ALL) for the batches */
    float threshold;             /* relative CSR threshold, as fd_ctx_set_threshold */
} fd_serve_opts;

typedef struct fd_serve_stats {
    uint64_t requests_done, beams_done, beamlets_done, batches;
    double p50_ms, p99_ms, mean_ms, max_ms; /* per-beam latency: request submit -> beam ready */
    double first_submit_s, last_done_s;     /* seconds since fd_serve_start */
    uint64_t peak_inflight_bytes;           /* admitted scratch, high-water mark */
    int peak_inflight_batches;
} fd_serve_stats;

int fd_serve_start(fd_ctx* ctx, const fd_serve_opts* opts, fd_server** out);
/* Copies the request (beamlets of n_beams beams concatenated, counts[k] each, 256 fluence floats
 * per beamlet) and queues it; returns at once with its id. */
int fd_serve_submit(fd_server* s, const fd_beam* beamlets, const float* fluence, const int* counts, int n_beams,
                    uint64_t* request_id);
/* *state: 0 queued or running, 1 finished, -1 failed (fd_serve_fetch reports the error). */
int fd_serve_poll(fd_server* s, uint64_t request_id, int* state);
/* Blocks until the request finishes or timeout_s passes (FD_ERR_CAPACITY on time-out). */
int fd_serve_wait(fd_server* s, uint64_t request_id, double timeout_s);
/* Waits, appends the request's beams to m as matrices (in beam order) and forgets the request. */
int fd_serve_fetch(fd_server* s, uint64_t request_id, fd_dosemat* m);
int fd_serve_stats_get(fd_server* s, fd_serve_stats* out);
int fd_serve_stats_reset(fd_server* s);
/* Finishes the queued work, joins the slot threads and frees the server (not the context).  The
   slots' device scratch and the output buffers stay with the context for the next server started
   on it (freed by fd_ctx_destroy). */
int fd_serve_stop(fd_server* s);

#ifdef __cplusplus
}
#endif

#endif /* FASTDOSE_HIP_H */
