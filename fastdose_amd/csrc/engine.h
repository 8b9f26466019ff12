// engine.h -- host-side engine of libfastdose_hip: device buffers, per-stream workspaces, the
// resident context and the stage runners shared by the C-ABI entry points and the serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "fd_internal.h"
#include "kernels.h"

namespace fdi {

#define FD_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ::fdi::set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            return FD_ERR_HIP;                                                              \
        }                                                                                   \
    } while (0)

#define FD_TRY(call)                \
    do {                            \
        int r_ = (call);            \
        if (r_ != FD_OK) return r_; \
    } while (0)

// Growable device allocation (contents are not kept when it grows).  With `stream` set, the
// allocations and frees are ordered on that stream (hipMallocAsync / hipFreeAsync) instead of
// synchronising the device: a server's slots grow their scratch while other slots' batches run,
// and a device-wide free there waits for those batches (and has stalled behind them for seconds).
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t cap = 0;  // elements
    hipStream_t stream = nullptr;
    int reserve(size_t n) {
        if (n <= cap) return FD_OK;
        size_t want = std::max(n, cap + cap / 2);
        if (ptr) {
            hipError_t e = stream ? hipFreeAsync(ptr, stream) : hipFree(ptr);
            ptr = nullptr;
            cap = 0;
            if (e != hipSuccess) {
                set_error(std::string("hipFree: ") + hipGetErrorString(e));
                return FD_ERR_HIP;
            }
        }
        hipError_t e = stream ? hipMallocAsync((void**)&ptr, want * sizeof(T), stream)
                              : hipMalloc((void**)&ptr, want * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return FD_ERR_HIP;
        }
        cap = want;
        return FD_OK;
    }
    void release() {
        if (ptr) (void)(stream ? hipFreeAsync(ptr, stream) : hipFree(ptr));
        ptr = nullptr;
        cap = 0;
    }
    size_t bytes() const { return cap * sizeof(T); }
    // exchange the storage (not the stream) with another buffer ordered on the same device
    void swap_storage(DevBuf& o) {
        std::swap(ptr, o.ptr);
        std::swap(cap, o.cap);
    }
};

// Growable device buffer that keeps its first `used` elements (used by the dose-matrix ensemble;
// doubles like MatCSREnsemble::addMat, IMRTSparseMatEns.cu:125-149).
template <typename T>
int grow_keep(DevBuf<T>& b, size_t used, size_t need, hipStream_t stream) {
    if (need <= b.cap) return FD_OK;
    size_t want = std::max<size_t>(need, b.cap ? b.cap * 2 : 1024);
    T* np = nullptr;
    FD_HIP(hipMalloc((void**)&np, want * sizeof(T)));
    if (used && b.ptr) FD_HIP(hipMemcpyAsync(np, b.ptr, used * sizeof(T), hipMemcpyDeviceToDevice, stream));
    FD_HIP(hipStreamSynchronize(stream));
    if (b.ptr) FD_HIP(hipFree(b.ptr));
    b.ptr = np;
    b.cap = want;
    return FD_OK;
}

// Scratch of one in-flight batch: descriptors, BEV volumes, cone partials, resampler staging and
// the batch's CSR rows.  One per stream; a context owns one, a server one per slot.
struct Workspace {
    BeamletDev* stage = nullptr;  // pinned descriptor staging
    size_t stage_cap = 0;
    float* flu_stage = nullptr;   // pinned fluence staging (batched path)
    size_t flu_cap = 0;
    hipEvent_t stage_done = nullptr;
    DevBuf<BeamletDev> desc;
    // what desc holds after the last upload_beams (valid until a batch overwrites it), and the
    // stream that upload was ordered on
    std::vector<BeamletDev> desc_host, desc_next;
    bool desc_valid = false;
    hipStream_t desc_stream = nullptr;
    DevBuf<float> partial;
    DevBuf<float> fluence, terma, dens_bev, dose_bev;
    std::vector<uint64_t> bev_offs;  // BEV offsets of the last batch's beamlets, packed (n + 1)
    DevBuf<uint32_t> cand_count, nnz_count;
    DevBuf<uint2> cand_rows;      // per segment: rows [y_lo, y_hi) holding candidates
    DevBuf<uint64_t> cand_off, nnz_off;
    DevBuf<uint64_t> tmp_cols;
    DevBuf<float> tmp_vals;
    DevBuf<uint64_t> rowoff;      // per-beamlet CSR row offsets of the last batch (n + 1)
    DevBuf<uint64_t> out_cols;    // CSR of the last batch
    DevBuf<float> out_vals;
    DevBuf<uint32_t> flags;       // [0]: resampler staging overflow
    DevBuf<uint32_t> rowmax;      // per-beamlet maximum (float bits) for the relative threshold
    DevBuf<uint64_t> scan_tmp;    // tile sums of the multi-workgroup scans
    // CCCS march tables of the last batch that used them: entries per (geometry, cone), record
    // streams, chunk headers, the pinned staging of the entries and each beamlet's geometry
    DevBuf<TabEntry> tab_ent;
    DevBuf<float4> tab_rec;
    DevBuf<int> tab_hdr;
    TabEntry* tab_stage = nullptr;
    size_t tab_stage_cap = 0;
    std::vector<int32_t> tab_geom;
    int tab_n_gt = 0;
    uint64_t* host_totals = nullptr;  // pinned: candidates, nonzeros, overflow flags (resampler, tables)
    uint64_t* host_rowoff = nullptr;  // pinned copy of rowoff
    size_t host_rowoff_cap = 0;
    double cand_per_beamlet = 0.0;    // running estimate that sizes the staging up front
    // kernel timers: slot k uses ev[2k] / ev[2k+1] (the cone sum starts at the CCCS end event);
    // per slot, the milliseconds and launches summed over the last API call that launched it
    hipEvent_t ev[2 * FD_T_COUNT] = {};
    bool ev_used[FD_T_COUNT] = {};
    bool ev_pending[FD_T_COUNT] = {};  // recorded, not yet added to acc_ms
    double acc_ms[FD_T_COUNT] = {};
    int acc_n[FD_T_COUNT] = {};
    uint64_t acc_call[FD_T_COUNT] = {};
    uint64_t call = 0;                 // API calls that launched work on this workspace
    bool timing = false;
    hipGraphExec_t gexec = nullptr;   // last instantiated batch graph (FD_OPT_GRAPH)
    hipStream_t own_stream = nullptr; // capture stream when the caller passes the null stream
    // second stream of a batch: the CCCS march tables and the resampler's candidate count run
    // there while Terma runs (fork, tables done, count done)
    hipStream_t aux_stream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, count_ev = nullptr;

    int init();
    void release();
    // Every device buffer of the workspace allocates and frees in the order of stream s (a server
    // slot's stream); nullptr: synchronous allocation.
    void set_alloc_stream(hipStream_t s) {
        desc.stream = partial.stream = fluence.stream = terma.stream = dens_bev.stream = dose_bev.stream = s;
        cand_count.stream = nnz_count.stream = flags.stream = rowmax.stream = s;
        cand_rows.stream = s;
        cand_off.stream = nnz_off.stream = tmp_cols.stream = rowoff.stream = out_cols.stream = scan_tmp.stream = s;
        tmp_vals.stream = out_vals.stream = s;
        tab_ent.stream = s;
        tab_rec.stream = s;
        tab_hdr.stream = s;
    }
    // Pinned host staging for batches of up to n beamlets (n_fwd forward cones), so that batches
    // up to that size never free and reallocate it (freeing pinned memory waits for the device).
    int reserve_staging(int n, int n_fwd);
    void mark(int slot, int which, hipStream_t s) {
        if (!timing || !ev[0]) return;
        if (which == 0) fold(slot);  // the previous launch's pair is about to be re-recorded
        (void)hipEventRecord(ev[2 * slot + which], s);
        if (which == 1) end_mark(slot);
    }
    // The end of slot's launch was recorded (by mark or by a launcher that records its own events).
    void end_mark(int slot) {
        ev_used[slot] = true;
        ev_pending[slot] = true;
        if (acc_call[slot] != call) {
            acc_call[slot] = call;
            acc_ms[slot] = 0.0;
            acc_n[slot] = 0;
        }
    }
    // Adds the recorded launch of slot (if any) to its sums; waits for its end event.
    int fold(int slot);
    void fold_all() {
        for (int k = 0; k < FD_T_COUNT; k++) (void)fold(k);
    }
    // Device bytes this workspace holds.
    size_t footprint() const;
};

// Device CSR entry buffers (column indices and values) of a server's output pool.
struct CsrBufs {
    DevBuf<uint64_t> cols;
    DevBuf<float> vals;
};

// CSR rows of one finished batch, as the host sees them.
struct BatchResult {
    std::vector<uint64_t> rowoff;  // n + 1 offsets into the entries' arrays
    uint64_t nnz = 0;
    bool in_sink = false;          // the entries went to the caller's sink (else ws.out_cols / ws.out_vals)
};

}  // namespace fdi

// Resident state of one device: density, spectrum and cone tables, plus the workspace the
// synchronous entry points use.
struct fd_ctx {
    int device = 0;
    fd_volume vol{};
    float* density = nullptr;
    float* density_pad = nullptr;  // per-voxel trilinear corner records of the density (Terma)
    fd_spectrum spectrum{};
    fd_kernel kernel{};
    fdi::SpectrumDev spec_dev{};
    float* spec_tab = nullptr;  // device spectrum table (spec_dev.tab), (f, df) pairs
    fdi::ConeTable cones{};
    fdi::ConeTable* cones_dev = nullptr;
    int n_fwd = 0;
    int options = 0;
    float threshold = 0.0f;  // relative CSR threshold (0: drop exact zeros only)
    int n_cu = 256;          // compute units of the device
    bool rcp_exact = false;  // the resampler's reciprocal is the IEEE quotient on this device
    fd_tuning tune{8192, 0, 16, 0.0f};  // launch parameters (fd_autotune; Terma segments also plain)
    fdi::Workspace ws;
    fdi::DevBuf<fdi::FlagBeamDev> flag_desc;  // beamlet-flag kernel inputs / outputs
    fdi::DevBuf<uint8_t> flag_out;
    // Device scratch of stopped servers, kept for the next server on this context: the slots'
    // workspaces and the output pool's CSR buffers keep their sizes, so the next server's slots
    // start without allocating (an allocation while another slot's batch runs has stalled both for
    // seconds).  While parked their buffers are ordered on park_stream; freed with the context.
    std::mutex park_mu;
    hipStream_t park_stream = nullptr;
    std::vector<fdi::Workspace> parked_ws;
    std::vector<fdi::CsrBufs> parked_csr;
};

struct fd_dosemat {
    uint64_t n_cols = 0;
    std::vector<uint64_t> rows;       // numRowsPerMat
    std::vector<uint64_t> nnz;        // NonZeroElements
    std::vector<uint64_t> nnz_off;    // start of each matrix in columns/values
    std::vector<uint64_t> off_off;    // OffsetBufferIdx
    fdi::DevBuf<uint64_t> offsets;    // offsetsBuffer
    fdi::DevBuf<uint64_t> columns;    // columnsBuffer
    fdi::DevBuf<float> values;        // valuesBuffer
    uint64_t used_off = 0, used_nnz = 0;
    int packed = -1;                  // FD_VAL_BF16 / FD_VAL_FP16 once packed, -1 otherwise
    fdi::DevBuf<uint16_t> values16;   // packed values (entry order of values)
    fdi::DevBuf<float> row_scale;     // per-row scale of the packed values
    uint64_t total_rows() const {
        uint64_t r = 0;
        for (uint64_t x : rows) r += x;
        return r;
    }
};

namespace fdi {

int check_beams(const fd_beam* beams, int n);
int max_long_dim(const fd_beam* beams, int n);

// Single stages on a workspace (descriptors uploaded by upload_beams first: BEV volumes `stride`
// floats apart).  bev_floats: the floats each of the batch's BEV buffers spans (n x stride for the
// single stages, the packed total in run_batch), which sizes the per-cone partials.  vec4: every
// beamlet's BEV offset is a multiple of four floats (the packed layout; a stride of 4k floats).
int upload_beams(Workspace& w, const fd_beam* beams, int n, uint64_t stride, hipStream_t stream);
int run_terma(const fd_ctx* ctx, Workspace& w, int options, int n, const float* fluence, float* terma, float* dens,
              uint64_t bev_floats, hipStream_t stream);
// The march tables `tab` (timing slot FD_T_CCCS_TABLES) on `stream`.
int run_cccs_tables(const fd_ctx* ctx, Workspace& w, const CccsTables& tab, hipStream_t stream);
int run_cccs(const fd_ctx* ctx, Workspace& w, int options, int n, int max_L, const float* terma, const float* dens,
             float* dose, uint64_t bev_floats, hipStream_t stream, const CccsTables* tab = nullptr,
             bool tab_built = false, bool vec4 = true);
// The whole resampler into the caller's CSR arrays `out` (offsets always; columns and values
// only when nnz <= out.capacity), with one host synchronisation (a second only when the estimated
// staging overflows).  thr > 0 keeps only entries >= thr * (the beamlet's maximum value).
int run_resample_csr(const fd_ctx* ctx, Workspace& w, const fd_beam* beams, int n, const float* dose, float extent,
                     float thr, const fd_csr& out, uint64_t* nnz_total, hipStream_t stream);
int run_resample_compact(const fd_ctx* ctx, Workspace& w, int n, float thr, uint64_t* offsets, const CsrDst& dst0,
                         const CsrDst& dst1, hipStream_t stream);

// The whole path for a batch of beamlets (host descriptors and host fluence, 256 floats each):
// Terma -> CCCS -> resampling, one host synchronisation (two when the resampler staging estimate
// was too small).  The CSR columns and values go to *sink when its capacity holds them (the
// device decides; out->in_sink tells), else to w.out_cols / w.out_vals.
// overlap: the march tables and the candidate count run on a second stream (w.aux_stream) joined
// by events, and the batch's staging runs on several host threads.  Only for callers that submit
// all their device work from one thread: cross-stream waits submitted from several threads at
// once (a server's slots) can stall behind each other when streams share hardware queues.
int run_batch(const fd_ctx* ctx, Workspace& w, int options, const fd_beam* beams, int n, const float* fluence,
              float extent, float thr, hipStream_t stream, BatchResult* out, const CsrDst* sink = nullptr,
              bool overlap = false);

// Device bytes a batch of these beamlets needs (BEV volumes, cone partials, resampler staging).
size_t batch_footprint(const fd_ctx* ctx, const fd_beam* beams, int n);

// Appends one matrix of `rows` rows and `nnz` nonzeros; returns where its arrays go.
int append_matrix(fd_dosemat* m, uint64_t rows, uint64_t nnz, hipStream_t stream, uint64_t** off_dst,
                  uint64_t** col_dst, float** val_dst);

}  // namespace fdi
