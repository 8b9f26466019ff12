// engine.hip -- stage runners of libfastdose_hip on a workspace (one stream each).
//
// The per-beam pipeline of IMRT::DoseMatConstruction (IMRTOpt/src/IMRTDoseMat.cu:183-268):
// TermaComputeCollective -> DoseComputeCollective -> BEV2PVCSInterp -> addMat, here for any
// batch of beamlets (of one or many beams) with one host synchronisation per batch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <thread>

#include "engine.h"
#include "geometry.h"

namespace fdi {

// Pinned host buffer of at least `need` elements (grows to twice its size at least: freeing pinned
// memory waits for the device).
template <typename T>
int grow_pinned(T*& p, size_t& cap, size_t need) {
    if (need <= cap) return FD_OK;
    const size_t want = std::max(need, 2 * cap);
    if (p) FD_HIP(hipHostFree(p));
    p = nullptr;
    cap = 0;
    FD_HIP(hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocDefault));
    cap = want;
    return FD_OK;
}

int Workspace::init() {
    FD_HIP(hipEventCreateWithFlags(&stage_done, hipEventDisableTiming));
    FD_HIP(hipHostMalloc((void**)&host_totals, 4 * sizeof(uint64_t), hipHostMallocDefault));
    // kernel timers only time: no system-scope fence (cache write-back and invalidation) at each
    // record; results are read after the stream has been synchronised
    for (auto& e : ev) FD_HIP(hipEventCreateWithFlags(&e, hipEventDefault | hipEventDisableSystemFence));
    return FD_OK;
}

void Workspace::release() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    if (own_stream) (void)hipStreamDestroy(own_stream);
    own_stream = nullptr;
    if (aux_stream) (void)hipStreamDestroy(aux_stream);
    aux_stream = nullptr;
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (count_ev) (void)hipEventDestroy(count_ev);
    fork_ev = join_ev = count_ev = nullptr;
    if (stage) (void)hipHostFree(stage);
    if (flu_stage) (void)hipHostFree(flu_stage);
    if (tab_stage) (void)hipHostFree(tab_stage);
    tab_stage = nullptr;
    tab_stage_cap = 0;
    if (host_totals) (void)hipHostFree(host_totals);
    if (host_rowoff) (void)hipHostFree(host_rowoff);
    host_rowoff = nullptr;
    host_rowoff_cap = 0;
    if (stage_done) (void)hipEventDestroy(stage_done);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    stage = nullptr;
    flu_stage = nullptr;
    host_totals = nullptr;
    stage_done = nullptr;
    for (auto& e : ev) e = nullptr;
    desc.release();
    partial.release();
    fluence.release();
    terma.release();
    dens_bev.release();
    dose_bev.release();
    cand_count.release();
    cand_rows.release();
    nnz_count.release();
    cand_off.release();
    nnz_off.release();
    tmp_cols.release();
    tmp_vals.release();
    rowoff.release();
    out_cols.release();
    out_vals.release();
    flags.release();
    rowmax.release();
    scan_tmp.release();
    tab_ent.release();
    tab_rec.release();
    tab_hdr.release();
}

int Workspace::fold(int slot) {
    if (!ev_pending[slot]) return FD_OK;
    // the cone sum's launch starts where the CCCS kernel's ends
    hipEvent_t a = (slot == FD_T_CONE_SUM) ? ev[2 * FD_T_CCCS + 1] : ev[2 * slot];
    hipEvent_t b = ev[2 * slot + 1];
    ev_pending[slot] = false;
    FD_HIP(hipEventSynchronize(b));
    float t = 0.0f;
    FD_HIP(hipEventElapsedTime(&t, a, b));
    acc_ms[slot] += t;
    acc_n[slot] += 1;
    return FD_OK;
}

size_t Workspace::footprint() const {
    return desc.bytes() + partial.bytes() + fluence.bytes() + terma.bytes() + dens_bev.bytes() + dose_bev.bytes() +
           cand_count.bytes() + cand_rows.bytes() + nnz_count.bytes() + cand_off.bytes() + nnz_off.bytes() +
           tmp_cols.bytes() + tmp_vals.bytes() + rowoff.bytes() + out_cols.bytes() + out_vals.bytes() + flags.bytes() +
           rowmax.bytes() + tab_ent.bytes() + tab_rec.bytes() + tab_hdr.bytes();
}

int Workspace::reserve_staging(int n, int n_fwd) {
    FD_TRY(grow_pinned(stage, stage_cap, (size_t)n));
    FD_TRY(grow_pinned(flu_stage, flu_cap, (size_t)n * kNpix));
    FD_TRY(grow_pinned(host_rowoff, host_rowoff_cap, (size_t)n + 1));
    FD_TRY(grow_pinned(tab_stage, tab_stage_cap, (size_t)n * std::max(n_fwd, 1)));
    return FD_OK;
}

int check_beams(const fd_beam* beams, int n) {
    if (n < 0 || (n > 0 && !beams)) {
        set_error("invalid beamlet array");
        return FD_ERR_ARG;
    }
    for (int i = 0; i < n; i++) {
        if (beams[i].fmap_size[0] != (uint32_t)kFmap || beams[i].fmap_size[1] != (uint32_t)kFmap) {
            set_error("the HIP kernels implement 16 x 16 BEV grids (subFluenceDim 16)");
            return FD_ERR_UNSUPPORTED;
        }
        if (beams[i].long_dim == 0 || beams[i].long_dim > (1u << 20)) {
            set_error("beamlet long_dim out of range (run fd_calc_range first)");
            return FD_ERR_ARG;
        }
        if (!(beams[i].sad > 0.0f) || !(beams[i].long_spacing > 0.0f) || !(beams[i].beamlet_size[0] > 0.0f) ||
            !(beams[i].beamlet_size[1] > 0.0f)) {
            set_error("beamlet sad, long_spacing and beamlet_size must be positive");
            return FD_ERR_ARG;
        }
        if (!(beams[i].lim_min >= 0x1p-60f)) {  // as fd_calc_range guarantees; the resampler needs depth > 0
            set_error("beamlet lim_min must be positive (the source lies outside the density bounding box)");
            return FD_ERR_ARG;
        }
    }
    return FD_OK;
}

int max_long_dim(const fd_beam* beams, int n) {
    int m = 0;
    for (int i = 0; i < n; i++) m = std::max(m, (int)beams[i].long_dim);
    return m;
}

namespace {

// Device descriptors of beamlets [i0, i1) with BEV volumes at offs[i] (packed), or `stride`
// floats apart when offs is null.  Beamlets of one beam share the orientation, so the rotation
// matrices are computed once per run of equal angles.
void build_descriptors(const fd_beam* beams, int i0, int i1, uint64_t stride, const uint64_t* offs,
                       BeamletDev* out) {
    for (int i = i0; i < i1; i++) {
        const uint64_t off = offs ? offs[i] : (uint64_t)i * stride;
        if (i > i0 && std::memcmp(beams[i].angles, beams[i - 1].angles, sizeof(beams[i].angles)) == 0) {
            std::memcpy(out[i].Rf, out[i - 1].Rf, sizeof(out[i].Rf));
            std::memcpy(out[i].Ri, out[i - 1].Ri, sizeof(out[i].Ri));
            fill_beamlet_dev_no_rotation(beams[i], off, out[i]);
        } else {
            fill_beamlet_dev(beams[i], off, out[i]);
        }
    }
}

// fn(part) for part = 0 .. parts - 1, parts > 1 of them on threads of their own.
template <class F>
void run_parts(int parts, const F& fn) {
    std::thread th[8];
    parts = std::min(parts, 8);
    for (int p = 1; p < parts; p++) th[p] = std::thread([&fn, p] { fn(p); });
    fn(0);
    for (int p = 1; p < parts; p++) th[p].join();
}

// Pinned staging of the descriptors (beams non-null; BEV offsets as build_descriptors) and
// optionally the fluence of a batch; waits until the previous upload from this workspace has left
// the staging area.
int stage_inputs(Workspace& w, const fd_beam* beams, int n, uint64_t stride, const uint64_t* offs,
                 const float* fluence_host, bool threads = false) {
    FD_HIP(hipEventSynchronize(w.stage_done));
    FD_TRY(grow_pinned(w.stage, w.stage_cap, (size_t)n));
    const size_t nf = fluence_host ? (size_t)n * kNpix : 0;
    FD_TRY(grow_pinned(w.flu_stage, w.flu_cap, nf));
    // descriptors (about 0.15 us each, the rotation matrices) and the fluence copy (one core
    // copies into pinned memory at a few GB/s) split over up to 8 threads for large batches
    const int parts = threads && n >= 2048 ? 8 : 1;
    run_parts(parts, [&](int p) {
        const int i0 = (int)((int64_t)n * p / parts), i1 = (int)((int64_t)n * (p + 1) / parts);
        if (beams) build_descriptors(beams, i0, i1, stride, offs, w.stage);
        if (nf)
            std::memcpy(w.flu_stage + (size_t)i0 * kNpix, fluence_host + (size_t)i0 * kNpix,
                        (size_t)(i1 - i0) * kNpix * sizeof(float));
    });
    return FD_OK;
}

// Voxels a beamlet's widened frustum covers (volume / voxel volume), the scale the resampler's
// candidate staging is sized from.
double frustum_voxels(const fd_beam& b, const fd_volume& vol) {
    const double vx = vol.voxel_size[0], vy = vol.voxel_size[1], vz = vol.voxel_size[2];
    const double margin = std::sqrt(vx * vx + vy * vy + vz * vz) + 0.02;
    const double ls = b.long_spacing, sad = b.sad;
    const double z0 = b.lim_min, z1 = b.lim_min + ls * b.long_dim;
    // integral of (w(z) + 2 margin)^2 dz with w(z) = 16 bs z / sad (square field)
    const double k = kFmap * (double)std::max(b.beamlet_size[0], b.beamlet_size[1]) / sad;
    auto F = [&](double z) {
        const double u = k * z + 2 * margin;
        return u * u * u / (3 * k);
    };
    return std::max(F(z1) - F(z0), 0.0) / (vx * vy * vz);
}

// Follow the observed candidates per frustum voxel, with headroom; shrink slowly.
void update_candidate_estimate(Workspace& w, double ratio) {
    w.cand_per_beamlet = std::max(1.15 * ratio, 0.95 * w.cand_per_beamlet);
    if (w.cand_per_beamlet > 1.5 * 1.15 * ratio) w.cand_per_beamlet = 1.5 * 1.15 * ratio;
}

}  // namespace

size_t batch_footprint(const fd_ctx* ctx, const fd_beam* beams, int n) {
    if (n <= 0) return 0;
    uint64_t bev = 0;  // BEV floats of the batch (packed volumes)
    double cand = 0;
    for (int i = 0; i < n; i++) {
        bev += (uint64_t)beams[i].long_dim * kNpix;
        cand += frustum_voxels(beams[i], ctx->vol);
    }
    const size_t segs = (size_t)n * ctx->vol.dim[2];
    return (size_t)bev * sizeof(float) * (3 + std::max(ctx->n_fwd, 4)) + (size_t)n * kNpix * sizeof(float) +
           (size_t)(2.0 * cand) * (2 * sizeof(uint64_t) + 2 * sizeof(float)) + segs * 32 + (size_t)n * 128;
}

int upload_beams(Workspace& w, const fd_beam* beams, int n, uint64_t stride, hipStream_t stream) {
    if (n == 0) return FD_OK;
    // The stages of one beam (fd_terma, fd_cccs, fd_resample_csr) pass the same beamlets: when the
    // device already holds these exact descriptors, ordered on this stream or already copied, the
    // upload (and the copy engine's hand-off between two kernels) is skipped.
    w.desc_next.resize((size_t)n);
    build_descriptors(beams, 0, n, stride, nullptr, w.desc_next.data());
    const size_t bytes = (size_t)n * sizeof(BeamletDev);
    if (w.desc_valid && w.desc_host.size() == (size_t)n &&
        std::memcmp(w.desc_host.data(), w.desc_next.data(), bytes) == 0 &&
        (w.desc_stream == stream || hipEventQuery(w.stage_done) == hipSuccess))
        return FD_OK;
    w.desc_valid = false;
    FD_TRY(stage_inputs(w, nullptr, n, stride, nullptr, nullptr));
    std::memcpy(w.stage, w.desc_next.data(), bytes);
    FD_TRY(w.desc.reserve(n));
    FD_HIP(hipMemcpyAsync(w.desc.ptr, w.stage, bytes, hipMemcpyHostToDevice, stream));
    FD_HIP(hipEventRecord(w.stage_done, stream));
    w.desc_host.swap(w.desc_next);
    w.desc_valid = true;
    w.desc_stream = stream;
    return FD_OK;
}

int reserve_cccs(const fd_ctx* ctx, Workspace& w, uint64_t bev_floats);

int run_terma(const fd_ctx* ctx, Workspace& w, int options, int n, const float* fluence, float* terma, float* dens,
              uint64_t bev_floats, hipStream_t stream) {
    // FP64 radiological path on rays longer than kLongRayVoxels voxels (or on all rays)
    const float vmin = std::min(ctx->vol.voxel_size[0], std::min(ctx->vol.voxel_size[1], ctx->vol.voxel_size[2]));
    const float rpl64_len = (options & FD_OPT_RPL_FP64_ALL) ? -1.0f
                            : (options & FD_OPT_RPL_FP64) ? kLongRayVoxels * vmin
                                                          : INFINITY;
    // the per-cone partial buffer (4 floats per BEV value at least) is free until the CCCS stage
    FD_TRY(reserve_cccs(ctx, w, bev_floats));
    w.mark(FD_T_TERMA, 0, stream);
    hipError_t e = launch_terma(w.desc.ptr, n, fluence, ctx->density_pad, ctx->vol, ctx->spec_dev, rpl64_len, terma,
                                dens, w.partial.ptr, ctx->tune.terma_segments, stream);
    w.mark(FD_T_TERMA, 1, stream);
    if (e != hipSuccess) {
        set_error(std::string("terma launch: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

// The per-cone partials of a batch whose BEV buffers hold bev_floats floats (n_fwd of them per BEV
// value, and at least 4: the Terma kernel's scratch).
int reserve_cccs(const fd_ctx* ctx, Workspace& w, uint64_t bev_floats) {
    FD_TRY(w.partial.reserve((size_t)bev_floats * std::max(ctx->n_fwd, 4)));
    return FD_OK;
}

int run_cccs_tables(const fd_ctx* ctx, Workspace& w, const CccsTables& tab, hipStream_t stream) {
    w.mark(FD_T_CCCS_TABLES, 0, stream);
    const hipError_t et = launch_cccs_tables(tab, ctx->cones_dev, ctx->n_fwd, ctx->cones.n_phi, stream);
    w.mark(FD_T_CCCS_TABLES, 1, stream);
    if (et != hipSuccess) {
        set_error(std::string("cccs tables launch: ") + hipGetErrorString(et));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

int run_cccs(const fd_ctx* ctx, Workspace& w, int options, int n, int max_L, const float* terma, const float* dens,
             float* dose, uint64_t bev_floats, hipStream_t stream, const CccsTables* tab, bool tab_built,
             bool vec4) {
    FD_TRY(reserve_cccs(ctx, w, bev_floats));
    if (tab && !tab_built) FD_TRY(run_cccs_tables(ctx, w, *tab, stream));
    hipEvent_t evs[3];
    const bool t = w.timing && w.ev[0];
    if (t) {
        FD_TRY(w.fold(FD_T_CCCS));  // before their events are recorded again
        FD_TRY(w.fold(FD_T_CONE_SUM));
        evs[0] = w.ev[2 * FD_T_CCCS];
        evs[1] = w.ev[2 * FD_T_CCCS + 1];
        evs[2] = w.ev[2 * FD_T_CONE_SUM + 1];
    }
    hipError_t e = launch_cccs(w.desc.ptr, n, max_L, ctx->cones_dev, ctx->n_fwd, ctx->cones.n_phi,
                               (options & FD_OPT_EXACT_G) != 0, terma, dens, w.partial.ptr, dose, vec4, stream,
                               t ? evs : nullptr, tab);
    if (t) {
        // the cone sum starts where the CCCS kernel stops (timings pair them)
        w.end_mark(FD_T_CCCS);
        w.end_mark(FD_T_CONE_SUM);
    }
    if (e != hipSuccess) {
        set_error(std::string("cccs launch: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

namespace {

int reserve_segments(const fd_ctx* ctx, Workspace& w, int n) {
    const size_t segs = (size_t)n * ctx->vol.dim[2];
    FD_TRY(w.cand_count.reserve(segs));
    FD_TRY(w.cand_rows.reserve(segs));
    FD_TRY(w.nnz_count.reserve(segs));
    FD_TRY(w.cand_off.reserve(segs + 1));
    FD_TRY(w.nnz_off.reserve(segs + 1));
    FD_TRY(w.flags.reserve(4));  // [0] overflow, [1] zero (read as one u64 with [0]), [2] table overflow
    FD_TRY(w.rowmax.reserve(n));
    FD_TRY(w.scan_tmp.reserve(scan_scratch_size(segs)));
    return FD_OK;
}

int launch_count(const fd_ctx* ctx, Workspace& w, int n, hipStream_t stream) {
    const size_t segs = (size_t)n * ctx->vol.dim[2];
    w.mark(FD_T_RESAMPLE_COUNT, 0, stream);
    hipError_t e = launch_resample_count(w.desc.ptr, n, ctx->vol, w.cand_count.ptr, w.cand_rows.ptr, stream);
    if (e == hipSuccess) e = launch_scan_u32(w.cand_count.ptr, segs, w.cand_off.ptr, w.cand_off.ptr + segs, w.scan_tmp.ptr, stream);
    w.mark(FD_T_RESAMPLE_COUNT, 1, stream);
    if (e != hipSuccess) {
        set_error(std::string("resample count: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

int launch_values(const fd_ctx* ctx, Workspace& w, int n, const float* dose, float extent, float thr, uint64_t cap,
                  hipStream_t stream) {
    const size_t segs = (size_t)n * ctx->vol.dim[2];
    const int Z = (int)ctx->vol.dim[2];
    w.mark(FD_T_RESAMPLE_VALUES, 0, stream);
    hipError_t e = launch_fill_u32(w.flags.ptr, 2, 0u, stream);
    if (e == hipSuccess && thr > 0.0f) e = launch_fill_u32(w.rowmax.ptr, (uint64_t)n, 0u, stream);
    if (e == hipSuccess)
        e = launch_resample_values(w.desc.ptr, n, ctx->vol, extent, dose, w.cand_off.ptr, w.cand_rows.ptr, cap, w.flags.ptr,
                                   w.tmp_cols.ptr, w.tmp_vals.ptr, w.nnz_count.ptr,
                                   thr > 0.0f ? w.rowmax.ptr : nullptr,
                                   ctx->rcp_exact && !(ctx->options & FD_OPT_IEEE_DIV), stream);
    if (thr > 0.0f) {
        // kept entries per segment into cand_count (free once cand_off is scanned), then offsets
        if (e == hipSuccess)
            e = launch_resample_threshold_count(n, Z, w.cand_off.ptr, w.nnz_count.ptr, w.tmp_vals.ptr, w.rowmax.ptr,
                                                thr, w.cand_count.ptr, stream);
        if (e == hipSuccess)
            e = launch_scan_u32(w.cand_count.ptr, segs, w.nnz_off.ptr, w.nnz_off.ptr + segs, w.scan_tmp.ptr, stream);
    } else if (e == hipSuccess) {
        e = launch_scan_u32(w.nnz_count.ptr, segs, w.nnz_off.ptr, w.nnz_off.ptr + segs, w.scan_tmp.ptr, stream);
    }
    w.mark(FD_T_RESAMPLE_VALUES, 1, stream);
    if (e != hipSuccess) {
        set_error(std::string("resample values: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

int reserve_host_rowoff(Workspace& w, int n) { return grow_pinned(w.host_rowoff, w.host_rowoff_cap, (size_t)n + 1); }

// totals of the last resampler pass -> pinned host: [0] candidates, [1] nonzeros, [2] overflow;
// with rows, also the batch's row offsets -> w.host_rowoff (reserved by the caller)
int fetch_totals(const fd_ctx* ctx, Workspace& w, int n, hipStream_t stream, bool rows) {
    const size_t segs = (size_t)n * ctx->vol.dim[2];
    if (rows)
        FD_HIP(hipMemcpyAsync(w.host_rowoff, w.rowoff.ptr, ((size_t)n + 1) * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, stream));
    FD_HIP(hipMemcpyAsync(w.host_totals, w.cand_off.ptr + segs, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    FD_HIP(hipMemcpyAsync(w.host_totals + 1, w.nnz_off.ptr + segs, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    FD_HIP(hipMemcpyAsync(w.host_totals + 2, w.flags.ptr, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    return FD_OK;
}

}  // namespace

int run_resample_csr(const fd_ctx* ctx, Workspace& w, const fd_beam* beams, int n, const float* dose, float extent,
                     float thr, const fd_csr& out, uint64_t* nnz_total, hipStream_t stream) {
    FD_TRY(reserve_segments(ctx, w, n));
    // Staging sized from the running candidate estimate, and the copy into the caller's arrays
    // (skipped on the device when they are too small) issued before the host learns the totals:
    // one synchronisation, a second pass only when the estimated staging overflows.
    double est = 0;
    for (int i = 0; i < n; i++) est += frustum_voxels(beams[i], ctx->vol);
    if (w.cand_per_beamlet <= 0) w.cand_per_beamlet = 2.0;
    uint64_t cap = std::max<uint64_t>((uint64_t)std::ceil(w.cand_per_beamlet * est), 1024);
    const uint64_t capacity = (out.columns && out.values) ? out.capacity : 0;
    auto pass = [&]() -> int {
        FD_TRY(w.tmp_cols.reserve(cap));
        FD_TRY(w.tmp_vals.reserve(cap));
        FD_TRY(launch_values(ctx, w, n, dose, extent, thr, cap, stream));
        FD_TRY(run_resample_compact(ctx, w, n, thr, out.offsets, CsrDst{out.columns, out.values, capacity}, CsrDst{},
                                    stream));
        FD_TRY(fetch_totals(ctx, w, n, stream, false));
        FD_HIP(hipStreamSynchronize(stream));
        return FD_OK;
    };
    FD_TRY(launch_count(ctx, w, n, stream));
    FD_TRY(pass());
    const uint64_t ncand = w.host_totals[0];
    if (w.host_totals[2]) {
        cap = ncand;
        FD_TRY(pass());
        if (w.host_totals[2]) {
            set_error("resampler staging overflow with exact sizing");
            return FD_ERR_HIP;
        }
    }
    update_candidate_estimate(w, est > 0 ? (double)ncand / est : 1.0);
    *nnz_total = w.host_totals[1];
    return FD_OK;
}

int run_resample_compact(const fd_ctx* ctx, Workspace& w, int n, float thr, uint64_t* offsets, const CsrDst& dst0,
                         const CsrDst& dst1, hipStream_t stream) {
    w.mark(FD_T_RESAMPLE_COMPACT, 0, stream);
    hipError_t e = launch_resample_compact(n, (int)ctx->vol.dim[2], w.cand_off.ptr, w.nnz_off.ptr, w.nnz_count.ptr,
                                           w.tmp_cols.ptr, w.tmp_vals.ptr, thr > 0.0f ? w.rowmax.ptr : nullptr, thr,
                                           offsets, dst0, dst1, stream);
    w.mark(FD_T_RESAMPLE_COMPACT, 1, stream);
    if (e != hipSuccess) {
        set_error(std::string("resample compact: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    return FD_OK;
}

namespace {

// Beamlets of a batch per march geometry at which the tables replace the march wavefronts by
// default: below it, walking every geometry's longest beamlet costs about what it saves.
constexpr int kTabMinShare = 4;

// March tables for a batch (FD_OPT_TABLES, or kTabMinShare beamlets per march geometry unless
// FD_OPT_NO_TABLES): assigns w.tab_geom, stages one TabEntry per (geometry, forward cone) in
// w.tab_stage and reserves the streams and headers.  Capacities: within one slice an azimuth
// crosses at most ceil(|dx / dz|) + 1 x and ceil(|dy / dz|) + 1 y cell boundaries plus the slice
// end (dx, dy, dz the lattice direction); twice that plus 4 bounds its steps per slice, a chunk
// takes at most cccs_chunk_steps() of them, and a walk covers at most Lmax slices.  A chunk's
// records never straddle a block of cccs_chunk_steps() records: the space skipped at a block end
// is smaller than the chunk placed after it, so a stream needs at most twice its records (plus a
// block).  Sets w.tab_n_gt (0: no tables).
int plan_tables(const fd_ctx* ctx, Workspace& w, int options, const fd_beam* beams, int n) {
    w.tab_n_gt = 0;
    const int n_phi = ctx->cones.n_phi, n_fwd = ctx->n_fwd;
    if ((options & FD_OPT_TABLES) && n_phi > cccs_table_max_phi()) {
        set_error("FD_OPT_TABLES: the march tables take at most " + std::to_string(cccs_table_max_phi()) +
                  " azimuths");
        return FD_ERR_UNSUPPORTED;
    }
    if ((options & FD_OPT_NO_TABLES) || n_fwd <= 0 || n_phi > cccs_table_max_phi()) return FD_OK;
    struct Geom {
        uint32_t bits[3];
        int Lmax;
    };
    std::vector<Geom> geoms;
    w.tab_geom.resize((size_t)n);
    // open-addressed index of the geometries by their bit patterns
    size_t slots = 64;
    while (slots < 2 * (size_t)n) slots *= 2;
    std::vector<int> slot_geom(slots, -1);
    for (int i = 0; i < n; i++) {
        uint32_t k[3];
        std::memcpy(&k[0], &beams[i].beamlet_size[0], 4);
        std::memcpy(&k[1], &beams[i].beamlet_size[1], 4);
        std::memcpy(&k[2], &beams[i].long_spacing, 4);
        uint64_t hsh = ((uint64_t)k[0] * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)k[1] * 0xC2B2AE3D27D4EB4Full) ^
                       ((uint64_t)k[2] * 0x165667B19E3779F9ull);
        size_t h = (size_t)(hsh ^ (hsh >> 29)) & (slots - 1);
        int g;
        for (;;) {
            g = slot_geom[h];
            if (g < 0 || std::memcmp(geoms[g].bits, k, sizeof(k)) == 0) break;
            h = (h + 1) & (slots - 1);
        }
        if (g < 0) {
            g = (int)geoms.size();
            geoms.push_back(Geom{{k[0], k[1], k[2]}, 0});
            slot_geom[h] = g;
        }
        geoms[g].Lmax = std::max(geoms[g].Lmax, (int)beams[i].long_dim);
        w.tab_geom[i] = g;
    }
    const int ng = (int)geoms.size();
    if (!(options & FD_OPT_TABLES) && n < kTabMinShare * ng) return FD_OK;
    const size_t n_gt = (size_t)ng * n_fwd;
    FD_TRY(grow_pinned(w.tab_stage, w.tab_stage_cap, n_gt));
    const int chunk = cccs_chunk_steps();
    uint64_t rec = 0, hdr = 0;
    for (int g = 0; g < ng; g++) {
        float b0, b1, ls;
        std::memcpy(&b0, &geoms[g].bits[0], 4);
        std::memcpy(&b1, &geoms[g].bits[1], 4);
        std::memcpy(&ls, &geoms[g].bits[2], 4);
        for (int t = 0; t < n_fwd; t++) {
            double steps = 0;
            for (int f = 0; f < n_phi; f++) {
                const double az = std::fabs((double)ctx->cones.uz[t][f]) / ls;
                const double ax = std::fabs((double)ctx->cones.ux[t][f]) / b0;
                const double ay = std::fabs((double)ctx->cones.uy[t][f]) / b1;
                steps = std::max(steps, 4.0 + 2.0 * std::ceil(ax / az) + 2.0 * std::ceil(ay / az));
            }
            const double L1 = (double)geoms[g].Lmax + 1.0;
            const double rec_cap = 2.0 * L1 * steps + chunk, chunk_cap = L1 * std::ceil(steps / chunk) + 2.0;
            if (!(rec_cap < 4e9 && chunk_cap < 4e9)) {
                set_error("CCCS march table beyond 2^32 records (cone too close to 90 degrees)");
                return FD_ERR_CAPACITY;
            }
            TabEntry& e = w.tab_stage[(size_t)g * n_fwd + t];
            e.bs0 = b0;
            e.bs1 = b1;
            e.ls = ls;
            e.Lmax = geoms[g].Lmax;
            e.rec_cap = (uint32_t)rec_cap;
            e.chunk_cap = (uint32_t)chunk_cap;
            e.rec_base = rec;
            e.hdr_base = hdr;
            e.pad = 0;
            rec += 2 * (uint64_t)n_phi * e.rec_cap;
            hdr += (uint64_t)e.chunk_cap * kTabHdr;
        }
    }
    FD_TRY(w.tab_ent.reserve(n_gt));
    FD_TRY(w.tab_rec.reserve(rec));
    FD_TRY(w.tab_hdr.reserve(hdr));
    w.tab_n_gt = (int)n_gt;
    return FD_OK;
}

}  // namespace

int run_batch(const fd_ctx* ctx, Workspace& w, int options, const fd_beam* beams, int n, const float* fluence,
              float extent, float thr, hipStream_t stream, BatchResult* out, const CsrDst* sink, bool overlap) {
    out->rowoff.assign(1, 0);
    out->nnz = 0;
    if (n == 0) return FD_OK;
    const int maxL = max_long_dim(beams, n);
    // packed BEV volumes: beamlet i at bev_offs[i], the sum of L_j * 256 over j < i
    w.bev_offs.resize((size_t)n + 1);
    w.bev_offs[0] = 0;
    double est = 0;
    for (int i = 0; i < n; i++) {
        w.bev_offs[i + 1] = w.bev_offs[i] + (uint64_t)beams[i].long_dim * kNpix;
        est += frustum_voxels(beams[i], ctx->vol);
    }
    const uint64_t bev = w.bev_offs[n];
    if (w.cand_per_beamlet <= 0) w.cand_per_beamlet = 2.0;  // candidates per frustum voxel, first guess
    uint64_t cap = std::max<uint64_t>((uint64_t)std::ceil(w.cand_per_beamlet * est), 1024);

    w.desc_valid = false;  // the batch's descriptors replace the single-stage ones
    FD_TRY(w.desc.reserve(n));
    FD_TRY(w.fluence.reserve((size_t)n * kNpix));
    FD_TRY(w.terma.reserve((size_t)bev));
    FD_TRY(w.dens_bev.reserve((size_t)bev));
    FD_TRY(w.dose_bev.reserve((size_t)bev));
    FD_TRY(w.rowoff.reserve((size_t)n + 1));
    FD_TRY(reserve_segments(ctx, w, n));
    FD_TRY(w.tmp_cols.reserve(cap));
    FD_TRY(w.tmp_vals.reserve(cap));
    FD_TRY(w.out_cols.reserve(cap));
    FD_TRY(w.out_vals.reserve(cap));

    FD_TRY(reserve_cccs(ctx, w, bev));
    FD_TRY(reserve_host_rowoff(w, n));
    FD_TRY(plan_tables(ctx, w, options, beams, n));
    FD_TRY(stage_inputs(w, beams, n, 0, w.bev_offs.data(), fluence, overlap));
    const CccsTables tabs{w.tab_ent.ptr, w.tab_rec.ptr, w.tab_hdr.ptr, w.flags.ptr + 2, w.tab_n_gt};
    if (overlap && !w.aux_stream) {
        FD_HIP(hipStreamCreateWithFlags(&w.aux_stream, hipStreamNonBlocking));
        FD_HIP(hipEventCreateWithFlags(&w.fork_ev, hipEventDisableTiming));
        FD_HIP(hipEventCreateWithFlags(&w.join_ev, hipEventDisableTiming));
        FD_HIP(hipEventCreateWithFlags(&w.count_ev, hipEventDisableTiming));
    }
    // the batch's CSR entries: into the caller's sink when they fit there, else the workspace
    CsrDst dst1{w.out_cols.ptr, w.out_vals.ptr, cap};
    CsrDst dst0 = sink ? *sink : dst1;
    if (w.tab_n_gt)
        for (int i = 0; i < n; i++) w.stage[i].geom = w.tab_geom[i];
    w.host_totals[3] = 0;

    // Everything from the uploads to the host copies of the totals is one stream-ordered sequence
    // with no host synchronisation, so it can run as a captured HIP graph (FD_OPT_GRAPH).
    auto enqueue = [&](hipStream_t st) -> int {
        FD_HIP(hipMemcpyAsync(w.desc.ptr, w.stage, (size_t)n * sizeof(BeamletDev), hipMemcpyHostToDevice, st));
        FD_HIP(hipMemcpyAsync(w.fluence.ptr, w.flu_stage, (size_t)n * kNpix * sizeof(float), hipMemcpyHostToDevice,
                              st));
        if (w.tab_n_gt)
            FD_HIP(hipMemcpyAsync(w.tab_ent.ptr, w.tab_stage, (size_t)w.tab_n_gt * sizeof(TabEntry),
                                  hipMemcpyHostToDevice, st));
        if (overlap) {
            // What depends on the geometries only runs on the second stream while Terma runs
            // (Terma leaves vector issue slots free, and a few hundred single-wavefront table
            // workgroups would leave the GPU idle): the CCCS march tables, then the resampler's
            // candidate count.
            FD_HIP(hipEventRecord(w.fork_ev, st));
            FD_HIP(hipStreamWaitEvent(w.aux_stream, w.fork_ev, 0));
            if (w.tab_n_gt) {
                FD_TRY(run_cccs_tables(ctx, w, tabs, w.aux_stream));
                FD_HIP(hipEventRecord(w.join_ev, w.aux_stream));
            }
            FD_TRY(launch_count(ctx, w, n, w.aux_stream));
            FD_HIP(hipEventRecord(w.count_ev, w.aux_stream));
        }
        FD_TRY(run_terma(ctx, w, options, n, w.fluence.ptr, w.terma.ptr, w.dens_bev.ptr, bev, st));
        if (overlap && w.tab_n_gt) FD_HIP(hipStreamWaitEvent(st, w.join_ev, 0));
        FD_TRY(run_cccs(ctx, w, options, n, maxL, w.terma.ptr, w.dens_bev.ptr, w.dose_bev.ptr, bev, st,
                        w.tab_n_gt ? &tabs : nullptr, overlap));
        if (overlap)
            FD_HIP(hipStreamWaitEvent(st, w.count_ev, 0));
        else
            FD_TRY(launch_count(ctx, w, n, st));
        FD_TRY(launch_values(ctx, w, n, w.dose_bev.ptr, extent, thr, cap, st));
        FD_TRY(run_resample_compact(ctx, w, n, thr, w.rowoff.ptr, dst0, dst1, st));
        FD_TRY(fetch_totals(ctx, w, n, st, true));
        if (w.tab_n_gt)  // the table overflow flag into the low half of host_totals[3]
            FD_HIP(hipMemcpyAsync(w.host_totals + 3, w.flags.ptr + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        return FD_OK;
    };
    if (options & FD_OPT_GRAPH) {
        if (!stream) {
            if (!w.own_stream) FD_HIP(hipStreamCreateWithFlags(&w.own_stream, hipStreamNonBlocking));
            stream = w.own_stream;
        }
        const bool timing = w.timing;
        w.timing = false;  // no event records inside the graph
        FD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const int r = enqueue(stream);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream, &g);
        w.timing = timing;
        if (r != FD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        FD_HIP(ec);
        if (w.gexec) {
            // same node sequence as the last batch: update its parameters in place
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res;
            if (hipGraphExecUpdate(w.gexec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(w.gexec);
                w.gexec = nullptr;
            }
        }
        hipError_t ei = hipSuccess;
        if (!w.gexec) ei = hipGraphInstantiate(&w.gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FD_HIP(ei);
        FD_HIP(hipGraphLaunch(w.gexec, stream));
    } else {
        FD_TRY(enqueue(stream));
    }
    FD_HIP(hipEventRecord(w.stage_done, stream));
    FD_HIP(hipStreamSynchronize(stream));
    if (w.tab_n_gt && w.host_totals[3]) {
        set_error("CCCS march table beyond its capacity");
        return FD_ERR_HIP;
    }
    const uint64_t ncand = w.host_totals[0];
    const double ratio = est > 0 ? (double)ncand / est : 1.0;
    if (w.host_totals[2]) {
        // staging estimate too small: exact size now, re-run the values pass
        cap = ncand;
        FD_TRY(w.tmp_cols.reserve(cap));
        FD_TRY(w.tmp_vals.reserve(cap));
        FD_TRY(w.out_cols.reserve(cap));
        FD_TRY(w.out_vals.reserve(cap));
        FD_TRY(launch_values(ctx, w, n, w.dose_bev.ptr, extent, thr, cap, stream));
        dst1 = CsrDst{w.out_cols.ptr, w.out_vals.ptr, cap};
        if (!sink) dst0 = dst1;
        FD_TRY(run_resample_compact(ctx, w, n, thr, w.rowoff.ptr, dst0, dst1, stream));
        FD_TRY(fetch_totals(ctx, w, n, stream, true));
        FD_HIP(hipStreamSynchronize(stream));
        if (w.host_totals[2]) {
            set_error("resampler staging overflow with exact sizing");
            return FD_ERR_HIP;
        }
    }
    update_candidate_estimate(w, ratio);
    w.fold_all();
    out->nnz = w.host_totals[1];
    out->in_sink = sink && out->nnz <= dst0.capacity && dst0.columns && dst0.values;
    out->rowoff.assign(w.host_rowoff, w.host_rowoff + n + 1);
    return FD_OK;
}

int append_matrix(fd_dosemat* m, uint64_t rows, uint64_t nnz, hipStream_t stream, uint64_t** off_dst,
                  uint64_t** col_dst, float** val_dst) {
    FD_TRY(grow_keep(m->offsets, m->used_off, m->used_off + rows + 1, stream));
    FD_TRY(grow_keep(m->columns, m->used_nnz, m->used_nnz + std::max<uint64_t>(nnz, 1), stream));
    FD_TRY(grow_keep(m->values, m->used_nnz, m->used_nnz + std::max<uint64_t>(nnz, 1), stream));
    m->packed = -1;  // a packed copy no longer covers the ensemble
    *off_dst = m->offsets.ptr + m->used_off;
    *col_dst = m->columns.ptr + m->used_nnz;
    *val_dst = m->values.ptr + m->used_nnz;
    m->rows.push_back(rows);
    m->nnz.push_back(nnz);
    m->off_off.push_back(m->used_off);
    m->nnz_off.push_back(m->used_nnz);
    m->used_off += rows + 1;
    m->used_nnz += nnz;
    return FD_OK;
}

}  // namespace fdi
