// serve.hip -- resident request loop with continuous batching over one fd_ctx.
//
// The reference has no serving path: DoseMatConstruction (IMRTOpt/src/IMRTDoseMat.cu:183-268)
// computes one beam at a time and synchronises after each.  Here requests (one or more beams of
// any beamlet count) are queued beam by beam in arrival order.  Each slot thread owns a stream
// and a workspace; when free it takes the next beams (whole beams, up to max_batch_beamlets),
// admits the batch if the scratch of everything in flight stays within the HBM budget (a batch
// is always admitted when nothing else is in flight), runs the whole path with run_batch, and
// hands the batch's CSR rows to its beams as one pooled block (the workspace's output buffers are
// swapped into it, nothing is copied or allocated); fd_serve_fetch copies a request's rows into the
// caller's ensemble on the server's fetch stream and returns the block to the pool once every beam
// of its batch has been fetched.  Beams that arrive while
// others run join the next batch of whichever slot frees first, so their work overlaps on the
// device with the batches already running.  A stopped server parks its slots' workspaces and its
// output pool on the context, and the next server there starts with them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"

namespace fdi {
namespace {

using Clock = std::chrono::steady_clock;

// The CSR rows of one finished batch.  The slot hands its workspace's output buffers over to the
// block (a swap, no copy) and takes the block's previous buffers in exchange; the block goes back to
// the server's pool when every beam of the batch has been fetched.
struct Block {
    DevBuf<uint64_t> cols;
    DevBuf<float> vals;
    std::vector<uint64_t> rowoff;  // n + 1 row offsets into cols / vals (host)
    int beams_left = 0;            // beams of the batch not fetched yet
};

struct BeamOut {
    Block* block = nullptr;
    uint64_t first_row = 0, rows = 0;
    bool done = false;
};

struct Request {
    uint64_t id = 0;
    std::vector<fd_beam> beamlets;
    std::vector<float> fluence;
    std::vector<int> counts;
    std::vector<uint64_t> first;  // first beamlet of each beam
    std::vector<BeamOut> out;
    Clock::time_point t_submit;
    int beams_left = 0;
    bool failed = false;
    std::string error;
};

struct BeamRef {
    std::shared_ptr<Request> req;
    int beam;
};

}  // namespace
}  // namespace fdi

using namespace fdi;

struct fd_server {
    fd_ctx* ctx = nullptr;
    fd_serve_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<BeamRef> queue;
    std::map<uint64_t, std::shared_ptr<Request>> requests;
    uint64_t next_id = 1;
    bool stopping = false;
    uint64_t inflight_bytes = 0;
    int inflight_batches = 0;
    std::vector<std::thread> threads;
    std::vector<Workspace> ws;
    std::vector<hipStream_t> streams;
    hipStream_t fetch_stream = nullptr;        // copies of fetched beams into callers' ensembles
    std::vector<std::unique_ptr<Block>> blocks;  // every block, for release at stop
    std::vector<Block*> free_blocks;
    Clock::time_point t0;
    // statistics
    std::vector<double> lat_ms;
    uint64_t requests_done = 0, beams_done = 0, beamlets_done = 0, batches = 0, peak_bytes = 0;
    int peak_batches = 0;
    double first_submit = -1, last_done = -1;
};

namespace {

double since(const fd_server* s, Clock::time_point t) {
    return std::chrono::duration<double>(t - s->t0).count();
}

// A block whose beams have all been fetched goes back to the pool (server lock held).
void release_beam(fd_server* s, BeamOut& o) {
    if (!o.block) return;
    if (--o.block->beams_left == 0) s->free_blocks.push_back(o.block);
    o.block = nullptr;
}

// A block from the pool, or a new one (server lock held).  The slot swaps its output buffers with
// the block's, so the largest free block is taken: the buffers the slot gets back then rarely have
// to grow for its next batch (growing frees them, and a hipFree waits for the whole device).
Block* take_block(fd_server* s) {
    if (!s->free_blocks.empty()) {
        size_t best = 0;
        for (size_t i = 1; i < s->free_blocks.size(); i++)
            if (s->free_blocks[i]->cols.cap > s->free_blocks[best]->cols.cap) best = i;
        Block* b = s->free_blocks[best];
        s->free_blocks[best] = s->free_blocks.back();
        s->free_blocks.pop_back();
        return b;
    }
    s->blocks.push_back(std::make_unique<Block>());
    Block* b = s->blocks.back().get();
    b->cols.stream = b->vals.stream = s->fetch_stream;  // a block's storage is freed on the fetch stream
    return b;
}

void slot_loop(fd_server* s, int slot) {
    (void)hipSetDevice(s->ctx->device);
    Workspace& w = s->ws[slot];
    hipStream_t st = s->streams[slot];
    std::vector<fd_beam> beams;
    std::vector<float> flu;
    BatchResult res;
    for (;;) {
        std::vector<BeamRef> batch;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->queue.empty()) {
                    if (s->stopping) return;
                    s->cv_work.wait(lk);
                    continue;
                }
                // pack whole beams in arrival order
                batch.clear();
                beams.clear();
                flu.clear();
                uint64_t n = 0;
                size_t qi = 0;
                while (qi < s->queue.size()) {
                    const BeamRef& r = s->queue[qi];
                    const uint64_t c = (uint64_t)r.req->counts[r.beam];
                    if (!batch.empty() && n + c > (uint64_t)s->opts.max_batch_beamlets) break;
                    batch.push_back(r);
                    n += c;
                    qi++;
                }
                for (const BeamRef& r : batch) {
                    const uint64_t f = r.req->first[r.beam], c = (uint64_t)r.req->counts[r.beam];
                    beams.insert(beams.end(), r.req->beamlets.begin() + f, r.req->beamlets.begin() + f + c);
                    flu.insert(flu.end(), r.req->fluence.begin() + f * kNpix, r.req->fluence.begin() + (f + c) * kNpix);
                }
                bytes = batch_footprint(s->ctx, beams.data(), (int)beams.size());
                if (s->inflight_batches == 0 || s->inflight_bytes + bytes <= s->opts.hbm_budget) {
                    for (size_t k = 0; k < batch.size(); k++) s->queue.pop_front();
                    s->inflight_bytes += bytes;
                    s->inflight_batches++;
                    s->peak_bytes = std::max(s->peak_bytes, s->inflight_bytes);
                    s->peak_batches = std::max(s->peak_batches, s->inflight_batches);
                    break;
                }
                s->cv_work.wait(lk);  // memory frees up when a batch finishes
            }
        }
        int rc = run_batch(s->ctx, w, s->opts.options, beams.data(), (int)beams.size(), flu.data(), s->opts.extent,
                           s->opts.threshold, st, &res);
        // run_batch returns with the batch finished on the device
        const std::string err = rc == FD_OK ? std::string() : std::string(fd_last_error());
        const Clock::time_point t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            Block* blk = nullptr;
            if (rc == FD_OK) {
                blk = take_block(s);
                blk->cols.swap_storage(w.out_cols);
                blk->vals.swap_storage(w.out_vals);
                blk->rowoff.swap(res.rowoff);
                blk->beams_left = (int)batch.size();
            }
            s->inflight_bytes -= bytes;
            s->inflight_batches--;
            s->batches++;
            uint64_t row = 0;
            for (size_t k = 0; k < batch.size(); k++) {
                Request& q = *batch[k].req;
                BeamOut& o = q.out[batch[k].beam];
                const uint64_t c = (uint64_t)q.counts[batch[k].beam];
                if (rc == FD_OK) {
                    o.block = blk;
                    o.first_row = row;
                    o.rows = c;
                } else {
                    q.failed = true;
                    q.error = err;
                }
                row += c;
                o.done = true;
                q.beams_left--;
                const double ms = std::chrono::duration<double, std::milli>(t_done - q.t_submit).count();
                s->lat_ms.push_back(ms);
                s->beams_done++;
                s->beamlets_done += c;
                if (q.beams_left == 0) s->requests_done++;
            }
            s->last_done = since(s, t_done);
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int fd_serve_start(fd_ctx* ctx, const fd_serve_opts* opts, fd_server** out) {
    if (!ctx || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    *out = nullptr;
    fd_serve_opts o{};
    if (opts) o = *opts;
    if (o.n_slots < 0 || o.n_slots > 16 || o.max_batch_beamlets < 0 || o.extent < 0.0f ||
        !(o.threshold >= 0.0f && o.threshold < 1.0f)) {
        set_error("n_slots must be in [0, 16], max_batch_beamlets and extent non-negative, threshold in [0, 1)");
        return FD_ERR_ARG;
    }
    if (o.options & ~(FD_OPT_EXACT_G | FD_OPT_RPL_FP64 | FD_OPT_RPL_FP64_ALL | FD_OPT_GRAPH | FD_OPT_TABLES |
                      FD_OPT_NO_TABLES) ||
        ((o.options & FD_OPT_TABLES) && (o.options & FD_OPT_NO_TABLES))) {
        set_error("serving options: FD_OPT_EXACT_G, FD_OPT_RPL_FP64(_ALL), FD_OPT_GRAPH and one of "
                  "FD_OPT_TABLES / FD_OPT_NO_TABLES only");
        return FD_ERR_ARG;
    }
    if (o.n_slots == 0) o.n_slots = 3;
    if (o.max_batch_beamlets == 0)  // the context's tuned batch size, if it was tuned
        o.max_batch_beamlets = ctx->tune.beamlets_per_s > 0.0f ? ctx->tune.batch_beamlets : 4096;
    if (ctx->tune.graph) o.options |= FD_OPT_GRAPH;
    if (o.extent == 0.0f) o.extent = 2.0f;
    FD_HIP(hipSetDevice(ctx->device));
    if (o.hbm_budget == 0) {
        // half the free HBM, counting the scratch parked on the context (this server reuses it)
        size_t free_b = 0, total_b = 0;
        FD_HIP(hipMemGetInfo(&free_b, &total_b));
        size_t parked_b = 0;
        {
            std::lock_guard<std::mutex> lk(ctx->park_mu);
            for (const Workspace& w : ctx->parked_ws) parked_b += w.footprint();
            for (const CsrBufs& c : ctx->parked_csr) parked_b += c.cols.bytes() + c.vals.bytes();
        }
        o.hbm_budget = (free_b + parked_b) / 2;
    }
    fd_server* s = new fd_server();
    s->ctx = ctx;
    s->opts = o;
    s->t0 = Clock::now();
    s->ws.resize(o.n_slots);
    s->streams.assign(o.n_slots, nullptr);
    {
        // slot scratch is allocated and freed in stream order (DevBuf); the device's pool keeps
        // what the slots free for reuse (by later batches and later servers) instead of returning
        // it to the system: allocating from the system while another slot's batch runs has stalled
        // both for seconds
        hipMemPool_t pool = nullptr;
        uint64_t keep = UINT64_MAX;
        if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess && pool)
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    for (int i = 0; i < o.n_slots; i++) {
        // a workspace parked by an earlier server on this context (its buffers keep their sizes)
        bool parked = false;
        {
            std::lock_guard<std::mutex> lk(ctx->park_mu);
            if (!ctx->parked_ws.empty()) {
                s->ws[i] = std::move(ctx->parked_ws.back());
                ctx->parked_ws.pop_back();
                parked = true;
            }
        }
        int r = parked ? FD_OK : s->ws[i].init();
        hipError_t e = r == FD_OK ? hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) : hipSuccess;
        if (r == FD_OK && e == hipSuccess) {
            s->ws[i].set_alloc_stream(s->streams[i]);
            r = s->ws[i].reserve_staging((int)o.max_batch_beamlets, ctx->n_fwd);
        }
        if (r != FD_OK || e != hipSuccess) {
            if (e != hipSuccess) set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            for (int j = 0; j <= i; j++) {
                s->ws[j].release();
                if (s->streams[j]) {
                    (void)hipStreamSynchronize(s->streams[j]);
                    (void)hipStreamDestroy(s->streams[j]);
                }
            }
            delete s;
            return r != FD_OK ? r : FD_ERR_HIP;
        }
    }
    {
        hipError_t e = hipStreamCreateWithFlags(&s->fetch_stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            for (int j = 0; j < o.n_slots; j++) {
                s->ws[j].release();
                (void)hipStreamSynchronize(s->streams[j]);
                (void)hipStreamDestroy(s->streams[j]);
            }
            delete s;
            return FD_ERR_HIP;
        }
    }
    {
        // the output pool of an earlier server on this context, as free blocks
        std::lock_guard<std::mutex> lk(ctx->park_mu);
        for (CsrBufs& c : ctx->parked_csr) {
            s->blocks.push_back(std::make_unique<Block>());
            Block* b = s->blocks.back().get();
            std::swap(b->cols, c.cols);
            std::swap(b->vals, c.vals);
            b->cols.stream = b->vals.stream = s->fetch_stream;
            s->free_blocks.push_back(b);
        }
        ctx->parked_csr.clear();
    }
    for (int i = 0; i < o.n_slots; i++) s->threads.emplace_back(slot_loop, s, i);
    *out = s;
    return FD_OK;
}

int fd_serve_submit(fd_server* s, const fd_beam* beamlets, const float* fluence, const int* counts, int n_beams,
                    uint64_t* request_id) {
    if (!s || !request_id || n_beams <= 0 || !counts) {
        set_error("null argument or no beams");
        return FD_ERR_ARG;
    }
    auto q = std::make_shared<Request>();
    q->counts.assign(counts, counts + n_beams);
    q->first.assign((size_t)n_beams + 1, 0);
    for (int k = 0; k < n_beams; k++) {
        if (counts[k] < 0) {
            set_error("negative beamlet count");
            return FD_ERR_ARG;
        }
        q->first[k + 1] = q->first[k] + (uint64_t)counts[k];
    }
    const uint64_t total = q->first[n_beams];
    if (total > (uint64_t)INT32_MAX) {
        set_error("too many beamlets");
        return FD_ERR_ARG;
    }
    FD_TRY(check_beams(beamlets, (int)total));
    if (total && !fluence) {
        set_error("null fluence");
        return FD_ERR_ARG;
    }
    q->beamlets.assign(beamlets, beamlets + total);
    q->fluence.assign(fluence, fluence + total * kNpix);
    q->out.resize(n_beams);
    q->beams_left = n_beams;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) {
            set_error("server is stopping");
            return FD_ERR_ARG;
        }
        q->id = s->next_id++;
        q->t_submit = Clock::now();
        if (s->first_submit < 0) s->first_submit = since(s, q->t_submit);
        s->requests[q->id] = q;
        for (int k = 0; k < n_beams; k++) s->queue.push_back(BeamRef{q, k});
        *request_id = q->id;
    }
    s->cv_work.notify_all();
    return FD_OK;
}

int fd_serve_poll(fd_server* s, uint64_t id, int* state) {
    if (!s || !state) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        set_error("unknown request id");
        return FD_ERR_ARG;
    }
    const Request& q = *it->second;
    *state = q.beams_left > 0 ? 0 : (q.failed ? -1 : 1);
    return FD_OK;
}

int fd_serve_wait(fd_server* s, uint64_t id, double timeout_s) {
    if (!s) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) {
        set_error("unknown request id");
        return FD_ERR_ARG;
    }
    std::shared_ptr<Request> q = it->second;
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                             std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9));
    while (q->beams_left > 0)
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && q->beams_left > 0) {
            set_error("request not finished within the time-out");
            return FD_ERR_CAPACITY;
        }
    return FD_OK;
}

int fd_serve_fetch(fd_server* s, uint64_t id, fd_dosemat* m) {
    if (!s || !m) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    FD_TRY(fd_serve_wait(s, id, 0));
    std::shared_ptr<Request> q;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->requests.find(id);
        if (it == s->requests.end()) {
            set_error("unknown request id");
            return FD_ERR_ARG;
        }
        q = it->second;
        s->requests.erase(it);
    }
    FD_HIP(hipSetDevice(s->ctx->device));
    int rc = FD_OK;
    if (q->failed) {
        set_error("request failed: " + q->error);
        rc = FD_ERR_HIP;
    }
    if (rc == FD_OK && q->beamlets.size() && m->n_cols != (uint64_t)s->ctx->vol.dim[0] * s->ctx->vol.dim[1] *
                                                             s->ctx->vol.dim[2]) {
        set_error("dose matrix column count differs from the context's voxel count");
        rc = FD_ERR_ARG;
    }
    // every beam's offsets (re-based to its first row) in one host array that outlives the copies
    std::vector<uint64_t> host_off;
    hipStream_t fs = s->fetch_stream;
    if (rc == FD_OK) {
        size_t total_off = 0;
        uint64_t total_nnz = 0;
        for (const BeamOut& o : q->out) {
            total_off += o.rows + 1;
            total_nnz += o.block->rowoff[o.first_row + o.rows] - o.block->rowoff[o.first_row];
        }
        host_off.reserve(total_off);
        // the ensemble grows once for the whole request: growing it beam by beam frees the old
        // buffers each time, and a hipFree waits for the whole device, slot streams included
        rc = grow_keep(m->offsets, m->used_off, m->used_off + total_off, fs);
        if (rc == FD_OK) rc = grow_keep(m->columns, m->used_nnz, m->used_nnz + std::max<uint64_t>(total_nnz, 1), fs);
        if (rc == FD_OK) rc = grow_keep(m->values, m->used_nnz, m->used_nnz + std::max<uint64_t>(total_nnz, 1), fs);
    }
    for (size_t k = 0; k < q->out.size() && rc == FD_OK; k++) {
        const BeamOut& o = q->out[k];
        const std::vector<uint64_t>& ro = o.block->rowoff;
        const uint64_t base = ro[o.first_row], nnz = ro[o.first_row + o.rows] - base;
        uint64_t *od, *cd;
        float* vd;
        rc = append_matrix(m, o.rows, nnz, fs, &od, &cd, &vd);
        if (rc != FD_OK) break;
        const size_t h = host_off.size();
        for (uint64_t j = 0; j <= o.rows; j++) host_off.push_back(ro[o.first_row + j] - base);
        hipError_t e = hipMemcpyAsync(od, host_off.data() + h, (o.rows + 1) * sizeof(uint64_t),
                                      hipMemcpyHostToDevice, fs);
        if (e == hipSuccess && nnz)
            e = hipMemcpyAsync(cd, o.block->cols.ptr + base, nnz * sizeof(uint64_t), hipMemcpyDeviceToDevice, fs);
        if (e == hipSuccess && nnz)
            e = hipMemcpyAsync(vd, o.block->vals.ptr + base, nnz * sizeof(float), hipMemcpyDeviceToDevice, fs);
        if (e != hipSuccess) {
            set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
            rc = FD_ERR_HIP;
        }
    }
    // the blocks go back to the pool only once the copies out of them have finished
    const hipError_t es = hipStreamSynchronize(fs);
    if (es != hipSuccess && rc == FD_OK) {
        set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(es));
        rc = FD_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (BeamOut& o : q->out) release_beam(s, o);
    }
    return rc;
}

int fd_serve_stats_get(fd_server* s, fd_serve_stats* out) {
    if (!s || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->requests_done = s->requests_done;
    out->beams_done = s->beams_done;
    out->beamlets_done = s->beamlets_done;
    out->batches = s->batches;
    out->first_submit_s = s->first_submit;
    out->last_done_s = s->last_done;
    out->peak_inflight_bytes = s->peak_bytes;
    out->peak_inflight_batches = s->peak_batches;
    if (!s->lat_ms.empty()) {
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double x = p * (double)(v.size() - 1);
            const size_t i = (size_t)x;
            const double f = x - (double)i;
            return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / (double)v.size();
        out->max_ms = v.back();
    }
    return FD_OK;
}

int fd_serve_stats_reset(fd_server* s) {
    if (!s) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_ms.clear();
    s->requests_done = s->beams_done = s->beamlets_done = s->batches = s->peak_bytes = 0;
    s->peak_batches = 0;
    s->first_submit = s->last_done = -1;
    return FD_OK;
}

int fd_serve_stop(fd_server* s) {
    if (!s) return FD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->threads) t.join();
    fd_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    // The slots' workspaces and the output pool's buffers are parked on the context for the next
    // server (without a parking stream they are freed).
    hipStream_t park = nullptr;
    {
        std::lock_guard<std::mutex> lk(ctx->park_mu);
        if (!ctx->park_stream && hipStreamCreateWithFlags(&ctx->park_stream, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            ctx->park_stream = nullptr;
        }
        park = ctx->park_stream;
    }
    for (size_t i = 0; i < s->ws.size(); i++) {
        if (s->streams[i]) (void)hipStreamSynchronize(s->streams[i]);
        Workspace& w = s->ws[i];
        if (park) {
            w.set_alloc_stream(park);
            w.desc_valid = false;
            w.desc_stream = nullptr;
            std::lock_guard<std::mutex> lk(ctx->park_mu);
            ctx->parked_ws.push_back(std::move(w));
            w = Workspace{};
        } else {
            w.release();  // frees ordered on the slot's stream
        }
        if (s->streams[i]) {
            (void)hipStreamSynchronize(s->streams[i]);
            (void)hipStreamDestroy(s->streams[i]);
        }
    }
    // every block, including those of requests never fetched
    if (s->fetch_stream) (void)hipStreamSynchronize(s->fetch_stream);
    for (auto& b : s->blocks) {
        if (park && (b->cols.ptr || b->vals.ptr)) {
            CsrBufs c;
            std::swap(c.cols, b->cols);
            std::swap(c.vals, b->vals);
            c.cols.stream = c.vals.stream = park;
            std::lock_guard<std::mutex> lk(ctx->park_mu);
            ctx->parked_csr.push_back(std::move(c));
        } else {
            b->cols.release();  // freed on the fetch stream
            b->vals.release();
        }
    }
    if (s->fetch_stream) {
        (void)hipStreamSynchronize(s->fetch_stream);
        (void)hipStreamDestroy(s->fetch_stream);
    }
    delete s;
    return FD_OK;
}

}  // extern "C"
