// api.hip -- extern "C" entry points of libfastdose_hip (declared in include/fastdose_hip.h).
//
// Host orchestration of the CCCS beamlet-dose path: context set-up (density_h2d, bind_spectrum,
// bind_kernel), the three stages (TermaComputeCollective, DoseComputeCollective,
// BEV2PVCSInterp + addMat) and the dose-matrix ensemble (MatCSREnsemble, DoseMatConstruction's
// per-beam iteration).  Reference: IMRTOpt/src/IMRTDoseMat.cu:14-301,
// IMRTOpt/src/IMRTSparseMatEns.cu:10-336, FastDoseLib/src/{density,spectrum,kernel}.cu.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>

#include "dosemat_files.h"
#include "engine.h"
#include "fd_internal.h"
#include "geometry.h"
#include "kernels.h"

namespace fdi {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

}  // namespace fdi

using namespace fdi;

namespace {

int check_ctx(fd_ctx* ctx) {
    if (!ctx) {
        set_error("null context");
        return FD_ERR_ARG;
    }
    FD_HIP(hipSetDevice(ctx->device));
    return FD_OK;
}

int check_stride(const fd_beam* beams, int n, uint64_t stride) {
    for (int i = 0; i < n; i++)
        if ((uint64_t)beams[i].long_dim * kNpix > stride) {
            set_error("stride smaller than long_dim * 256 of a beamlet");
            return FD_ERR_ARG;
        }
    return FD_OK;
}

ConeTable build_cones(const fd_kernel& k, int* n_fwd) {
    ConeTable c;
    std::memset(&c, 0, sizeof(c));
    // forward cones (DoseComp.cu:265 skips theta_mid > pi/2), widest first for load balance
    std::vector<int> fwd;
    for (int i = 0; i < k.n_theta; i++)
        if (!((double)k.theta_mid[i] > 1.57079632679489661923)) fwd.push_back(i);
    std::stable_sort(fwd.begin(), fwd.end(), [&](int a, int b) { return k.theta_mid[a] > k.theta_mid[b]; });
    c.n_fwd = (int)fwd.size();
    c.n_phi = k.n_phi;
    for (int t = 0; t < c.n_fwd; t++) {
        const int i = fwd[t];
        c.theta_idx[t] = i;
        c.A_a[t] = k.A[i] / k.a[i];
        c.B_b[t] = k.B[i] / k.b[i];
        c.a[t] = k.a[i];
        c.b[t] = k.b[i];
        c.na[t] = -k.a[i] * 1.4426950408889634f;
        c.nb[t] = -k.b[i] * 1.4426950408889634f;
        const double st = std::sin((double)k.theta_mid[i]), ct = std::cos((double)k.theta_mid[i]);
        for (int f = 0; f < k.n_phi; f++) {
            const double sp = std::sin((double)k.phi[f]), cp = std::cos((double)k.phi[f]);
            c.ux[t][f] = (float)(st * cp);
            c.uy[t][f] = (float)(st * sp);
            c.uz[t][f] = (float)ct;
        }
    }
    *n_fwd = c.n_fwd;
    return c;
}

bool write_file(const std::string& path, const void* data, size_t bytes) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.is_open()) return false;
    if (bytes) f.write((const char*)data, (std::streamsize)bytes);
    return (bool)f;
}

}  // namespace

extern "C" {

const char* fd_last_error(void) { return g_error.c_str(); }
const char* fd_version(void) { return "fastdose_hip 0.1 (gfx950)"; }

int fd_calc_range(fd_beam* beam, const fd_volume* vol) {
    if (!beam || !vol) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    return calc_range(*beam, *vol);
}

int fd_beamlet_init(const fd_bundle* bundle, int ix, int iy, const fd_volume* vol, fd_beam* out,
                    float* fluence_out) {
    if (!bundle || !vol || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (ix < 0 || iy < 0 || ix >= bundle->fluence_dim[0] || iy >= bundle->fluence_dim[1]) {
        set_error("beamlet index outside the fluence map");
        return FD_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    return beamlet_init(*bundle, ix, iy, *vol, *out, fluence_out);
}

int fd_beamlets_init(const fd_bundle* bundle, const uint8_t* flags, const fd_volume* vol, fd_beam* out,
                     float* fluence_out, int* n_out) {
    if (!bundle || !flags || !vol || !out || !n_out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    const int fx = bundle->fluence_dim[0], fy = bundle->fluence_dim[1];
    const size_t npix = (size_t)bundle->sub_dim[0] * bundle->sub_dim[1];
    int count = 0;
    for (int j = 0; j < fy; j++)
        for (int i = 0; i < fx; i++) {
            if (!flags[i + j * fx]) continue;
            std::memset(&out[count], 0, sizeof(fd_beam));
            int r = beamlet_init(*bundle, i, j, *vol, out[count], fluence_out ? fluence_out + npix * count : nullptr);
            if (r != FD_OK) return r;
            count++;
        }
    *n_out = count;
    return FD_OK;
}

int fd_ptv_isocenter(const uint8_t* mask, const fd_volume* vol, float iso[3]) {
    if (!mask || !vol || !iso) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    return ptv_isocenter(mask, *vol, iso);
}

int fd_beamlet_flags_host(const fd_bundle* bundle, const uint8_t* ptv_mask, const fd_volume* vol,
                          uint8_t* flags) {
    if (!bundle || !ptv_mask || !vol || !flags) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    return beamlet_flags(*bundle, ptv_mask, *vol, flags);
}

int fd_beamlet_flags(fd_ctx* ctx, const uint8_t* mask_dev, const fd_bundle* bundles, int n_bundles,
                     uint8_t* flags_host, void* stream) {
    FD_TRY(check_ctx(ctx));
    if (n_bundles < 0 || n_bundles > 65535 || (n_bundles > 0 && (!mask_dev || !bundles || !flags_host))) {
        set_error("null argument or n_bundles outside [0, 65535]");
        return FD_ERR_ARG;
    }
    if (n_bundles == 0) return FD_OK;
    std::vector<FlagBeamDev> d((size_t)n_bundles);
    uint64_t total = 0;
    int max_pixels = 0;
    for (int b = 0; b < n_bundles; b++) {
        const int fx = bundles[b].fluence_dim[0], fy = bundles[b].fluence_dim[1];
        if (fx <= 0 || fy <= 0 || fx > 1024 || fy > 1024) {
            set_error("fluence_dim must be in [1, 1024]");
            return FD_ERR_ARG;
        }
        (void)flag_beam_setup(bundles[b], ctx->vol, d[b]);  // a miss leaves L = 0: all flags 0
        d[b].out_off = (uint32_t)total;
        total += (uint64_t)fx * fy;
        max_pixels = std::max(max_pixels, fx * fy);
    }
    if (total >= (1ull << 32)) {
        set_error("too many flags");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    FD_TRY(ctx->flag_desc.reserve(d.size()));
    FD_TRY(ctx->flag_out.reserve(total));
    FD_HIP(hipMemcpyAsync(ctx->flag_desc.ptr, d.data(), d.size() * sizeof(FlagBeamDev), hipMemcpyHostToDevice, s));
    hipError_t e = launch_beamlet_flags(ctx->flag_desc.ptr, n_bundles, max_pixels, mask_dev, ctx->vol,
                                        ctx->flag_out.ptr, s);
    if (e != hipSuccess) {
        set_error(std::string("beamlet flags launch: ") + hipGetErrorString(e));
        return FD_ERR_HIP;
    }
    FD_HIP(hipMemcpyAsync(flags_host, ctx->flag_out.ptr, total, hipMemcpyDeviceToHost, s));
    FD_HIP(hipStreamSynchronize(s));
    return FD_OK;
}

int fd_read_spectrum(const char* path, fd_spectrum* out) {
    if (!path || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    std::ifstream f(path);
    if (!f.is_open()) {
        set_error(std::string("cannot open spectrum file ") + path);
        return FD_ERR_IO;
    }
    std::memset(out, 0, sizeof(*out));
    std::string line;
    float total = 0.0f;
    int n = 0;
    while (std::getline(f, line)) {
        if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;  // blank rows
        if (n >= FD_MAX_TABLE) {
            set_error("more than 24 spectrum rows");
            return FD_ERR_ARG;
        }
        std::istringstream row(line);
        std::string label;
        row >> out->energy[n] >> out->fluence[n] >> out->mu[n] >> out->mu_en[n] >> label;
        if (row.fail() && !row.eof()) {
            set_error("malformed spectrum row: " + line);
            return FD_ERR_IO;
        }
        total += out->fluence[n];
        n++;
    }
    out->n = n;
    for (int i = 0; i < n; i++) out->fluence[i] /= total;
    return FD_OK;
}

int fd_read_kernel(const char* path, int n_phi, fd_kernel* out) {
    if (!path || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (n_phi <= 0 || n_phi > kMaxPhi) {
        set_error("n_phi must be in [1, 24]");
        return FD_ERR_ARG;
    }
    std::ifstream f(path);
    if (!f.is_open()) {
        set_error(std::string("cannot open kernel file ") + path);
        return FD_ERR_IO;
    }
    std::memset(out, 0, sizeof(*out));
    std::string line;
    std::getline(f, line);  // header
    int n = 0;
    while (std::getline(f, line)) {
        if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        if (n >= FD_MAX_TABLE) {
            set_error("more than 24 kernel rows");
            return FD_ERR_ARG;
        }
        std::istringstream row(line);
        row >> out->theta_begin[n] >> out->theta_end[n] >> out->A[n] >> out->a[n] >> out->B[n] >> out->b[n];
        if (row.fail()) {
            set_error("malformed kernel row: " + line);
            return FD_ERR_IO;
        }
        out->theta_mid[n] = (out->theta_begin[n] + out->theta_end[n]) / 2;
        n++;
    }
    out->n_theta = n;
    out->n_phi = n_phi;
    const float step = 2 * 3.141592654f / n_phi;
    for (int i = 0; i < n_phi; i++) out->phi[i] = (float)((i + 0.5) * step);
    return FD_OK;
}

// Whether the resampler's reciprocal (hardware reciprocal + one Newton step) equals the IEEE
// quotient for every float with exponent in [-64, 64) on `device` -- the depths a sample can have
// (calc_range keeps them >= 2^-60).  Checked once per device and process (about 10^9 inputs, a
// millisecond); where it does not hold the resampler corrects the one significand where the
// step can tie.
static hipError_t rcp_exact_on_device(int device, bool& exact) {
    static std::mutex mu;
    static int state[256] = {0};  // 0 unknown, 1 exact, 2 not exact
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0 || device >= 256) {
        exact = false;
        return hipSuccess;
    }
    if (state[device] == 0) {
        uint32_t* d = nullptr;
        uint32_t h = 1;
        hipError_t e = hipMalloc((void**)&d, sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(d, 0, sizeof(uint32_t));
        if (e == hipSuccess) e = launch_rcp_check(-64, 128, d, nullptr);
        if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(uint32_t), hipMemcpyDeviceToHost);
        if (d) (void)hipFree(d);
        if (e != hipSuccess) return e;
        state[device] = h == 0 ? 1 : 2;
    }
    exact = state[device] == 1;
    return hipSuccess;
}

int fd_ctx_create(int device, const fd_volume* vol, const float* density_host, const fd_spectrum* spectrum,
                  const fd_kernel* kernel, fd_ctx** out) {
    if (!vol || !density_host || !spectrum || !kernel || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    *out = nullptr;
    if (spectrum->n <= 0 || spectrum->n > FD_MAX_TABLE) {
        set_error("spectrum must have 1..24 bins");
        return FD_ERR_ARG;
    }
    if (kernel->n_theta <= 0 || kernel->n_theta > FD_MAX_TABLE || kernel->n_phi <= 0 || kernel->n_phi > kMaxPhi) {
        set_error("kernel must have 1..24 zenith and 1..24 azimuth angles");
        return FD_ERR_ARG;
    }
    for (int d = 0; d < 3; d++)
        if (vol->dim[d] == 0 || !(vol->voxel_size[d] > 0.0f)) {
            set_error("volume dims and voxel sizes must be positive");
            return FD_ERR_ARG;
        }
    if (vol->dim[1] > 8192) {
        set_error("volume y dimension above 8192 (resampler row table lives in LDS)");
        return FD_ERR_UNSUPPORTED;
    }
    for (int i = 0; i < kernel->n_theta; i++) {
        const double th = kernel->theta_mid[i];
        if (!(th > 1.57079632679489661923) && std::cos(th) < kMinCosTheta) {
            set_error("forward cone too close to 90 degrees (cos(theta_mid) < 0.01)");
            return FD_ERR_UNSUPPORTED;
        }
    }
    int ndev = 0;
    FD_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("device index out of range");
        return FD_ERR_ARG;
    }
    FD_HIP(hipSetDevice(device));
    int n_cu = 0;
    FD_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
    fd_ctx* c = new fd_ctx();
    c->device = device;
    c->n_cu = std::max(n_cu, 1);
    c->vol = *vol;
    c->spectrum = *spectrum;
    c->kernel = *kernel;
    c->spec_dev.n = spectrum->n;
    for (int e = 0; e < spectrum->n; e++) {
        c->spec_dev.w[e] = spectrum->fluence[e] * spectrum->energy[e] * spectrum->mu[e];
        c->spec_dev.mu[e] = spectrum->mu[e];
    }
    c->cones = build_cones(*kernel, &c->n_fwd);
    // spectrum table over the radiological depth: up to (volume diagonal) x (max density)
    std::vector<float> tab;
    {
        const size_t nv = (size_t)vol->dim[0] * vol->dim[1] * vol->dim[2];
        float dmax = 0.0f;
        for (size_t i = 0; i < nv; i++) dmax = std::max(dmax, density_host[i]);
        double diag = 0.0;
        for (int d = 0; d < 3; d++) diag += std::pow((double)vol->dim[d] * vol->voxel_size[d], 2);
        const double rmax = std::sqrt(diag) * std::max((double)dmax, 1e-3) * 1.05 + 1.0;
        const int ntab = (int)std::min<double>(std::max<double>(std::ceil(rmax / kSpecTableStep) + 2, 256), 1 << 22);
        const double h = rmax / (ntab - 2);
        std::vector<double> f(ntab);
        for (int i = 0; i < ntab; i++) {
            double acc = 0.0;
            for (int e = 0; e < spectrum->n; e++)
                acc += (double)c->spec_dev.w[e] * std::exp(-(double)c->spec_dev.mu[e] * (i * h));
            f[i] = acc;
        }
        tab.resize(2 * (size_t)ntab);
        for (int i = 0; i < ntab; i++) {
            const double nxt = i + 1 < ntab ? f[i + 1] : f[i];
            tab[2 * i] = (float)f[i];
            tab[2 * i + 1] = (float)(nxt - f[i]);
        }
        c->spec_dev.tab_n = ntab;
        c->spec_dev.tab_inv_h = (float)(1.0 / h);
    }
    const size_t nvox = (size_t)vol->dim[0] * vol->dim[1] * vol->dim[2];
    auto fail = [&](hipError_t e, const char* what) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        fd_ctx_destroy(c);
        return FD_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&c->density, nvox * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc density");
    if ((e = hipMemcpy(c->density, density_host, nvox * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "density upload");
    {
        // per-voxel trilinear corner records for the Terma ray trace (built on the device)
        const size_t nrec = density_records_floats(*vol);
        if ((e = hipMalloc((void**)&c->density_pad, nrec * sizeof(float))) != hipSuccess)
            return fail(e, "hipMalloc density corner records");
        if ((e = launch_expand_density(c->density, *vol, c->density_pad, nullptr)) != hipSuccess ||
            (e = hipDeviceSynchronize()) != hipSuccess)
            return fail(e, "density corner records");
    }
    if ((e = rcp_exact_on_device(c->device, c->rcp_exact)) != hipSuccess) return fail(e, "reciprocal check");
    if ((e = hipMalloc((void**)&c->spec_tab, tab.size() * sizeof(float))) != hipSuccess)
        return fail(e, "hipMalloc spectrum table");
    if ((e = hipMemcpy(c->spec_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "spectrum table upload");
    c->spec_dev.tab = c->spec_tab;
    if ((e = hipMalloc((void**)&c->cones_dev, sizeof(ConeTable))) != hipSuccess) return fail(e, "hipMalloc cones");
    if ((e = hipMemcpy(c->cones_dev, &c->cones, sizeof(ConeTable), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "cone upload");
    {
        const int r = c->ws.init();
        if (r != FD_OK) {
            fd_ctx_destroy(c);
            return r;
        }
    }
    *out = c;
    return FD_OK;
}

int fd_ctx_destroy(fd_ctx* ctx) {
    if (!ctx) return FD_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->density) (void)hipFree(ctx->density);
    if (ctx->density_pad) (void)hipFree(ctx->density_pad);
    if (ctx->cones_dev) (void)hipFree(ctx->cones_dev);
    if (ctx->spec_tab) (void)hipFree(ctx->spec_tab);
    ctx->ws.release();
    ctx->flag_desc.release();
    ctx->flag_out.release();
    for (auto& w : ctx->parked_ws) w.release();  // frees ordered on park_stream
    for (auto& c : ctx->parked_csr) {
        c.cols.release();
        c.vals.release();
    }
    if (ctx->park_stream) {
        (void)hipStreamSynchronize(ctx->park_stream);
        (void)hipStreamDestroy(ctx->park_stream);
    }
    delete ctx;
    return FD_OK;
}

int fd_ctx_set_options(fd_ctx* ctx, int options) {
    if (!ctx) {
        set_error("null context");
        return FD_ERR_ARG;
    }
    if (options & ~(FD_OPT_EXACT_G | FD_OPT_RPL_FP64 | FD_OPT_TIMING | FD_OPT_RPL_FP64_ALL | FD_OPT_GRAPH |
                    FD_OPT_IEEE_DIV | FD_OPT_TABLES | FD_OPT_NO_TABLES)) {
        set_error("unknown option bits");
        return FD_ERR_ARG;
    }
    if ((options & FD_OPT_TABLES) && (options & FD_OPT_NO_TABLES)) {
        set_error("FD_OPT_TABLES and FD_OPT_NO_TABLES exclude each other");
        return FD_ERR_ARG;
    }
    ctx->options = options;
    ctx->ws.timing = (options & FD_OPT_TIMING) != 0;
    return FD_OK;
}

int fd_ctx_set_threshold(fd_ctx* ctx, float rel) {
    if (!ctx) {
        set_error("null context");
        return FD_ERR_ARG;
    }
    if (!(rel >= 0.0f && rel < 1.0f)) {
        set_error("relative threshold must be in [0, 1)");
        return FD_ERR_ARG;
    }
    ctx->threshold = rel;
    return FD_OK;
}

int fd_ctx_timings(fd_ctx* ctx, float* ms, int n) {
    if (!ctx || !ms || n < 0) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    Workspace& w = ctx->ws;
    for (int k = 0; k < FD_T_COUNT; k++) FD_TRY(w.fold(k));
    for (int k = 0; k < n; k++) ms[k] = (k < FD_T_COUNT && w.ev_used[k]) ? (float)w.acc_ms[k] : -1.0f;
    return FD_OK;
}

int fd_ctx_timing_launches(fd_ctx* ctx, int* launches, int n) {
    if (!ctx || !launches || n < 0) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    Workspace& w = ctx->ws;
    for (int k = 0; k < FD_T_COUNT; k++) FD_TRY(w.fold(k));
    for (int k = 0; k < n; k++) launches[k] = (k < FD_T_COUNT && w.ev_used[k]) ? w.acc_n[k] : 0;
    return FD_OK;
}

int fd_ctx_density(fd_ctx* ctx, const float** density_dev) {
    if (!ctx || !density_dev) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    *density_dev = ctx->density;
    return FD_OK;
}

int fd_terma(fd_ctx* ctx, const fd_beam* beams, int n, const float* fluence_dev, float* terma_dev,
             float* density_bev_dev, uint64_t stride, void* stream) {
    FD_TRY(check_ctx(ctx));
    ctx->ws.call++;
    FD_TRY(check_beams(beams, n));
    FD_TRY(check_stride(beams, n, stride));
    if (n == 0) return FD_OK;
    if (!fluence_dev || !terma_dev || !density_bev_dev) {
        set_error("null device buffer");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    FD_TRY(upload_beams(ctx->ws, beams, n, stride, s));
    return run_terma(ctx, ctx->ws, ctx->options, n, fluence_dev, terma_dev, density_bev_dev, (uint64_t)n * stride, s);
}

int fd_cccs(fd_ctx* ctx, const fd_beam* beams, int n, const float* terma_dev, const float* density_bev_dev,
            float* dose_bev_dev, uint64_t stride, void* stream) {
    FD_TRY(check_ctx(ctx));
    ctx->ws.call++;
    FD_TRY(check_beams(beams, n));
    FD_TRY(check_stride(beams, n, stride));
    if (n == 0) return FD_OK;
    if (!terma_dev || !density_bev_dev || !dose_bev_dev) {
        set_error("null device buffer");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    FD_TRY(upload_beams(ctx->ws, beams, n, stride, s));
    return run_cccs(ctx, ctx->ws, ctx->options, n, max_long_dim(beams, n), terma_dev, density_bev_dev, dose_bev_dev,
                    (uint64_t)n * stride, s, nullptr, false, stride % 4 == 0);
}

int fd_resample_csr(fd_ctx* ctx, const fd_beam* beams, int n, const float* dose_bev_dev, uint64_t stride,
                    float extent, fd_csr* out, void* stream) {
    FD_TRY(check_ctx(ctx));
    ctx->ws.call++;
    FD_TRY(check_beams(beams, n));
    FD_TRY(check_stride(beams, n, stride));
    if (!out || !out->offsets || (n > 0 && !dose_bev_dev)) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (!(extent >= 0.0f)) {
        set_error("extent must be non-negative");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    out->rows = (uint64_t)n;
    if (n == 0) {
        out->nnz = 0;
        FD_HIP(launch_fill_u64(out->offsets, 1, 0, s));
        return FD_OK;
    }
    FD_TRY(upload_beams(ctx->ws, beams, n, stride, s));
    uint64_t nnz = 0;
    FD_TRY(run_resample_csr(ctx, ctx->ws, beams, n, dose_bev_dev, extent, ctx->threshold, *out, &nnz, s));
    out->nnz = nnz;
    if (nnz > out->capacity || (nnz > 0 && (!out->columns || !out->values))) {
        set_error("CSR capacity too small; nnz reported in out->nnz");
        return FD_ERR_CAPACITY;
    }
    return FD_OK;
}

int fd_dense_to_csr(const float* dense_dev, uint64_t rows, uint64_t cols, fd_csr* out, void* stream) {
    if (!out || !out->offsets || (rows > 0 && cols > 0 && !dense_dev)) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (cols >= (1ull << 32) || rows >= (1ull << 31)) {
        set_error("dense matrix too large for the per-row u32 counts");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    out->rows = rows;
    uint32_t* cnt = nullptr;
    uint64_t* off = nullptr;
    uint64_t* host = nullptr;
    auto cleanup = [&]() {
        if (cnt) (void)hipFree(cnt);
        if (off) (void)hipFree(off);
        if (host) (void)hipHostFree(host);
    };
    auto hip_fail = [&](hipError_t e, const char* what) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        cleanup();
        return FD_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc((void**)&cnt, std::max<uint64_t>(rows, 1) * sizeof(uint32_t))) != hipSuccess)
        return hip_fail(e, "hipMalloc");
    if ((e = hipMalloc((void**)&off, (rows + 1) * sizeof(uint64_t))) != hipSuccess) return hip_fail(e, "hipMalloc");
    if ((e = hipHostMalloc((void**)&host, sizeof(uint64_t), hipHostMallocDefault)) != hipSuccess)
        return hip_fail(e, "hipHostMalloc");
    if (rows > 0 && cols > 0) {
        if ((e = launch_dense_count(dense_dev, rows, cols, cnt, s)) != hipSuccess) return hip_fail(e, "count");
    } else if (rows > 0) {
        if ((e = launch_fill_u32(cnt, rows, 0, s)) != hipSuccess) return hip_fail(e, "fill");
    }
    if ((e = launch_scan_u32(cnt, rows, off, off + rows, nullptr, s)) != hipSuccess) return hip_fail(e, "scan");
    if ((e = hipMemcpyAsync(host, off + rows, sizeof(uint64_t), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_fail(e, "copy");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "sync");
    out->nnz = host[0];
    if (out->nnz > out->capacity || (out->nnz > 0 && (!out->columns || !out->values))) {
        cleanup();
        set_error("CSR capacity too small; nnz reported in out->nnz");
        return FD_ERR_CAPACITY;
    }
    if (rows > 0 && cols > 0 && out->nnz > 0)
        if ((e = launch_dense_fill(dense_dev, rows, cols, off, out->columns, out->values, s)) != hipSuccess)
            return hip_fail(e, "fill");
    if ((e = launch_write_offsets(off, off + rows, rows, out->offsets, s)) != hipSuccess) return hip_fail(e, "offsets");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "sync");
    cleanup();
    return FD_OK;
}

int fd_cccs_segments(fd_ctx* ctx, const fd_beam* beams, int n, int64_t* nseg_out) {
    if (!ctx || !nseg_out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    FD_TRY(check_beams(beams, n));
    // The same FP32 lattice walk as the kernel, without the physics.
    for (int i = 0; i < n; i++) {
        const fd_beam& B = beams[i];
        int64_t total = 0;
        for (int t = 0; t < ctx->n_fwd; t++) {
            const int nphi = ctx->cones.n_phi;
            float dxs[kMaxPhi], dys[kMaxPhi], dzs[kMaxPhi], ixs[kMaxPhi], iys[kMaxPhi], izs[kMaxPhi];
            float px[kMaxPhi], py[kMaxPhi], pz[kMaxPhi];
            for (int f = 0; f < nphi; f++) {
                const float vx = ctx->cones.ux[t][f] / B.beamlet_size[0];
                const float vy = ctx->cones.uy[t][f] / B.beamlet_size[1];
                const float vz = ctx->cones.uz[t][f] / B.long_spacing;
                const float rn = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
                dxs[f] = vx * rn; dys[f] = vy * rn; dzs[f] = vz * rn;
                ixs[f] = 1.0f / dxs[f]; iys[f] = 1.0f / dys[f]; izs[f] = 1.0f / dzs[f];
                px[f] = 0.5f; py[f] = 0.5f; pz[f] = 0.0f;
            }
            auto step = [&](float c, float d, float inv) {
                return d > 0.0f ? (ceilf(c + kLargerEps) - c) * inv
                                : (d < 0.0f ? (floorf(c - kLargerEps) - c) * inv : INFINITY);
            };
            int s = 0;
            while (s < (int)B.long_dim) {
                for (int f = 0; f < nphi; f++) {
                    for (int it = 0; it < kMaxStepsPerSlice; it++) {
                        const float sx = step(px[f], dxs[f], ixs[f]);
                        const float sy = step(py[f], dys[f], iys[f]);
                        const float sz = step(pz[f], dzs[f], izs[f]);
                        const float st = fminf(fminf(sx, sy), sz);
                        float nx = fmaf(st, dxs[f], px[f]), ny = fmaf(st, dys[f], py[f]), nz = fmaf(st, dzs[f], pz[f]);
                        const bool flag = fabsf(nz - roundf(nz)) < kEps;
                        if (flag) {
                            nx = fmaf(sz, dxs[f], px[f]);
                            ny = fmaf(sz, dys[f], py[f]);
                            nz = roundf(fmaf(sz, dzs[f], pz[f]));
                        }
                        px[f] = nx; py[f] = ny; pz[f] = nz;
                        total += kNpix;
                        if (flag) break;
                    }
                }
                s = std::max((int)roundf(pz[0]), s + 1);
            }
        }
        nseg_out[i] = total;
    }
    return FD_OK;
}

int fd_dosemat_create(uint64_t n_cols, uint64_t est_nnz, fd_dosemat** out) {
    if (!out || n_cols == 0) {
        set_error("invalid argument");
        return FD_ERR_ARG;
    }
    fd_dosemat* m = new fd_dosemat();
    m->n_cols = n_cols;
    if (est_nnz) {
        int r = grow_keep(m->columns, 0, est_nnz, nullptr);
        if (r == FD_OK) r = grow_keep(m->values, 0, est_nnz, nullptr);
        if (r != FD_OK) {
            fd_dosemat_destroy(m);
            return r;
        }
    }
    *out = m;
    return FD_OK;
}

int fd_dosemat_destroy(fd_dosemat* m) {
    if (!m) return FD_OK;
    m->offsets.release();
    m->columns.release();
    m->values.release();
    m->values16.release();
    m->row_scale.release();
    delete m;
    return FD_OK;
}

int fd_dosemat_add_beams(fd_ctx* ctx, fd_dosemat* m, const fd_beam* beamlets, const float* fluence,
                         const int* counts, int n_beams, int max_chunk, float extent, void* stream) {
    FD_TRY(check_ctx(ctx));
    ctx->ws.call++;
    if (!m || n_beams < 0 || (n_beams > 0 && !counts)) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    std::vector<uint64_t> first((size_t)n_beams + 1, 0);
    for (int k = 0; k < n_beams; k++) {
        if (counts[k] < 0) {
            set_error("negative beamlet count");
            return FD_ERR_ARG;
        }
        first[k + 1] = first[k] + (uint64_t)counts[k];
    }
    const uint64_t total = first[n_beams];
    if (total > (uint64_t)INT32_MAX) {
        set_error("too many beamlets");
        return FD_ERR_ARG;
    }
    FD_TRY(check_beams(beamlets, (int)total));
    if (total > 0 && !fluence) {
        set_error("null fluence");
        return FD_ERR_ARG;
    }
    const uint64_t nvox = (uint64_t)ctx->vol.dim[0] * ctx->vol.dim[1] * ctx->vol.dim[2];
    if (nvox != m->n_cols) {
        set_error("dose matrix column count differs from the context's voxel count");
        return FD_ERR_ARG;
    }
    if (max_chunk <= 0) max_chunk = ctx->tune.batch_beamlets;
    const int options = ctx->options | (ctx->tune.graph ? FD_OPT_GRAPH : 0);
    hipStream_t s = (hipStream_t)stream;
    std::vector<uint64_t> host_off;
    BatchResult res;
    int k0 = 0;
    while (k0 < n_beams) {
        // whole beams up to max_chunk beamlets (a larger single beam goes alone)
        int k1 = k0;
        uint64_t n = 0;
        while (k1 < n_beams && (k1 == k0 || n + (uint64_t)counts[k1] <= (uint64_t)max_chunk)) n += counts[k1++];
        // the batch's entries straight into the ensemble's free space when they fit there
        const uint64_t have = std::min<uint64_t>(m->columns.cap, m->values.cap);
        const uint64_t room = have > m->used_nnz ? have - m->used_nnz : 0;
        const CsrDst sink{m->columns.ptr + m->used_nnz, m->values.ptr + m->used_nnz, room};
        FD_TRY(run_batch(ctx, ctx->ws, options, beamlets + first[k0], (int)n, fluence + first[k0] * kNpix, extent,
                         ctx->threshold, s, &res, room ? &sink : nullptr, true));
        // one matrix per beam: offsets re-based to each beam's first row
        host_off.clear();
        host_off.reserve(n + (k1 - k0));
        if (!res.in_sink) {
            FD_TRY(grow_keep(m->columns, m->used_nnz, m->used_nnz + std::max<uint64_t>(res.nnz, 1), s));
            FD_TRY(grow_keep(m->values, m->used_nnz, m->used_nnz + std::max<uint64_t>(res.nnz, 1), s));
        }
        if (res.nnz && !res.in_sink) {
            FD_HIP(hipMemcpyAsync(m->columns.ptr + m->used_nnz, ctx->ws.out_cols.ptr, res.nnz * sizeof(uint64_t),
                                  hipMemcpyDeviceToDevice, s));
            FD_HIP(hipMemcpyAsync(m->values.ptr + m->used_nnz, ctx->ws.out_vals.ptr, res.nnz * sizeof(float),
                                  hipMemcpyDeviceToDevice, s));
        }
        const uint64_t off_base = m->used_off;
        for (int k = k0; k < k1; k++) {
            const uint64_t a = first[k] - first[k0], c = (uint64_t)counts[k];
            const uint64_t base = res.rowoff[a];
            m->rows.push_back(c);
            m->nnz.push_back(res.rowoff[a + c] - base);
            m->off_off.push_back(off_base + host_off.size());
            m->nnz_off.push_back(m->used_nnz + base);
            for (uint64_t j = 0; j <= c; j++) host_off.push_back(res.rowoff[a + j] - base);
        }
        m->packed = -1;  // a packed copy no longer covers the ensemble
        FD_TRY(grow_keep(m->offsets, m->used_off, m->used_off + host_off.size(), s));
        FD_HIP(hipMemcpyAsync(m->offsets.ptr + m->used_off, host_off.data(), host_off.size() * sizeof(uint64_t),
                              hipMemcpyHostToDevice, s));
        FD_HIP(hipStreamSynchronize(s));
        m->used_off += host_off.size();
        m->used_nnz += res.nnz;
        k0 = k1;
    }
    return FD_OK;
}

int fd_dosemat_add_beam(fd_ctx* ctx, fd_dosemat* m, const fd_beam* beamlets, int n, const float* fluence,
                        float extent, void* stream) {
    if (n < 0) {
        set_error("negative beamlet count");
        return FD_ERR_ARG;
    }
    return fd_dosemat_add_beams(ctx, m, beamlets, fluence, &n, 1, n, extent, stream);
}

int fd_ctx_get_tuning(const fd_ctx* ctx, fd_tuning* out) {
    if (!ctx || !out) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    *out = ctx->tune;
    return FD_OK;
}

int fd_ctx_set_tuning(fd_ctx* ctx, const fd_tuning* t) {
    FD_TRY(check_ctx(ctx));
    if (!t || t->batch_beamlets < 1 || (t->graph != 0 && t->graph != 1) ||
        (t->terma_segments != 1 && t->terma_segments != 2 && t->terma_segments != 4 && t->terma_segments != 8 &&
         t->terma_segments != 16)) {
        set_error("tuning: batch_beamlets >= 1, graph 0 or 1, terma_segments 1, 2, 4, 8 or 16");
        return FD_ERR_ARG;
    }
    ctx->tune = *t;
    return FD_OK;
}

int fd_autotune(fd_ctx* ctx, const fd_beam* beamlets, const float* fluence, const int* counts, int n_beams,
                float extent, fd_tuning* out, void* stream) {
    FD_TRY(check_ctx(ctx));
    if (!out || n_beams <= 0 || !counts) {
        set_error("autotune needs beams and an output");
        return FD_ERR_ARG;
    }
    uint64_t total = 0;
    for (int k = 0; k < n_beams; k++) total += counts[k] > 0 ? (uint64_t)counts[k] : 0;
    if (total == 0) {
        set_error("autotune needs at least one beamlet");
        return FD_ERR_ARG;
    }
    const uint64_t nvox = (uint64_t)ctx->vol.dim[0] * ctx->vol.dim[1] * ctx->vol.dim[2];
    const fd_tuning saved = ctx->tune;
    // one setting, run twice on a fresh matrix each time; beamlets per second of the faster run
    auto trial = [&](const fd_tuning& t, double& bps) -> int {
        ctx->tune = t;
        double best_s = 1e30;
        for (int rep = 0; rep < 2; rep++) {
            fd_dosemat* m = nullptr;
            int rc = fd_dosemat_create(nvox, 0, &m);
            if (rc != FD_OK) return rc;
            const auto t0 = std::chrono::steady_clock::now();
            rc = fd_dosemat_add_beams(ctx, m, beamlets, fluence, counts, n_beams, 0, extent, stream);
            if (rc == FD_OK && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) rc = FD_ERR_HIP;
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            fd_dosemat_destroy(m);
            if (rc != FD_OK) return rc;
            best_s = std::min(best_s, sec);
        }
        bps = (double)total / best_s;
        return FD_OK;
    };
    fd_tuning best = saved;
    best.beamlets_per_s = 0.0f;
    double best_bps = 0.0, bps = 0.0;
    int rc = FD_OK;
    auto consider = [&](fd_tuning t) {
        if (rc != FD_OK) return;
        rc = trial(t, bps);
        if (rc == FD_OK && bps > best_bps) {
            best_bps = bps;
            best = t;
        }
    };
    consider(best);  // also warms up the workspace
    for (int b : {1024, 2048, 4096, 8192, 16384}) {
        if (b == best.batch_beamlets || (uint64_t)b >= 2 * total + 1024) continue;
        fd_tuning t = best;
        t.batch_beamlets = b;
        consider(t);
    }
    {
        fd_tuning t = best;
        t.graph = 1 - t.graph;
        consider(t);
    }
    for (int g : {4, 8, 16}) {
        if (g == best.terma_segments) continue;
        fd_tuning t = best;
        t.terma_segments = g;
        consider(t);
    }
    if (rc != FD_OK) {
        ctx->tune = saved;
        return rc;
    }
    best.beamlets_per_s = (float)best_bps;
    ctx->tune = best;
    *out = best;
    return FD_OK;
}

int fd_dosemat_add_csr(fd_dosemat* m, const fd_csr* csr, void* stream) {
    if (!m || !csr || !csr->offsets) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    uint64_t *off_dst, *col_dst;
    float* val_dst;
    FD_TRY(append_matrix(m, csr->rows, csr->nnz, s, &off_dst, &col_dst, &val_dst));
    FD_HIP(hipMemcpyAsync(off_dst, csr->offsets, (csr->rows + 1) * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    if (csr->nnz) {
        FD_HIP(hipMemcpyAsync(col_dst, csr->columns, csr->nnz * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        FD_HIP(hipMemcpyAsync(val_dst, csr->values, csr->nnz * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    FD_HIP(hipStreamSynchronize(s));
    return FD_OK;
}

int fd_dosemat_info(const fd_dosemat* m, uint64_t* n_matrices, uint64_t* total_rows, uint64_t* total_nnz) {
    if (!m) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    uint64_t r = 0;
    for (uint64_t x : m->rows) r += x;
    if (n_matrices) *n_matrices = m->rows.size();
    if (total_rows) *total_rows = r;
    if (total_nnz) *total_nnz = m->used_nnz;
    return FD_OK;
}

int fd_dosemat_matrix_size(const fd_dosemat* m, uint64_t k, uint64_t* rows, uint64_t* nnz) {
    if (!m || k >= m->rows.size()) {
        set_error("matrix index out of range");
        return FD_ERR_ARG;
    }
    if (rows) *rows = m->rows[k];
    if (nnz) *nnz = m->nnz[k];
    return FD_OK;
}

int fd_dosemat_copy_matrix(const fd_dosemat* m, uint64_t k, uint64_t* offsets, uint64_t* columns, float* values) {
    if (!m || k >= m->rows.size()) {
        set_error("matrix index out of range");
        return FD_ERR_ARG;
    }
    FD_HIP(hipDeviceSynchronize());
    if (offsets)
        FD_HIP(hipMemcpy(offsets, m->offsets.ptr + m->off_off[k], (m->rows[k] + 1) * sizeof(uint64_t),
                         hipMemcpyDeviceToHost));
    if (m->nnz[k]) {
        if (columns)
            FD_HIP(hipMemcpy(columns, m->columns.ptr + m->nnz_off[k], m->nnz[k] * sizeof(uint64_t),
                             hipMemcpyDeviceToHost));
        if (values)
            FD_HIP(hipMemcpy(values, m->values.ptr + m->nnz_off[k], m->nnz[k] * sizeof(float), hipMemcpyDeviceToHost));
    }
    return FD_OK;
}

int fd_dosemat_tofile(const fd_dosemat* m, const char* folder) {
    if (!m || !folder) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    const std::string dir(folder);
    mkdir(dir.c_str(), 0777);
    const size_t nm = m->rows.size();
    std::vector<uint64_t> cumu_nnz(nm), cumu_rows(nm);
    uint64_t a = 0, r = 0;
    for (size_t i = 0; i < nm; i++) {
        a += m->nnz[i];
        r += m->rows[i];
        cumu_nnz[i] = a;
        cumu_rows[i] = r;
    }
    const std::pair<const std::vector<uint64_t>*, const char*> small[] = {
        {&m->nnz, "NonZeroElements"},   {&cumu_nnz, "CumuNonZeroElements"}, {&m->rows, "numRowsPerMat"},
        {&cumu_rows, "CumuNumRowsPerMat"}, {&m->off_off, "OffsetBufferIdx"}};
    for (const auto& f : small)
        if (!write_file(dir + "/" + f.second + ".bin", f.first->data(), nm * sizeof(uint64_t))) {
            set_error("cannot write " + dir + "/" + f.second + ".bin");
            return FD_ERR_IO;
        }
    FD_HIP(hipDeviceSynchronize());
    std::vector<uint64_t> off(m->used_off), col(m->used_nnz);
    std::vector<float> val(m->used_nnz);
    if (m->used_off) FD_HIP(hipMemcpy(off.data(), m->offsets.ptr, m->used_off * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (m->used_nnz) {
        FD_HIP(hipMemcpy(col.data(), m->columns.ptr, m->used_nnz * sizeof(uint64_t), hipMemcpyDeviceToHost));
        FD_HIP(hipMemcpy(val.data(), m->values.ptr, m->used_nnz * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (!write_file(dir + "/offsetsBuffer.bin", off.data(), off.size() * sizeof(uint64_t)) ||
        !write_file(dir + "/columnsBuffer.bin", col.data(), col.size() * sizeof(uint64_t)) ||
        !write_file(dir + "/valuesBuffer.bin", val.data(), val.size() * sizeof(float))) {
        set_error("cannot write the CSR buffers under " + dir);
        return FD_ERR_IO;
    }
    if (m->packed == FD_VAL_BF16 || m->packed == FD_VAL_FP16) {
        const uint64_t R = m->total_rows();
        std::vector<uint16_t> q(m->used_nnz);
        std::vector<float> sc(R);
        if (m->used_nnz)
            FD_HIP(hipMemcpy(q.data(), m->values16.ptr, m->used_nnz * sizeof(uint16_t), hipMemcpyDeviceToHost));
        if (R) FD_HIP(hipMemcpy(sc.data(), m->row_scale.ptr, R * sizeof(float), hipMemcpyDeviceToHost));
        const std::string vname = m->packed == FD_VAL_BF16 ? "/valuesBuffer.bf16.bin" : "/valuesBuffer.fp16.bin";
        if (!write_file(dir + vname, q.data(), q.size() * sizeof(uint16_t)) ||
            !write_file(dir + "/beamletScale.bin", sc.data(), sc.size() * sizeof(float))) {
            set_error("cannot write the packed values under " + dir);
            return FD_ERR_IO;
        }
    }
    return FD_OK;
}

int fd_dosemat_fromfile(fd_dosemat* m, const char* folder) {
    if (!m || !folder) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    // Everything is read and checked on the host before the ensemble or the device is touched: a
    // folder written for another volume, or a truncated one, must not become an out-of-range index
    // in a kernel (the D^T X transpose counts columns with atomics, the copies follow the offsets).
    FolderCSR f;
    FD_TRY(read_dosemat_folder(folder, m->n_cols, false, f));
    const uint64_t n_off = f.n_off, n_nnz = f.n_nnz;
    const std::vector<char>& off_b = f.off_b;
    const std::vector<char>& col_b = f.col_b;
    const std::vector<char>& val_b = f.val_b;
    fdi::DevBuf<uint64_t> offsets, columns;
    fdi::DevBuf<float> values;
    auto upload = [&]() -> int {
        FD_TRY(offsets.reserve(std::max<uint64_t>(n_off, 1)));
        FD_TRY(columns.reserve(std::max<uint64_t>(n_nnz, 1)));
        FD_TRY(values.reserve(std::max<uint64_t>(n_nnz, 1)));
        if (n_off) FD_HIP(hipMemcpy(offsets.ptr, off_b.data(), off_b.size(), hipMemcpyHostToDevice));
        if (n_nnz) {
            FD_HIP(hipMemcpy(columns.ptr, col_b.data(), col_b.size(), hipMemcpyHostToDevice));
            FD_HIP(hipMemcpy(values.ptr, val_b.data(), val_b.size(), hipMemcpyHostToDevice));
        }
        return FD_OK;
    };
    const int rc = upload();
    if (rc != FD_OK) {
        offsets.release();
        columns.release();
        values.release();
        return rc;
    }
    // swap the new contents in only now
    std::swap(m->offsets, offsets);
    std::swap(m->columns, columns);
    std::swap(m->values, values);
    offsets.release();
    columns.release();
    values.release();
    m->values16.release();
    m->row_scale.release();
    m->rows.swap(f.rows);
    m->nnz.swap(f.nnz);
    m->off_off.swap(f.off_off);
    m->nnz_off.swap(f.nnz_off);
    m->used_off = n_off;
    m->used_nnz = n_nnz;
    m->packed = -1;
    return FD_OK;
}

int fd_write_fluence_map(const char* path, const uint8_t* flags, uint64_t n_beams, uint64_t per_beam) {
    if (!path || (!flags && n_beams * per_beam)) {
        set_error("null argument");
        return FD_ERR_ARG;
    }
    if (!write_file(path, flags, n_beams * per_beam)) {
        set_error(std::string("cannot open file: ") + path);
        return FD_ERR_IO;
    }
    return FD_OK;
}

}  // extern "C"
