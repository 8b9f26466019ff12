// fastdose_imrt -- command-line drop-in for FastDose's dose mode (`IMRT --mode 0`).
//
// Accepts the reference's flag set (IMRTOpt/src/IMRTArgs.cpp:12-88) and follows IMRTOpt/main.cu:33-84:
// structures and masks -> density + bbox (IMRTInit.cu:276-352) -> spectrum and kernel tables ->
// beam bundles (IMRTBeamBundle.cu:15-127: PTV-centroid isocenter, beam list in degrees, beamlet/PTV
// flags, canonical beamlets) -> DoseMatConstruction (IMRTDoseMat.cu:14-301) -> doseMatFolder files
// and fluenceMap.bin.  Masks are raw u8 volumes `<masks>/<structure>.bin` (x fastest), the layout of
// the reference's InputMask folders, in place of the HDF5 roi_list reader.  Only mode 0 exists here.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>

#include "fastdose_hip.h"

namespace {

struct Flag {
    const char* name;
    bool multi;
    const char* def;  // nullptr: required (mode 0) or optional without default
    bool required;
    const char* help;
};

// IMRTArgs.cpp:12-88 (flags of mode 1 are accepted and ignored)
const Flag kFlags[] = {
    {"phantomDim", true, nullptr, true, "The phantom dimension"},
    {"voxelSize", true, nullptr, true, "The isotropic resolution [cm]"},
    {"SAD", false, nullptr, true, "Source-to-axis distance [cm]"},
    {"density", false, nullptr, true, "The path to the density raw file"},
    {"structures", true, nullptr, true, "The list of structures"},
    {"masks", false, nullptr, true, "Folder with one raw u8 mask <structure>.bin per structure"},
    {"primaryROI", false, nullptr, true, "The ROI to cover in dose calculation (the PTV)"},
    {"bboxROI", false, nullptr, true, "The region within which to calculate the dose"},
    {"structureInfo", false, "", false, "(optimisation only)"},
    {"params", false, "", false, "(optimisation only)"},
    {"beamlist", false, nullptr, true, "The path to the beamlist (gantry couch collimator, degrees)"},
    {"mode", false, "0", false, "0 for dose calculation"},
    {"deviceIdx", false, "0", false, "The device index"},
    {"spectrum", false, nullptr, true, "The path to the spectrum"},
    {"kernel", false, nullptr, true, "The path to the exponential CCCS kernel"},
    {"nPhi", false, "8", false, "The number of phi angles in convolution"},
    {"fluenceDim", false, "20", false, "Fluence map dimension"},
    {"beamletSize", false, "0.5", false, "Beamlet size in cm"},
    {"subFluenceDim", false, "16", false, "Subdivided fluence dimension"},
    {"subFluenceOn", false, "4", false, "Fluence pixels on in the subdivided fluence map"},
    {"longSpacing", false, "0.25", false, "Longitudinal voxel size in the dose calculation"},
    {"extent", false, "2.0", false, "Dose interpolation extent [cm]"},
    {"EstNonZeroElementsPerMat", false, "6000000", false, "Estimated non-zero elements per beam"},
    {"outputFolder", true, nullptr, true, "Folder that receives doseMatFolder"},
    {"planFolder", false, "", false, "(optimisation only)"},
    {"nBeamsReserve", false, "500", false, "Reserve space for beam allocation"},
    {"beamIdxDebug", false, "0", false, "(debug only)"},
    {"SIB", false, "0", false, "(optimisation only)"},
    {"referenceDose", false, "", false, "(optimisation only)"},
    {"beamletBatch", false, "8192", false, "Beamlets per batched launch across beams (0: one beam at a time)"},
};

using Args = std::map<std::string, std::vector<std::string>>;

void usage() {
    std::cout << "fastdose_imrt: CCCS beamlet dose matrix on an MI355X (IMRT --mode 0 flags)\n";
    for (const Flag& f : kFlags)
        std::cout << "  --" << f.name << (f.multi ? " v1 v2 ..." : " v") << (f.required ? "  (required)" : "")
                  << (f.def && *f.def ? std::string("  [") + f.def + "]" : std::string()) << "\n      " << f.help
                  << "\n";
}

const Flag* find_flag(const std::string& n) {
    for (const Flag& f : kFlags)
        if (n == f.name) return &f;
    return nullptr;
}

bool parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc;) {
        std::string tok = argv[i];
        if (tok.rfind("--", 0) != 0) {
            std::cerr << "unexpected argument " << tok << "\n";
            return false;
        }
        std::string name = tok.substr(2);
        const Flag* f = find_flag(name);
        if (!f) {
            std::cerr << "unknown option --" << name << "\n";
            return false;
        }
        std::vector<std::string> vals;
        i++;
        while (i < argc && std::string(argv[i]).rfind("--", 0) != 0) {
            vals.push_back(argv[i]);
            i++;
            if (!f->multi) break;
        }
        if (vals.empty()) {
            std::cerr << "option --" << name << " needs a value\n";
            return false;
        }
        a[name] = vals;
    }
    for (const Flag& f : kFlags) {
        if (a.count(f.name)) continue;
        if (f.required) {
            std::cerr << "the option '--" << f.name << "' is required but missing\n";
            return false;
        }
        if (f.def) a[f.name] = {f.def};
    }
    return true;
}

float fget(const Args& a, const char* n, size_t i = 0) { return std::stof(a.at(n).at(i)); }
int iget(const Args& a, const char* n, size_t i = 0) { return std::stoi(a.at(n).at(i)); }
const std::string& sget(const Args& a, const char* n, size_t i = 0) { return a.at(n).at(i); }

bool read_raw(const std::string& path, void* dst, size_t bytes) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    f.read((char*)dst, (std::streamsize)bytes);
    return (size_t)f.gcount() == bytes;
}

int fail(const std::string& what) {
    std::cerr << what << std::endl;
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 1 || std::string(argv[1]) == "--help") {
        usage();
        return 0;
    }
    Args a;
    if (!parse(argc, argv, a)) return 1;
    if (iget(a, "mode") != 0) return fail("only --mode 0 (dose calculation) is provided by this build");
    if (a["phantomDim"].size() != 3 || a["voxelSize"].size() != 3) return fail("phantomDim and voxelSize take 3 values");
    if (a["outputFolder"].size() != 1) return fail("Only one entry is expected in the argument \"outputFolder\"");

    fd_volume vol;
    std::memset(&vol, 0, sizeof(vol));
    for (int i = 0; i < 3; i++) {
        vol.dim[i] = (uint32_t)iget(a, "phantomDim", i);
        vol.voxel_size[i] = fget(a, "voxelSize", i);
    }
    const size_t nvox = (size_t)vol.dim[0] * vol.dim[1] * vol.dim[2];

    // structures: primary ROI and bbox ROI masks (IMRTInit.cu:73-95, 276-352)
    auto mask_path = [&](const std::string& name) { return sget(a, "masks") + "/" + name + ".bin"; };
    bool have_ptv = false, have_bbox = false;
    for (const auto& s : a["structures"]) {
        have_ptv |= s == sget(a, "primaryROI");
        have_bbox |= s == sget(a, "bboxROI");
    }
    if (!have_ptv) return fail("Cannot find the primary ROI, " + sget(a, "primaryROI") + " in structs");
    if (!have_bbox) return fail("The bounding box structure " + sget(a, "bboxROI") + " was not found");
    std::vector<uint8_t> ptv(nvox), bbox(nvox);
    if (!read_raw(mask_path(sget(a, "primaryROI")), ptv.data(), nvox)) return fail("Cannot read " + mask_path(sget(a, "primaryROI")));
    if (!read_raw(mask_path(sget(a, "bboxROI")), bbox.data(), nvox)) return fail("Cannot read " + mask_path(sget(a, "bboxROI")));
    // getBBox (IMRTInit.cu:328-352)
    uint32_t lo[3] = {vol.dim[0] - 1, vol.dim[1] - 1, vol.dim[2] - 1}, hi[3] = {0, 0, 0};
    bool any = false;
    for (uint32_t k = 0; k < vol.dim[2]; k++)
        for (uint32_t j = 0; j < vol.dim[1]; j++)
            for (uint32_t i = 0; i < vol.dim[0]; i++)
                if (bbox[i + (size_t)vol.dim[0] * (j + (size_t)vol.dim[1] * k)]) {
                    const uint32_t c[3] = {i, j, k};
                    for (int d = 0; d < 3; d++) {
                        lo[d] = std::min(lo[d], c[d]);
                        hi[d] = std::max(hi[d], c[d]);
                    }
                    any = true;
                }
    if (!any) return fail("the bounding-box mask is empty");
    for (int d = 0; d < 3; d++) {
        vol.bbox_start[d] = lo[d];
        vol.bbox_dim[d] = hi[d] + 1 - lo[d];
    }
    std::vector<float> density(nvox);
    if (!read_raw(sget(a, "density"), density.data(), nvox * sizeof(float)))
        return fail("Could not open file: " + sget(a, "density"));

    fd_spectrum spec;
    fd_kernel kern;
    if (fd_read_spectrum(sget(a, "spectrum").c_str(), &spec)) return fail(fd_last_error());
    if (fd_read_kernel(sget(a, "kernel").c_str(), iget(a, "nPhi"), &kern)) return fail(fd_last_error());

    // beam bundles (IMRTBeamBundle.cu:15-127)
    fd_bundle proto;
    std::memset(&proto, 0, sizeof(proto));
    if (fd_ptv_isocenter(ptv.data(), &vol, proto.isocenter)) return fail(fd_last_error());
    std::cout << "Isocenter: (" << proto.isocenter[0] << ", " << proto.isocenter[1] << ", " << proto.isocenter[2]
              << ") [cm]\n";
    proto.beamlet_size[0] = proto.beamlet_size[1] = fget(a, "beamletSize");
    proto.fluence_dim[0] = proto.fluence_dim[1] = iget(a, "fluenceDim");
    proto.sad = fget(a, "SAD");
    proto.long_spacing = fget(a, "longSpacing");
    proto.sub_dim[0] = proto.sub_dim[1] = iget(a, "subFluenceDim");
    proto.sub_on[0] = proto.sub_on[1] = iget(a, "subFluenceOn");
    std::vector<fd_bundle> bundles;
    {
        std::ifstream f(sget(a, "beamlist"));
        if (!f.is_open()) return fail("Cannot open file: " + sget(a, "beamlist"));
        std::string row;
        const float pi = 3.141592654f;
        while (std::getline(f, row)) {
            if (row.find_first_not_of(" \t\r\n") == std::string::npos) continue;
            std::istringstream iss(row);
            float deg[3] = {0, 0, 0};
            iss >> deg[0] >> deg[1] >> deg[2];
            fd_bundle b = proto;
            for (int d = 0; d < 3; d++) b.angles[d] = deg[d] * pi / 180;
            bundles.push_back(b);
        }
    }
    const int fmap = proto.fluence_dim[0] * proto.fluence_dim[1];
    std::vector<std::vector<uint8_t>> flags(bundles.size(), std::vector<uint8_t>(fmap));
    for (size_t i = 0; i < bundles.size(); i++)
        if (fd_beamlet_flags_host(&bundles[i], ptv.data(), &vol, flags[i].data()) && fd_last_error()[0])
            std::cerr << "beam " << i + 1 << ": " << fd_last_error() << "\n";

    fd_ctx* ctx = nullptr;
    if (fd_ctx_create(iget(a, "deviceIdx"), &vol, density.data(), &spec, &kern, &ctx)) return fail(fd_last_error());
    fd_dosemat* m = nullptr;
    if (fd_dosemat_create(nvox, (uint64_t)std::stoull(sget(a, "EstNonZeroElementsPerMat")), &m))
        return fail(fd_last_error());
    const float extent = fget(a, "extent");
    const size_t npix = (size_t)proto.sub_dim[0] * proto.sub_dim[1];
    const int batch = iget(a, "beamletBatch");
    // every beam's active beamlets, concatenated in beam order
    std::vector<fd_beam> beams((size_t)fmap * bundles.size());
    std::vector<float> fluence((size_t)fmap * bundles.size() * npix);
    std::vector<int> counts(bundles.size(), 0);
    size_t total = 0;
    for (size_t i = 0; i < bundles.size(); i++) {
        int n = 0;
        if (fd_beamlets_init(&bundles[i], flags[i].data(), &vol, beams.data() + total, fluence.data() + total * npix,
                             &n))
            return fail("beam " + std::to_string(i + 1) + ": " + fd_last_error());
        counts[i] = n;
        total += n;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (batch > 0) {
        if (fd_dosemat_add_beams(ctx, m, beams.data(), fluence.data(), counts.data(), (int)bundles.size(), batch,
                                 extent, nullptr))
            return fail(fd_last_error());
    } else {
        size_t first = 0;
        for (size_t i = 0; i < bundles.size(); i++) {
            if (fd_dosemat_add_beam(ctx, m, beams.data() + first, counts[i], fluence.data() + first * npix, extent,
                                    nullptr))
                return fail("beam " + std::to_string(i + 1) + ": " + fd_last_error());
            first += counts[i];
        }
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t i = 0; i < bundles.size(); i++) {
        uint64_t rows = 0, nnz = 0;
        fd_dosemat_matrix_size(m, i, &rows, &nnz);
        std::cout << "Beam " << i + 1 << " / " << bundles.size() << ": " << rows << " active beamlets, nnz " << nnz
                  << "\n";
    }
    std::cout << "Dose calculation time: " << s << " s (" << total << " beamlets)\n";

    const std::string out = sget(a, "outputFolder");
    mkdir(out.c_str(), 0777);
    const std::string folder = out + "/doseMatFolder";
    if (fd_dosemat_tofile(m, folder.c_str())) return fail(fd_last_error());
    std::vector<uint8_t> all;
    for (const auto& f : flags) all.insert(all.end(), f.begin(), f.end());
    if (fd_write_fluence_map((folder + "/fluenceMap.bin").c_str(), all.data(), bundles.size(), fmap))
        return fail(fd_last_error());
    std::cout << "Saved " << folder << "\n";
    fd_dosemat_destroy(m);
    fd_ctx_destroy(ctx);
    return 0;
}
