// Host-only code of libfastdose_hip under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Random and degenerate inputs through the beam geometry (BEAM_h::calc_range, beam.cu:134-221;
// BeamBundle::beamletInit, IMRTBeamBundle.cu:341-569; beamletFlagInit, IMRTBeamBundle.cu:130-316;
// the PTV isocenter, IMRTBeamBundle.cu:36-65; the kernels' beamlet descriptors) and through the
// doseMatFolder reader and merger (scripts/doseMerge.py:9-63) on valid, truncated and corrupted
// folders.  Any sanitizer report aborts the run; the driver also checks the documented status of
// every call and exits non-zero on a mismatch.
//
//     sanitize_driver <scratch directory>
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "dosemat_files.h"
#include "geometry.h"

namespace fdi {
static std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace fdi

using namespace fdi;

static int failures = 0;
#define EXPECT(cond, what)                                                     \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::fprintf(stderr, "check failed: %s (%s)\n", what, #cond);      \
            failures++;                                                        \
        }                                                                      \
    } while (0)

static void write_u64(const std::string& path, const std::vector<uint64_t>& v) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    f.write((const char*)v.data(), (std::streamsize)(v.size() * 8));
}
static void write_f32(const std::string& path, const std::vector<float>& v) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    f.write((const char*)v.data(), (std::streamsize)(v.size() * 4));
}

// A valid two-matrix folder over n_cols voxels (rows 3 and 2, ascending columns per row).
static void write_folder(const std::string& dir, uint64_t n_cols, std::mt19937_64& rng) {
    mkdir(dir.c_str(), 0777);
    std::vector<uint64_t> nnz, rows = {3, 2}, off, cols, cumu_nnz, cumu_rows, oidx;
    std::vector<float> vals;
    uint64_t a = 0, r = 0;
    for (uint64_t m = 0; m < 2; m++) {
        oidx.push_back(off.size());
        uint64_t k = 0;
        off.push_back(0);
        for (uint64_t row = 0; row < rows[m]; row++) {
            const uint64_t n = 1 + rng() % 4;
            uint64_t c = rng() % (n_cols / 8);
            for (uint64_t j = 0; j < n; j++) {
                cols.push_back(c);
                vals.push_back(0.5f + (float)j);
                c += 1 + rng() % 3;
            }
            k += n;
            off.push_back(k);
        }
        nnz.push_back(k);
        a += k;
        r += rows[m];
        cumu_nnz.push_back(a);
        cumu_rows.push_back(r);
    }
    write_u64(dir + "/NonZeroElements.bin", nnz);
    write_u64(dir + "/CumuNonZeroElements.bin", cumu_nnz);
    write_u64(dir + "/numRowsPerMat.bin", rows);
    write_u64(dir + "/CumuNumRowsPerMat.bin", cumu_rows);
    write_u64(dir + "/OffsetBufferIdx.bin", oidx);
    write_u64(dir + "/offsetsBuffer.bin", off);
    write_u64(dir + "/columnsBuffer.bin", cols);
    write_f32(dir + "/valuesBuffer.bin", vals);
    std::vector<char> flags(8, 1);
    std::ofstream(dir + "/fluenceMap.bin", std::ios::binary).write(flags.data(), 8);
}

static void geometry_fuzz(std::mt19937_64& rng, int iters) {
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    int ok_range = 0, ok_init = 0;
    for (int it = 0; it < iters; it++) {
        fd_volume vol{};
        for (int d = 0; d < 3; d++) {
            vol.voxel_size[d] = 0.1f + 0.4f * U(rng);
            vol.dim[d] = 1 + (uint32_t)(rng() % 48);
            vol.bbox_start[d] = (uint32_t)(rng() % vol.dim[d]);
            vol.bbox_dim[d] = 1 + (uint32_t)(rng() % (vol.dim[d] - vol.bbox_start[d]));
        }
        const size_t nvox = (size_t)vol.dim[0] * vol.dim[1] * vol.dim[2];
        fd_bundle bb{};
        for (int d = 0; d < 3; d++) bb.isocenter[d] = (U(rng) * 1.4f - 0.2f) * vol.dim[d] * vol.voxel_size[d];
        bb.beamlet_size[0] = bb.beamlet_size[1] = 0.2f + 0.8f * U(rng);
        bb.fluence_dim[0] = bb.fluence_dim[1] = 1 + (int)(rng() % 12);
        bb.sad = 50.0f + 100.0f * U(rng);
        for (int d = 0; d < 3; d++) bb.angles[d] = (U(rng) * 4.0f - 2.0f) * 3.14159265f;
        if (it % 7 == 0) bb.angles[1] = bb.angles[2] = 0.0f;  // axis-aligned cases
        bb.long_spacing = 0.1f + 0.4f * U(rng);
        const bool small = it % 5 == 0;
        bb.sub_dim[0] = bb.sub_dim[1] = small ? 8 : 16;
        bb.sub_on[0] = bb.sub_on[1] = small ? 2 : 4;
        // central beam: calc_range
        fd_beam beam{};
        std::memcpy(beam.isocenter, bb.isocenter, sizeof(beam.isocenter));
        beam.beamlet_size[0] = bb.beamlet_size[0] * bb.fluence_dim[0];
        beam.beamlet_size[1] = bb.beamlet_size[1] * bb.fluence_dim[1];
        beam.fmap_size[0] = beam.fmap_size[1] = 16;
        beam.sad = bb.sad;
        std::memcpy(beam.angles, bb.angles, sizeof(beam.angles));
        beam.long_spacing = bb.long_spacing;
        const int rc = calc_range(beam, vol);
        EXPECT(rc == FD_OK || rc == FD_ERR_GEOM, "calc_range status");
        if (rc == FD_OK) {
            ok_range++;
            EXPECT(beam.long_dim > 0 && beam.lim_max >= beam.lim_min, "calc_range result");
        }
        // every beamlet of the bundle: canonical form + range, descriptors
        std::vector<float> flu((size_t)bb.sub_dim[0] * bb.sub_dim[1]);
        for (int iy = 0; iy < bb.fluence_dim[1]; iy++)
            for (int ix = 0; ix < bb.fluence_dim[0]; ix++) {
                fd_beam out{};
                const int r2 = beamlet_init(bb, ix, iy, vol, out, flu.data());
                EXPECT(r2 == FD_OK || r2 == FD_ERR_GEOM, "beamlet_init status");
                if (r2 == FD_OK) {
                    ok_init++;
                    BeamletDev d{};
                    fill_beamlet_dev(out, (uint64_t)(iy * bb.fluence_dim[0] + ix) * 256, d);
                    EXPECT(d.L == (int32_t)out.long_dim, "descriptor long_dim");
                }
            }
        // flags and isocenter on a random PTV mask (sometimes empty)
        std::vector<uint8_t> mask(nvox, 0), flags((size_t)bb.fluence_dim[0] * bb.fluence_dim[1], 7);
        const uint64_t density = it % 4 == 0 ? 0 : 1 + rng() % 5;
        for (size_t v = 0; v < nvox; v++) mask[v] = density && rng() % (density * 3) == 0;
        const int r3 = beamlet_flags(bb, mask.data(), vol, flags.data());
        EXPECT(r3 == FD_OK || r3 == FD_ERR_GEOM, "beamlet_flags status");
        for (uint8_t f : flags) EXPECT(r3 != FD_OK || f <= 1, "flag values");
        FlagBeamDev fd{};
        const int r4 = flag_beam_setup(bb, vol, fd);
        EXPECT(r4 == FD_OK || r4 == FD_ERR_GEOM, "flag_beam_setup status");
        float iso[3];
        const int r5 = ptv_isocenter(mask.data(), vol, iso);
        bool any = false;
        for (uint8_t m : mask) any = any || m;
        EXPECT(any ? r5 == FD_OK : r5 != FD_OK, "ptv_isocenter status");
    }
    std::printf("geometry: %d iterations, %d central beams in range, %d beamlets initialised\n", iters, ok_range,
                ok_init);
    EXPECT(ok_range > iters / 30 && ok_init > iters, "fuzz reached the successful paths");
}

static void folder_checks(const std::string& root, std::mt19937_64& rng) {
    const uint64_t n_cols = 4096;
    const std::string a = root + "/a", b = root + "/b", m = root + "/merged";
    write_folder(a, n_cols, rng);
    write_folder(b, n_cols, rng);
    FolderCSR f;
    EXPECT(read_dosemat_folder(a, n_cols, true, f) == FD_OK && f.has_fluence_map, "valid folder");
    const char* ins[2] = {a.c_str(), b.c_str()};
    EXPECT(fd_dosemat_merge_folders(ins, 2, m.c_str()) == FD_OK, "merge");
    FolderCSR g;
    EXPECT(read_dosemat_folder(m, n_cols, true, g) == FD_OK && g.rows.size() == 4, "merged folder");
    // corruptions, each on a fresh copy of a valid folder; all must give FD_ERR_IO
    // (the merger checks indices only: it has no voxel count to check columns against)
    auto corrupt = [&](const std::string& name, const std::string& file, auto mutate, bool merge_fails = true) {
        const std::string d = root + "/" + name;
        write_folder(d, n_cols, rng);
        std::vector<char> bytes;
        {
            std::ifstream in(d + "/" + file, std::ios::binary);
            bytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
        }
        mutate(bytes);
        std::ofstream(d + "/" + file, std::ios::binary | std::ios::trunc).write(bytes.data(), (std::streamsize)bytes.size());
        FolderCSR h;
        EXPECT(read_dosemat_folder(d, n_cols, true, h) == FD_ERR_IO, name.c_str());
        const char* one[1] = {d.c_str()};
        EXPECT((fd_dosemat_merge_folders(one, 1, (d + "_m").c_str()) == FD_ERR_IO) == merge_fails,
               (name + " merge").c_str());
    };
    auto set_u64 = [](std::vector<char>& v, size_t i, uint64_t x) {
        if (v.size() >= 8 * (i + 1)) std::memcpy(v.data() + 8 * i, &x, 8);
    };
    corrupt("truncated_columns", "columnsBuffer.bin", [](std::vector<char>& v) { v.resize(v.size() - 8); });
    corrupt("odd_offsets", "offsetsBuffer.bin", [](std::vector<char>& v) { v.resize(v.size() - 3); });
    corrupt("decreasing", "offsetsBuffer.bin", [&](std::vector<char>& v) { set_u64(v, 2, 0); set_u64(v, 1, 5); });
    corrupt("nonzero_start", "offsetsBuffer.bin", [&](std::vector<char>& v) { set_u64(v, 0, 1); });
    corrupt("huge_index", "OffsetBufferIdx.bin", [&](std::vector<char>& v) { set_u64(v, 1, ~0ull - 1); });
    corrupt("huge_rows", "numRowsPerMat.bin", [&](std::vector<char>& v) { set_u64(v, 0, ~0ull); });
    corrupt("huge_nnz", "NonZeroElements.bin", [&](std::vector<char>& v) { set_u64(v, 0, ~0ull); });
    corrupt("column_range", "columnsBuffer.bin", [&](std::vector<char>& v) { set_u64(v, 1, 4096); }, false);
    corrupt("empty_values", "valuesBuffer.bin", [](std::vector<char>& v) { v.clear(); });
    // a missing file and a fluence map in one input only
    const std::string c = root + "/c";
    write_folder(c, n_cols, rng);
    std::remove((c + "/valuesBuffer.bin").c_str());
    FolderCSR h;
    EXPECT(read_dosemat_folder(c, n_cols, true, h) == FD_ERR_IO, "missing file");
    const std::string d = root + "/d";
    write_folder(d, n_cols, rng);
    std::remove((d + "/fluenceMap.bin").c_str());
    const char* mix[2] = {a.c_str(), d.c_str()};
    EXPECT(fd_dosemat_merge_folders(mix, 2, (root + "/mix").c_str()) == FD_ERR_IO, "fluence map in one input");
    EXPECT(fd_dosemat_merge_folders(nullptr, 1, m.c_str()) == FD_ERR_ARG, "null inputs");
    std::printf("folders: valid, merged and 11 corrupted sets checked\n");
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch dir>\n", argv[0]);
        return 2;
    }
    std::mt19937_64 rng(1234);
    geometry_fuzz(rng, 600);
    folder_checks(argv[1], rng);
    if (failures) std::fprintf(stderr, "%d checks failed\n", failures);
    return failures ? 1 : 0;
}
