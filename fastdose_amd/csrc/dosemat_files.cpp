// dosemat_files.cpp -- host-side reading, checking and merging of doseMatFolder file sets.
//
// The on-disk layout is the one MatCSREnsemble::tofile writes (IMRTOpt/src/IMRTSparseMatEns.cu:
// 168-239) and scripts/readDose.m:14-53 reads: per-matrix u64 tables (NonZeroElements,
// numRowsPerMat, OffsetBufferIdx, ...), local u64 row offsets of every matrix back to back, u64
// voxel columns and f32 values, plus fluenceMap.bin (beamletFlagSave, IMRTBeamBundle.cu:572-601)
// in the same folder.
#include <cstring>
#include <fstream>
#include <string>
#include <vector>
#include <sys/stat.h>

#include "dosemat_files.h"

namespace fdi {

namespace {

bool read_all(const std::string& path, std::vector<char>& out, bool* missing = nullptr) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) {
        if (missing) *missing = true;
        return false;
    }
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    out.resize((size_t)n);
    if (n) f.read(out.data(), n);
    return (bool)f;
}

bool write_all(const std::string& path, const void* data, size_t bytes) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.is_open()) return false;
    if (bytes) f.write((const char*)data, (std::streamsize)bytes);
    return (bool)f;
}

uint64_t u64_at(const std::vector<char>& b, size_t i) {
    uint64_t v;
    std::memcpy(&v, b.data() + 8 * i, 8);
    return v;
}

}  // namespace

int read_dosemat_folder(const std::string& dir, uint64_t n_cols, bool want_fluence_map, FolderCSR& out) {
    std::vector<char> nnz_b, rows_b, oidx_b;
    if (!read_all(dir + "/NonZeroElements.bin", nnz_b) || !read_all(dir + "/numRowsPerMat.bin", rows_b) ||
        !read_all(dir + "/OffsetBufferIdx.bin", oidx_b) || !read_all(dir + "/offsetsBuffer.bin", out.off_b) ||
        !read_all(dir + "/columnsBuffer.bin", out.col_b) || !read_all(dir + "/valuesBuffer.bin", out.val_b)) {
        set_error("cannot read the dose-matrix folder " + dir);
        return FD_ERR_IO;
    }
    out.has_fluence_map = false;
    if (want_fluence_map) {
        bool missing = false;
        if (read_all(dir + "/fluenceMap.bin", out.fluence_b, &missing))
            out.has_fluence_map = true;
        else if (!missing) {
            set_error("cannot read " + dir + "/fluenceMap.bin");
            return FD_ERR_IO;
        }
    }
    auto bad = [&](const std::string& why) {
        set_error("dose-matrix folder " + dir + ": " + why);
        return FD_ERR_IO;
    };
    const size_t nm = nnz_b.size() / 8;
    if (nnz_b.size() % 8 || rows_b.size() != nnz_b.size() || oidx_b.size() != nnz_b.size() ||
        out.off_b.size() % 8 || out.col_b.size() % 8 || out.val_b.size() % 4 ||
        out.val_b.size() / 4 != out.col_b.size() / 8)
        return bad("file sizes do not match one another");
    out.n_off = out.off_b.size() / 8;
    out.n_nnz = out.col_b.size() / 8;
    out.rows.resize(nm);
    out.nnz.resize(nm);
    out.off_off.resize(nm);
    out.nnz_off.resize(nm);
    uint64_t acc = 0;
    for (size_t i = 0; i < nm; i++) {
        const uint64_t rows = u64_at(rows_b, i), nnz = u64_at(nnz_b, i), oo = u64_at(oidx_b, i);
        const std::string m = "matrix " + std::to_string(i);
        if (nnz > out.n_nnz - acc) return bad("NonZeroElements exceeds the columns buffer");
        if (rows >= out.n_off || oo > out.n_off - rows - 1) return bad(m + "'s offsets run past offsetsBuffer");
        // local offsets: start at 0, never decrease, end at the matrix's nonzero count
        uint64_t prev = u64_at(out.off_b, oo);
        if (prev != 0) return bad(m + "'s offsets do not start at 0");
        for (uint64_t r = 1; r <= rows; r++) {
            const uint64_t o = u64_at(out.off_b, oo + r);
            if (o < prev) return bad(m + "'s offsets decrease");
            prev = o;
        }
        if (prev != nnz) return bad(m + "'s last offset differs from its nonzero count");
        out.rows[i] = rows;
        out.nnz[i] = nnz;
        out.off_off[i] = oo;
        out.nnz_off[i] = acc;
        acc += nnz;
    }
    if (acc != out.n_nnz) return bad("NonZeroElements does not add up to the columns buffer");
    if (n_cols) {
        for (uint64_t e = 0; e < out.n_nnz; e++) {
            const uint64_t c = u64_at(out.col_b, e);
            if (c >= n_cols)
                return bad("column index " + std::to_string(c) + " beyond the " + std::to_string(n_cols) + " voxels");
        }
    }
    return FD_OK;
}

}  // namespace fdi

using namespace fdi;

extern "C" int fd_dosemat_merge_folders(const char* const* folders, int n_folders, const char* out_folder) {
    if (!folders || n_folders <= 0 || !out_folder) {
        set_error("merge needs at least one input folder and an output folder");
        return FD_ERR_ARG;
    }
    // read and check every input before anything is written
    std::vector<FolderCSR> in((size_t)n_folders);
    int with_map = 0;
    for (int k = 0; k < n_folders; k++) {
        if (!folders[k]) {
            set_error("null input folder");
            return FD_ERR_ARG;
        }
        const int r = read_dosemat_folder(folders[k], 0, true, in[k]);
        if (r != FD_OK) return r;
        with_map += in[k].has_fluence_map ? 1 : 0;
    }
    if (with_map != 0 && with_map != n_folders) {
        set_error("fluenceMap.bin is present in some input folders only");
        return FD_ERR_IO;
    }
    std::vector<uint64_t> nnz, rows, cumu_nnz, cumu_rows, oidx;
    std::vector<char> off, col, val, flu;
    uint64_t off_base = 0, a = 0, r = 0;
    for (const FolderCSR& f : in) {
        for (size_t i = 0; i < f.rows.size(); i++) {
            nnz.push_back(f.nnz[i]);
            rows.push_back(f.rows[i]);
            a += f.nnz[i];
            r += f.rows[i];
            cumu_nnz.push_back(a);
            cumu_rows.push_back(r);
            oidx.push_back(off_base + f.off_off[i]);
        }
        off_base += f.n_off;
        off.insert(off.end(), f.off_b.begin(), f.off_b.end());
        col.insert(col.end(), f.col_b.begin(), f.col_b.end());
        val.insert(val.end(), f.val_b.begin(), f.val_b.end());
        flu.insert(flu.end(), f.fluence_b.begin(), f.fluence_b.end());
    }
    const std::string dir(out_folder);
    mkdir(dir.c_str(), 0777);
    const std::pair<const std::vector<uint64_t>*, const char*> small[] = {
        {&nnz, "NonZeroElements"}, {&cumu_nnz, "CumuNonZeroElements"}, {&rows, "numRowsPerMat"},
        {&cumu_rows, "CumuNumRowsPerMat"}, {&oidx, "OffsetBufferIdx"}};
    for (const auto& f : small)
        if (!write_all(dir + "/" + f.second + ".bin", f.first->data(), f.first->size() * sizeof(uint64_t))) {
            set_error("cannot write " + dir + "/" + f.second + ".bin");
            return FD_ERR_IO;
        }
    if (!write_all(dir + "/offsetsBuffer.bin", off.data(), off.size()) ||
        !write_all(dir + "/columnsBuffer.bin", col.data(), col.size()) ||
        !write_all(dir + "/valuesBuffer.bin", val.data(), val.size()) ||
        (with_map && !write_all(dir + "/fluenceMap.bin", flu.data(), flu.size()))) {
        set_error("cannot write the merged buffers under " + dir);
        return FD_ERR_IO;
    }
    return FD_OK;
}
