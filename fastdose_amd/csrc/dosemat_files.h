// dosemat_files.h -- host-side reading and checking of doseMatFolder file sets (dosemat_files.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fd_internal.h"

namespace fdi {

// One doseMatFolder as read from disk, checked: per-matrix tables, the raw offsets / columns /
// values buffers, and fluenceMap.bin when asked for and present.
struct FolderCSR {
    std::vector<uint64_t> rows, nnz, off_off, nnz_off;
    std::vector<char> off_b, col_b, val_b, fluence_b;
    uint64_t n_off = 0, n_nnz = 0;
    bool has_fluence_map = false;
};

// Reads the folder and checks that every index it holds stays in range: table sizes agree, each
// matrix's local offsets lie inside offsetsBuffer, start at 0, never decrease and end at its
// nonzero count, the counts add up to the columns buffer, and (n_cols > 0) every column is below
// n_cols.  FD_ERR_IO with a message otherwise; nothing on the device is touched.
int read_dosemat_folder(const std::string& dir, uint64_t n_cols, bool want_fluence_map, FolderCSR& out);

}  // namespace fdi
