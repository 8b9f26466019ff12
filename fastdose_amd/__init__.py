"""fastdose_amd -- MI355X-native CCCS beamlet dose calculation (FastDoseLib's dose path).

Host API in :mod:`fastdose_amd.host`; C-ABI bindings in :mod:`fastdose_amd._lib`; synthetic
benchmark inputs in :mod:`fastdose_amd.synth`; build with ``python -m fastdose_amd.build``.
"""
from ._lib import (Beam, Bundle, CSR, FastDoseError, Kernel, Spectrum, Volume, FD_OPT_EXACT_G, FD_OPT_RPL_FP64,
                   FD_OPT_RPL_FP64_ALL, FD_OPT_TIMING, FD_OPT_GRAPH, FD_OPT_IEEE_DIV, FD_OPT_TABLES, FD_OPT_NO_TABLES,
                   FD_ERR_UNSUPPORTED, FD_VAL_BF16, FD_VAL_F32, FD_VAL_FP16, LIB_PATH, load)
from .host import (Context, DoseMat, Server, Spmm, beam_array, bundle_beamlets, beamlet_flag_init, beamlet_init, beamlets_init, calc_range,
                   dense_to_csr, dose_mat_construction, merge_dosemat_folders, ptv_isocenter, read_dosemat_folder, read_kernel,
                   read_spectrum, write_fluence_map)

__all__ = [
    "Beam", "Bundle", "CSR", "Context", "DoseMat", "Server", "Spmm", "FD_VAL_F32", "FD_VAL_BF16", "FD_VAL_FP16", "FastDoseError", "Kernel", "Spectrum", "Volume",
    "FD_OPT_EXACT_G", "FD_OPT_RPL_FP64", "FD_OPT_RPL_FP64_ALL", "FD_OPT_TIMING", "FD_OPT_GRAPH", "FD_OPT_IEEE_DIV", "FD_OPT_TABLES", "FD_OPT_NO_TABLES", "FD_ERR_UNSUPPORTED",
    "LIB_PATH", "load", "beam_array", "bundle_beamlets", "beamlet_flag_init", "beamlet_init",
    "beamlets_init", "calc_range", "dense_to_csr", "dose_mat_construction", "merge_dosemat_folders", "ptv_isocenter",
    "read_dosemat_folder", "read_kernel", "read_spectrum", "write_fluence_map",
]
