"""densecrf_ho_amd: dense-CRF inference (permutohedral filter, QP, LP, Pn
superpixel terms) on AMD MI355X (gfx950) through hand-written HIP kernels.

The compute lives in ``libdensecrf_ho_amd.so`` (C ABI, ``include/densecrf_ho_amd.h``);
this package is a thin host-side mirror of the reference's C++ interface.
Importing it does not require a GPU; calling into the engine does, and fails
with an error when the library or the device is missing.
"""

from .api import (DenseCRF, DenseCRF2D, DeviceArray, LP_inf_params, Permutohedral, PottsCompatibility,  # noqa: F401
                  DiagonalCompatibility, MatrixCompatibility, DenseCRFBatch,
                  NO_NORMALIZATION, NORMALIZE_BEFORE, NORMALIZE_AFTER, NORMALIZE_SYMMETRIC, EngineError,
                  profile_enable, profile_filter, profile_report, synchronize, autotune_enable, autotune_report,
                  set_reference_path, reference_path,
                  load_unary, load_image, save_map, superpixel_terms, less_confident_pixels,
                  PairwisePotential, CONST_KERNEL, DIAG_KERNEL, FULL_KERNEL, UnaryEnergy, ConstUnaryEnergy,
                  LogisticUnaryEnergy)
from ._lib import lib_path, load  # noqa: F401

__all__ = ["DenseCRF", "DenseCRF2D", "DenseCRFBatch", "DeviceArray", "profile_enable", "profile_filter", "profile_report", "synchronize", "autotune_enable", "autotune_report", "set_reference_path", "reference_path", "LP_inf_params", "Permutohedral", "PottsCompatibility", "DiagonalCompatibility",
           "MatrixCompatibility", "EngineError",
           "NO_NORMALIZATION", "NORMALIZE_BEFORE", "NORMALIZE_AFTER", "NORMALIZE_SYMMETRIC", "load", "lib_path",
           "load_unary", "load_image", "save_map", "superpixel_terms", "less_confident_pixels",
           "PairwisePotential", "CONST_KERNEL", "DIAG_KERNEL", "FULL_KERNEL", "UnaryEnergy", "ConstUnaryEnergy",
           "LogisticUnaryEnergy"]
