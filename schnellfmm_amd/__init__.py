"""MI355X-native streaming LFMM2 hot path (schnelLFMM drop-in) on HIP/gfx950.

The numerics live in libslfmm.so (C ABI in include/slfmm.h); this package
holds its Python binding and a host-side mirror of the reference API.
"""
from .api import (CENTER_ONLY, EIGENSTRAT, PREC_FP64, PREC_I8, BedFile, Context, Lfmm2Config,  # noqa: F401
                  Lfmm2Problem, NmfEstResult, NmfParams, TestResults, center_covariates, estimate_factors,
                  estimate_factors_streaming, fit_lfmm2, fit_lfmm2_batch, fit_lfmm2_nmf, nmf_impute_estimation,
                  precompute, safe_inv,
                  subset_indices, test_associations, test_associations_fused)
from .output import BimRecords, parse_output_format, write_results  # noqa: F401
from .samples import FamRecord, align_est_bed_samples, load_covariates, read_fam  # noqa: F401
