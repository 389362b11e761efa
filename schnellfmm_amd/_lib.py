"""ctypes binding of libslfmm.so (the C ABI declared in include/slfmm.h).

The product path has no CPU fallback: if the HIP library is missing this
module raises at import of any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SLFMM_LIBRARY may name another build of the same library (e.g. one built
# with different tuning macros); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("SLFMM_LIBRARY") or os.path.join(HERE, "libslfmm.so")

u64, i64, i32, u32, dbl = C.c_uint64, C.c_int64, C.c_int32, C.c_uint32, C.c_double
vp = C.c_void_p


class Config(C.Structure):
    _fields_ = [("k", u64), ("lam", dbl), ("chunk_size", u64), ("oversampling", u64),
                ("n_power_iter", u64), ("seed", u64), ("norm", i32), ("scale_cov", i32),
                ("impute", i32), ("precision", i32)]


class NmfConfig(C.Structure):
    _fields_ = [("k", u64), ("n_iter", u64), ("eps", dbl), ("cv_rate", dbl),
                ("chunk_size", u64), ("seed", u64)]


class AssocOut(C.Structure):
    _fields_ = [("beta", vp), ("tstat", vp), ("r2", vp), ("pvalue", vp), ("hist", vp),
                ("gif_per_trait", vp), ("gif", vp)]


class SynthSpec(C.Structure):
    """slfmm_synth_spec: options of the generalised synthetic panel."""
    _fields_ = [("n_pops", u32), ("fst_lo", dbl), ("fst_hi", dbl), ("maf_beta", i32), ("maf_a", dbl),
                ("maf_b", dbl), ("maf_min", dbl), ("ancestry", vp), ("het_rate", dbl), ("miss_rate", dbl),
                ("sample_miss", vp), ("snp_miss_a", dbl), ("snp_miss_b", dbl), ("miss_scale", dbl),
                ("seed", u64)]


class Tuning(C.Structure):
    """slfmm_tuning: launch shapes (tile and workgroup-count choices)."""
    _fields_ = [("snp_two_tiles_min_blocks", u32), ("fwd_target_wg", u32), ("gram_max_blocks", u32),
                ("batch_fit_ms", dbl), ("batch_gb_per_ms", dbl)]


class Problem(C.Structure):
    """slfmm_problem: the arguments of slfmm_fit_lfmm2 for one problem of a batch."""
    _fields_ = [("est", vp), ("subset_idx", vp), ("n_idx", u64), ("g", vp), ("x", vp), ("d", u64),
                ("cfg", Config), ("b0", vp), ("u_hat", vp), ("out", AssocOut)]


# name -> (restype, argtypes)
SIGNATURES = {
    "slfmm_config_default": (None, [vp]),
    "slfmm_nmf_config_default": (None, [vp]),
    "slfmm_version": (C.c_char_p, []),
    "slfmm_create": (C.c_int, [C.c_int, vp]),
    "slfmm_destroy": (None, [vp]),
    "slfmm_last_error": (C.c_char_p, [vp]),
    "slfmm_synchronize": (C.c_int, [vp]),
    "slfmm_bed_open": (C.c_int, [vp, C.c_char_p, vp, vp]),
    "slfmm_geno_load_bed": (C.c_int, [vp, C.c_char_p, vp, u64, vp]),
    "slfmm_geno_upload": (C.c_int, [vp, vp, u64, u64, vp, u64, vp]),
    "slfmm_geno_wrap_device": (C.c_int, [vp, vp, u64, u64, u64, vp]),
    "slfmm_geno_open_bed_host": (C.c_int, [vp, C.c_char_p, u64, vp]),
    "slfmm_geno_wrap_host": (C.c_int, [vp, vp, u64, u64, u64, vp]),
    "slfmm_geno_read_packed": (C.c_int, [vp, vp, u64, u64, vp]),
    "slfmm_write_tsv": (C.c_int, [vp, C.c_char_p, C.c_char_p, vp, vp, vp, u64, u64, vp, vp, vp, vp, vp,
                                  C.c_uint32, C.c_int]),
    "slfmm_format_e6": (C.c_int, [vp, u64, vp, u64]),
    "slfmm_geno_free": (None, [vp]),
    "slfmm_geno_info": (C.c_int, [vp, vp, vp, vp]),
    "slfmm_decode_stats": (C.c_int, [vp, vp, vp, u64, vp, vp, vp]),
    "slfmm_decode": (C.c_int, [vp, vp, vp, u64, i32, vp, vp]),
    "slfmm_decode_raw": (C.c_int, [vp, vp, vp, u64, vp]),
    "slfmm_center_covariates": (C.c_int, [vp, u64, u64, i32, vp]),
    "slfmm_safe_inv": (C.c_int, [vp, vp, u64, vp, vp]),
    "slfmm_precompute": (C.c_int, [vp, vp, u64, u64, dbl, vp, vp, vp, vp]),
    "slfmm_estimate_factors": (C.c_int, [vp, vp, vp, u64, vp, u64, vp, vp, vp, u32]),
    "slfmm_test_associations": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, u32]),
    "slfmm_test_associations_noalloc": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, u32]),
    "slfmm_fit_lfmm2": (C.c_int, [vp, vp, vp, u64, vp, vp, u64, vp, vp, vp, vp, u32]),
    "slfmm_fit_lfmm2_batch": (C.c_int, [vp, vp, u64, C.c_int, u32]),
    "slfmm_nmf_estimate": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp, u32]),
    "slfmm_fit_lfmm2_nmf": (C.c_int, [vp, vp, vp, u64, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp, vp, u32]),
    "slfmm_debug_mfma_i8": (C.c_int, [vp, vp, vp, vp]),
    "slfmm_debug_i8_layout": (C.c_int, [i32, i32, i32, vp, vp, u64]),
    "slfmm_debug_mfma_rate": (C.c_int, [vp, i32, u64, vp]),
    "slfmm_set_tuning": (C.c_int, [vp, vp]),
    "slfmm_get_tuning": (C.c_int, [vp, vp]),
    "slfmm_profile_enable": (C.c_int, [vp, i32]),
    "slfmm_profile_read": (C.c_int, [vp, C.c_char_p, vp, vp, i32]),
    "slfmm_graph_stats": (C.c_int, [vp, vp, vp]),
    "slfmm_synth_bn": (C.c_int, [vp, u64, u64, u32, dbl, dbl, dbl, dbl, dbl, u64, vp]),
    "slfmm_synth_bn_ex": (C.c_int, [vp, u64, u64, vp, vp]),
}

_lib = None


def lib():
    """Load libslfmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run python -m schnellfmm_amd.build "
                               "(there is no CPU fallback for the LFMM2 hot path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SlfmmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def check(ctx, rc: int) -> None:
    if rc != 0:
        msg = lib().slfmm_last_error(ctx).decode(errors="replace")   # NULL ctx: the thread's last error
        raise SlfmmError(rc, msg)
