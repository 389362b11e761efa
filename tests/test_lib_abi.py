"""CPU checks of the C-ABI library: it loads and exports every symbol that
include/slfmm.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "slfmm.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(slfmm_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("slfmm_create", "slfmm_fit_lfmm2", "slfmm_test_associations", "slfmm_estimate_factors",
                 "slfmm_decode", "slfmm_safe_inv", "slfmm_precompute", "slfmm_geno_load_bed"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from schnellfmm_amd._lib import LIB_PATH, SIGNATURES
    if not os.path.exists(LIB_PATH):
        pytest.fail("libslfmm.so is not built (run python -m schnellfmm_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(declared_symbols()) <= set(SIGNATURES)


def test_host_only_entry_points():
    import numpy as np
    from schnellfmm_amd import center_covariates
    import oracle as O
    x = np.random.default_rng(0).standard_normal((17, 3)) * 3 + 1
    for scale in (False, True):
        assert np.allclose(center_covariates(x, scale), O.center_covariates(x, scale), rtol=0, atol=1e-13)


def test_subset_indices_matches_reference_vectors():
    import json
    from schnellfmm_amd import subset_indices
    kat = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_kat.json")))
    for c in kat["subset_indices"]:
        assert list(subset_indices(c["spec"], c["n"])) == c["idx"]
