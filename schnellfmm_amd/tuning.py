"""Launch-shape table of the HIP library (slfmm_tuning).

The library picks some tile and workgroup-count shapes per problem from a
small table: how many 32-SNP tiles a wave of the int8 SNP-side kernels
carries (1 or 2, by the work of the grid), the workgroup target of the int8 RSVD
forward pass, the number of partial blocks of the tall-skinny Gram, the cost
model of the batched-fit schedule and the default stream count of batched
fits.  ``python -m schnellfmm_amd.autotune`` measures the candidates on the
device and writes ``tuning.json`` next to this file; every Context created by
this package applies that table (SLFMM_TUNING names another file, "none"
keeps the library defaults).  The tile choice leaves every result bit for
bit unchanged; the workgroup targets only regroup FP64 partial sums
(rounding level).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")

# the library's built-in values (slfmm_ctx defaults)
DEFAULTS = {"snp_two_tiles_min_blocks": 0, "fwd_target_wg": 512, "gram_max_blocks": 1024,
            "batch_fit_ms": 1.4, "batch_gb_per_ms": 2.3, "batch_streams": 8}
LIBRARY_KEYS = ("snp_two_tiles_min_blocks", "fwd_target_wg", "gram_max_blocks", "batch_fit_ms", "batch_gb_per_ms")

_table: Optional[dict] = None


def table_path() -> Optional[str]:
    p = os.environ.get("SLFMM_TUNING")
    if p is None:
        return DEFAULT_PATH if os.path.exists(DEFAULT_PATH) else None
    return None if p.lower() == "none" else p


def validate(params: dict) -> dict:
    """Checked copy of a parameter set (unknown keys rejected, missing keys from DEFAULTS)."""
    bad = set(params) - set(DEFAULTS)
    if bad:
        raise ValueError(f"unknown tuning keys: {sorted(bad)}")
    out = dict(DEFAULTS)
    out.update(params)
    for k in ("snp_two_tiles_min_blocks", "fwd_target_wg", "gram_max_blocks", "batch_streams"):
        out[k] = int(out[k])
    if not (1 <= out["fwd_target_wg"] <= 65536 and 1 <= out["gram_max_blocks"] <= 65536):
        raise ValueError("fwd_target_wg and gram_max_blocks must be 1..65536")
    if not (0 <= out["snp_two_tiles_min_blocks"] < 2 ** 32):
        raise ValueError("snp_two_tiles_min_blocks must fit 32 bits")
    if not (1 <= out["batch_streams"] <= 16):
        raise ValueError("batch_streams must be 1..16")
    if not (out["batch_fit_ms"] >= 0.0 and out["batch_gb_per_ms"] > 0.0):
        raise ValueError("batch_fit_ms must be >= 0 and batch_gb_per_ms > 0")
    return out


def params() -> dict:
    """The parameters in force for new contexts (the table's, else the defaults)."""
    global _table
    if _table is None:
        path = table_path()
        if path is None:
            _table = dict(DEFAULTS)
        else:
            with open(path) as f:
                _table = validate(json.load(f)["params"])
    return dict(_table)


def reload() -> dict:
    global _table
    _table = None
    return params()


def to_c(p: dict):
    from ._lib import Tuning
    t = Tuning()
    for k in LIBRARY_KEYS:
        setattr(t, k, p[k])
    return t


def apply(ctx, p: Optional[dict] = None) -> None:
    """Set the launch shapes of a Context (the table's by default)."""
    from ._lib import lib
    t = to_c(validate(p) if p is not None else params())
    ctx.check(lib().slfmm_set_tuning(ctx.h, C.byref(t)))


def read(ctx) -> dict:
    from ._lib import Tuning, lib
    t = Tuning()
    ctx.check(lib().slfmm_get_tuning(ctx.h, C.byref(t)))
    return {k: getattr(t, k) for k in LIBRARY_KEYS}
