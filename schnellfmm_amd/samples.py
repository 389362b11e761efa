"""Sample alignment before the path: .fam records, the covariate file loader
(mean-imputing missing covariate values) and the estimation-panel alignment
(src/bed.rs:39-46 and :456-478; src/main.rs:291-571).

The kept .fam indices these return are the `sample_keep` argument of
BedFile.open (BedFile::subset_samples, src/bed.rs:155-159).
"""
from __future__ import annotations

import math
import os
import sys
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class FamRecord:
    """FamRecord (src/bed.rs:39-46)."""
    fid: str
    iid: str
    father: str
    mother: str
    sex: int
    pheno: str


def _lines(path: str) -> List[str]:
    """BufRead::lines: split on '\\n', one trailing '\\r' dropped, no empty last line."""
    with open(path, "rb") as f:
        text = f.read().decode()
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    return [ln[:-1] if ln.endswith("\r") else ln for ln in lines]


def read_fam(path: str) -> List[FamRecord]:
    """parse_fam (src/bed.rs:456-478)."""
    recs = []
    for i, line in enumerate(_lines(path)):
        f = line.split()
        if len(f) < 6:
            raise ValueError(f"FAM line {i + 1} has {len(f)} fields, expected 6")
        s = f[4][1:] if f[4].startswith("+") else f[4]
        if not (s.isascii() and s.isdigit()) or int(s) > 255:
            raise ValueError(f"FAM line {i + 1}: invalid sex '{f[4]}'")
        recs.append(FamRecord(f[0], f[1], f[2], f[3], int(s), f[5]))
    return recs


def fam_path_for(bed_path: str) -> str:
    return (bed_path[:-4] if bed_path.endswith(".bed") else bed_path) + ".fam"


def guess_delimiter(path: str) -> str:
    """src/main.rs:339-345: comma for a .csv extension, tab otherwise."""
    return "," if os.path.splitext(os.path.basename(path))[1] == ".csv" else "\t"


def _parse_f64(s: str) -> float:
    """str::parse::<f64> (Rust's grammar: optional sign, decimal digits with an
    optional point and exponent, or inf/infinity/nan in any case; no
    underscores or surrounding whitespace)."""
    if not s or not s.isascii() or "_" in s or s != s.strip():
        raise ValueError(s)
    return float(s)


def load_covariates(path: str, fam: Sequence[FamRecord], intersect: bool = False,
                    verbose: bool = False) -> Tuple[List[str], np.ndarray, List[int]]:
    """load_covariates (src/main.rs:347-571): header row (first cell names the
    sample column), one row per sample ID, numeric covariates; '.', 'NA' and
    empty cells are missing and take their column's observed mean.  Returns
    (covariate names, n_kept x d matrix in .fam order, kept .fam indices)."""
    delim = guess_delimiter(path)
    if verbose:
        print(f"  Detected {'CSV' if delim == ',' else 'TSV'} format (delimiter: {delim!r})", file=sys.stderr)
    try:
        lines = _lines(path)
    except OSError as e:
        raise OSError(f"Failed to open covariate file: {path}") from e
    if not lines:
        raise ValueError("Covariate file is empty")
    header = lines[0].split(delim)
    if len(header) < 2:
        raise ValueError(f"Covariate file header has only {len(header)} field(s) - expected at least a sample ID "
                         f"column and one covariate column (delimiter: {delim!r})")
    names = [h.strip() for h in header[1:]]
    d = len(names)
    data = {}
    order = []
    for i, line in enumerate(lines[1:]):
        if not line.strip():
            continue
        fields = line.split(delim)
        if len(fields) != d + 1:
            raise ValueError(f"Line {i + 2} has {len(fields)} fields, expected {d + 1} "
                             f"(1 sample ID + {d} covariates)")
        sid = fields[0].strip()
        vals = []
        for j, cell in enumerate(fields[1:]):
            t = cell.strip()
            if t == "" or t == "." or t.lower() == "na":
                vals.append(math.nan)
                continue
            try:
                v = _parse_f64(t)
            except ValueError:
                raise ValueError(f"Failed to parse value '{t}' for covariate '{names[j]}' on line {i + 2} "
                                 f"(sample '{sid}')") from None
            if math.isinf(v):
                raise ValueError(f"Infinite value ({'inf' if v > 0 else '-inf'}) for covariate '{names[j]}' on "
                                 f"line {i + 2} (sample '{sid}'): Inf values will propagate through all matrix "
                                 f"operations")
            vals.append(v)
        if sid in data:
            raise ValueError(f"Duplicate sample ID '{sid}' in covariate file (line {i + 2})")
        data[sid] = vals
        order.append(sid)
    if not data:
        raise ValueError("Covariate file has no data rows")
    print(f"  Read {len(data)} samples from covariate file", file=sys.stderr)
    n = len(fam)
    kept, rows = [], []
    for r, rec in enumerate(fam):
        if rec.iid in data:
            kept.append(r)
            rows.append(data[rec.iid])
        elif not intersect:
            raise ValueError(
                f"Sample '{rec.iid}' (FID='{rec.fid}') from .fam file not found in covariate file.\n"
                f".fam has {n} samples, covariate file has {len(data)}.\n"
                f"First few covariate sample IDs: [{', '.join(order[:5])}]\n"
                f"Hint: use --intersect-samples to use only samples present in both files.")
    matched = len(kept)
    if matched == 0:
        raise ValueError(
            f"No samples in common between .fam file ({n} samples) and covariate file ({len(data)} samples).\n"
            f"First few .fam IIDs: [{', '.join(r.iid for r in fam[:5])}]\n"
            f"First few covariate IDs: [{', '.join(order[:5])}]")
    if len(data) > matched:
        print(f"  Warning: {len(data) - matched} sample(s) in covariate file not present in .fam file (ignored)",
              file=sys.stderr)
    if n > matched:
        print(f"  Warning: {n - matched} sample(s) in .fam file not present in covariate file "
              f"(dropped with --intersect-samples)", file=sys.stderr)
    print(f"  Matched {matched} samples to .fam order", file=sys.stderr)
    x = np.array(rows, dtype=np.float64).reshape(matched, d)
    for j in range(d):
        col = x[:, j]
        miss = np.isnan(col)
        if not miss.any():
            continue
        s, cnt = 0.0, 0
        for v in col:                     # sequential f64 sum in row order
            if not math.isnan(v):
                s += v
                cnt += 1
        mean = s / cnt if cnt else 0.0
        col[miss] = mean
        print(f"  Warning: covariate '{names[j]}' has {int(miss.sum())} missing value(s) ('.', 'NA', or empty), "
              f"imputed to column mean ({mean:.6f})", file=sys.stderr)
    return names, x, kept


def align_est_bed_samples(est_fam: Sequence[FamRecord], main_fam: Sequence[FamRecord]) -> Optional[List[int]]:
    """align_est_bed_samples (src/main.rs:291-334): the estimation panel's
    physical sample indices in the main panel's order, or None when its
    samples already are the main panel's, in order."""
    pos = {}
    for i, r in enumerate(est_fam):
        pos[r.iid] = i                    # a later duplicate IID wins, as in the reference's map
    idx, missing = [], []
    for r in main_fam:
        if r.iid in pos:
            idx.append(pos[r.iid])
        else:
            missing.append(r.iid)
    if missing:
        more = ", ..." if len(missing) > 5 else ""
        raise ValueError(f"est-bed is missing {len(missing)} sample(s) that are in the main BED file: "
                         f"[{', '.join(missing[:5])}]{more}")
    if len(idx) == len(est_fam) and all(i == j for i, j in enumerate(idx)):
        return None
    print(f"  Reordering/subsetting est-bed samples to match main BED ({len(est_fam)} → {len(idx)} samples)",
          file=sys.stderr)
    return idx
