"""CPU restatement of schnelLFMM's output stage (test infrastructure only).

The checker for ``slfmm_write_tsv`` / ``slfmm_format_e6`` and the Parquet
writer of ``schnellfmm_amd.api``.  Follows:
  * parse_bim                  src/bed.rs:480-503
  * coalesce_output_tsv        src/testing.rs:740-801 (header :757-761, rows
                               :777-796, every number `{:.6e}`)
  * build_calibrated_schema    src/testing.rs:621-640
  * coalesce_output_parquet    src/testing.rs:818-926: worker w encodes chunks
                               w, w + W, w + 2W, ... (W = min(n_workers,
                               n_chunks), at least 1) and the worker files are
                               concatenated in worker order (:861-909).

Pinning (tests/test_output_oracle.py):
  * rust_e6 -- Rust's documented `{:e}` form (exponent without '+' or leading
    zeros, "NaN", "inf", signed zero) on hand-checked values, and the digit
    choice against exact rational arithmetic (nearest 7-significant-digit
    decimal, ties to even) on random and tie-constructed doubles.
  * results_tsv / read_bim -- column order and header against the reference's
    own TSV reader (tests/integration.rs:28-68: d = (cols - 6) / 3, p/beta/t
    triples per covariate, r2 triple last), bim validation messages against
    src/bed.rs:488-498.
  * parquet_row_order -- brute-force simulation of the worker loop.
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import List, Optional, Sequence, Tuple

import numpy as np


def rust_e6(x: float) -> str:
    """Rust's `format!("{:.6e}", x)` for an f64 (src/testing.rs:788, :795).

    Python's '%.6e' rounds the exact binary value correctly with ties to even,
    as core::fmt's exact mode does; only the exponent spelling differs."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "-inf" if x < 0 else "inf"
    mant, exp = ("%.6e" % x).split("e")
    return f"{mant}e{int(exp)}"


def e6_exact(x: float) -> str:
    """The same string from first principles: the nearest decimal with seven
    significant digits to the exact rational value of x, ties to even."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "-inf" if x < 0 else "inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0.000000e0"
    a = Fraction(abs(x))
    e = math.floor(math.log10(abs(x)))
    for _ in range(3):                       # make 10^e <= a < 10^(e+1) exactly
        if a >= Fraction(10) ** (e + 1):
            e += 1
        elif a < Fraction(10) ** e:
            e -= 1
    scaled = a * Fraction(10) ** (6 - e)     # in [1e6, 1e7)
    q, r = divmod(scaled.numerator, scaled.denominator)
    twice = 2 * r
    if twice > scaled.denominator or (twice == scaled.denominator and q % 2 == 1):
        q += 1
    if q == 10 ** 7:
        q, e = 10 ** 6, e + 1
    d = str(q)
    return f"{sign}{d[0]}.{d[1:]}e{e}"


def read_bim(path: str) -> List[Tuple[str, str, float, int, str, str]]:
    """parse_bim (src/bed.rs:480-503): (chrom, snp_id, cm, pos, a1, a2) per line;
    BufRead::lines splits on '\n' and drops one trailing '\r'."""
    with open(path, "rb") as f:
        text = f.read().decode()
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    recs = []
    for i, line in enumerate(lines):
        if line.endswith("\r"):
            line = line[:-1]
        fields = line.split()
        if len(fields) < 6:
            raise ValueError(f"BIM line {i + 1} has {len(fields)} fields, expected 6")
        try:
            cm = float(fields[2])
        except ValueError:
            raise ValueError(f"BIM line {i + 1}: invalid cM '{fields[2]}'")
        digits = fields[3][1:] if fields[3].startswith("+") else fields[3]
        if not (digits.isascii() and digits.isdigit()) or int(digits) >= 2 ** 64:
            raise ValueError(f"BIM line {i + 1}: invalid position '{fields[3]}'")
        recs.append((fields[0], fields[1], cm, int(digits), fields[4], fields[5]))
    return recs


def results_tsv(chrom: Sequence[str], pos: Sequence[int], snp_id: Sequence[str],
                cov_names: Optional[Sequence[str]], pvalue: np.ndarray, beta: np.ndarray,
                tstat: np.ndarray, r2: np.ndarray) -> str:
    """Text of coalesce_output_tsv (src/testing.rs:740-801) for per-SNP arrays
    pvalue/beta/tstat (p x d) and r2 (p x 3)."""
    p, d = np.asarray(tstat).reshape(len(chrom), -1).shape
    names = list(cov_names) if cov_names is not None else [f"cov_{j}" for j in range(d)]
    out = ["chr\tpos\tsnp_id"]
    for nm in names:
        out.append(f"\tp_{nm}\tbeta_{nm}\tt_{nm}")
    out.append("\tr2_cov\tr2_latent\tr2_resid\n")
    pv, be, tt, rr = (np.asarray(a, np.float64).reshape(p, -1) for a in (pvalue, beta, tstat, r2))
    for j in range(p):
        row = [f"{chrom[j]}\t{int(pos[j])}\t{snp_id[j]}"]
        for c in range(d):
            row.append(f"\t{rust_e6(pv[j, c])}\t{rust_e6(be[j, c])}\t{rust_e6(tt[j, c])}")
        row.append(f"\t{rust_e6(rr[j, 0])}\t{rust_e6(rr[j, 1])}\t{rust_e6(rr[j, 2])}\n")
        out.append("".join(row))
    return "".join(out)


def parse_results_tsv(text: str):
    """The reference's own reader of that table (tests/integration.rs:28-68)."""
    lines = text.split("\n")
    header = lines[0].split("\t")
    d = (len(header) - 3 - 3) // 3
    rows = [ln.split("\t") for ln in lines[1:] if ln]
    pv = np.array([[float(r[3 + 3 * j]) for j in range(d)] for r in rows]).reshape(len(rows), d)
    be = np.array([[float(r[4 + 3 * j]) for j in range(d)] for r in rows]).reshape(len(rows), d)
    tt = np.array([[float(r[5 + 3 * j]) for j in range(d)] for r in rows]).reshape(len(rows), d)
    r2 = np.array([[float(r[3 + 3 * d + k]) for k in range(3)] for r in rows]).reshape(len(rows), 3)
    return header, pv, be, tt, r2


def calibrated_columns(cov_names: Sequence[str]) -> List[Tuple[str, str]]:
    """build_calibrated_schema (src/testing.rs:621-640): (name, arrow type)."""
    cols = [("chr", "string"), ("pos", "int64"), ("snp_id", "string")]
    cols += [(f"p_{c}", "double") for c in cov_names]
    cols += [(f"beta_{c}", "double") for c in cov_names]
    cols += [(f"t_{c}", "double") for c in cov_names]
    cols += [("r2_cov", "double"), ("r2_latent", "double"), ("r2_resid", "double")]
    return cols


def parquet_row_order(p: int, chunk_size: int, n_workers: int) -> np.ndarray:
    """SNP index of each output row of coalesce_output_parquet
    (src/testing.rs:828, :861-888, :909): worker w writes chunks w, w+W, ...;
    the worker files are concatenated in worker order."""
    n_chunks = (p + chunk_size - 1) // chunk_size if p else 0
    w_count = max(min(n_workers, n_chunks), 1)
    order = []
    for w in range(w_count):
        seq = w
        while seq < n_chunks:
            order.extend(range(seq * chunk_size, min(p, (seq + 1) * chunk_size)))
            seq += w_count
    return np.asarray(order, dtype=np.int64)
