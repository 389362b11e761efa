"""Command-line front end: the reference CLI's flags (src/main.rs:12-107) and
its flow (src/main.rs:114-289) over this library.

    python -m schnellfmm_amd -b panel.bed -c covariates.tsv -k 20 -o out

.bed + covariate file -> sample matching -> optional estimation panel
(--est-bed, aligned by IID) or SNP thinning (--est-rate) -> fit on the GPU
(mean or NMF imputation) -> <out>.tsv / <out>.parquet and <out>.summary.txt.
Progress lines go to stderr with the reference's wording; a failure prints
"Error: <message>" and exits with status 1, as an anyhow error from the
reference's main does.

Flags beyond the reference's: --precision (fp64, the reference's arithmetic,
or mixed: int8 MFMA genotype streams with FP64 solves, max |dlog10 p| <= 1e-3
against FP64) and --host-panel (keep the .bed in host memory and stream it to
the device in chunks, for panels larger than HBM).
"""
from __future__ import annotations

import argparse
import math
import os
import sys
from decimal import Decimal
from typing import List, Optional, Sequence

import numpy as np

from . import samples as SM


def _eprint(*a) -> None:
    print(*a, file=sys.stderr, flush=True)


def rust_display(x: float) -> str:
    """Rust's `{}` for f64: the shortest round-trip digits in positional
    form, no exponent, no trailing ".0" (1e-5 -> "0.00001", 2.0 -> "2")."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    t = format(Decimal(repr(float(x))), "f")
    if "." in t:
        t = t.rstrip("0").rstrip(".")
    return t


def rust_exp1(x: float) -> str:
    """Rust's `{:.1e}` for a finite f64 (1e-5 -> "1.0e-5")."""
    m, e = f"{x:.1e}".split("e")
    return f"{m}e{int(e)}"


def _default_threads() -> int:
    """default_threads (src/main.rs:109-113): available parallelism minus one."""
    return max((os.cpu_count() or 2) - 1, 0)


def _norm(s: str) -> str:
    if s not in ("center-only", "eigenstrat"):
        raise argparse.ArgumentTypeError(f"invalid value '{s}' (possible values: center-only, eigenstrat)")
    return s


def _format(s: str) -> str:
    from .output import parse_output_format
    try:
        return parse_output_format(s)
    except ValueError as e:
        raise argparse.ArgumentTypeError(str(e))


def build_parser() -> argparse.ArgumentParser:
    """The Cli struct (src/main.rs:12-107): same flags, short forms and defaults."""
    ap = argparse.ArgumentParser(prog="lfmm2", description="Latent Factor Mixed Model v2 - GWAS with latent confounders")
    ap.add_argument("-b", "--bed", required=True, help="PLINK .bed file (with .bim/.fam)")
    ap.add_argument("-c", "--cov", required=True,
                    help="Covariate/phenotype file (CSV or TSV, auto-detected from extension). "
                         "First row = header, first column = sample ID.")
    ap.add_argument("-k", type=int, required=True, help="Number of latent factors")
    ap.add_argument("-l", "--lambda", dest="lam", type=float, default=1e-5, help="Ridge penalty")
    ap.add_argument("-o", "--out", default="lfmm2_out", help="Output prefix")
    ap.add_argument("-f", "--format", type=_format, default="tsv",
                    help="Output format: tsv (default) or parquet (zstd-compressed, DuckDB-ready)")
    ap.add_argument("-t", "--threads", type=int, default=_default_threads(),
                    help="Worker threads (0 is treated as 1): formatting threads of the results "
                         "writer and the Parquet row order")
    ap.add_argument("--chunk-size", type=int, default=10_000, help="SNPs per chunk")
    ap.add_argument("--oversampling", type=int, default=10, help="RSVD oversampling")
    ap.add_argument("--power-iter", type=int, default=2, help="RSVD power iterations")
    ap.add_argument("--seed", type=int, default=42, help="RNG seed")
    ap.add_argument("--est-bed", default=None, help="Separate .bed for estimation (LD-pruned)")
    ap.add_argument("--est-rate", type=float, default=None, help="Thin estimation SNPs at this rate")
    ap.add_argument("--intersect-samples", action="store_true",
                    help="Use only samples present in both the .fam and covariate files. "
                         "Without this flag, a missing sample in the covariate file is an error.")
    ap.add_argument("-v", "--verbose", action="store_true", help="Verbose progress output")
    ap.add_argument("--norm", type=_norm, default="eigenstrat", help="SNP normalization mode")
    ap.add_argument("--scale-cov", action="store_true",
                    help="Scale covariate columns to unit variance (in addition to centering).")
    imp = ap.add_mutually_exclusive_group()
    imp.add_argument("--mean-impute", action="store_true",
                     help="Use mean imputation for missing genotypes (default).")
    imp.add_argument("--nmf-impute", action="store_true", help="Use NMF low-rank imputation for missing genotypes.")
    ap.add_argument("--nmf-k", type=int, default=None, help="NMF rank (default: same as -k).")
    ap.add_argument("--nmf-iter", type=int, default=10, help="NMF multiplicative update iterations (default: 10).")
    ap.add_argument("--nmf-cv-rate", type=float, default=0.0005,
                    help="Fraction of genotypes held out per NMF iteration for CV (default: 0.0005).")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="fp64 (the reference's arithmetic) or mixed (int8 MFMA genotype streams, FP64 solves)")
    ap.add_argument("--host-panel", action="store_true",
                    help="Keep the .bed in host memory and stream it to the GPU in chunks (panels larger than HBM)")
    return ap


def _bed_dims(path: str):
    """BedFile::open's validation (src/bed.rs:78-135) on the host: magic,
    exact size, .fam/.bim line counts."""
    import ctypes as C
    from ._lib import check, lib
    n, p = C.c_uint64(), C.c_uint64()
    check(None, lib().slfmm_bed_open(None, path.encode(), C.byref(n), C.byref(p)))
    return int(n.value), int(p.value)


def _rate_indices(rate: float, n_snps: int) -> np.ndarray:
    """subset_indices for SubsetSpec::Rate (src/parallel.rs:113-122): every
    ceil(1/rate)-th SNP from 0; a step beyond the panel keeps SNP 0 alone."""
    inv = 1.0 / rate if rate > 0 else math.inf
    step = math.ceil(inv) if math.isfinite(inv) else n_snps + 1
    return np.arange(0, n_snps, max(step, 1), dtype=np.int64)


def write_summary(path: str, res, n_samples: int, n_snps: int, k: int, cov_names: Sequence[str], lam: float,
                  threads: int, nmf: bool) -> None:
    """The summary file of src/main.rs:248-283."""
    lines = [f"GIF: {res.gif:.6f}", f"n_samples: {n_samples}", f"n_snps: {n_snps}", f"K: {k}",
             f"d: {len(cov_names)}", f"covariates: {', '.join(cov_names)}", f"lambda: {rust_display(lam)}",
             f"threads: {threads}", f"imputation: {'nmf' if nmf else 'mean'}"]
    if res.nmf_cv is not None:
        lines.append(f"nmf_iterations: {len(res.nmf_cv)}")
        lines += [f"nmf_cv_mae_iter_{i + 1}: {v:.6f}" for i, v in enumerate(res.nmf_cv)]
        if res.nmf_cv_mean is not None:
            lines.append(f"mean_impute_cv_mae: {res.nmf_cv_mean:.6f}")
    if res.gwas_cv is not None:
        mae, mean_mae, cnt = res.gwas_cv
        lines += [f"nmf_gwas_cv_mae: {mae:.6f}", f"nmf_gwas_cv_mean_mae: {mean_mae:.6f}", f"nmf_gwas_cv_count: {cnt}"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def run(a: argparse.Namespace):
    """main (src/main.rs:114-289).  Returns the fit's TestResults."""
    from . import api as A
    from .output import write_results

    _eprint(f"Loading BED file: {a.bed}")
    n_phys, n_snps = _bed_dims(a.bed)
    _eprint(f"  {n_phys} samples x {n_snps} SNPs")
    fam = SM.read_fam(SM.fam_path_for(a.bed))

    _eprint(f"Loading covariates from: {a.cov}")
    cov_names, x, kept = SM.load_covariates(a.cov, fam, a.intersect_samples, a.verbose)
    d = x.shape[1]
    keep: Optional[List[int]] = kept if len(kept) < len(fam) else None
    main_fam = fam if keep is None else [fam[i] for i in keep]
    n = len(main_fam)
    _eprint(f"  {n} samples x {d} covariates: [{', '.join(cov_names)}]")

    if a.est_bed is not None and a.est_rate is not None:
        raise ValueError("Cannot specify both --est-bed and --est-rate")
    est_keep = None
    if a.est_bed is not None:
        _eprint(f"Loading estimation BED file: {a.est_bed}")
        _, est_snps = _bed_dims(a.est_bed)
        est_keep = SM.align_est_bed_samples(SM.read_fam(SM.fam_path_for(a.est_bed)), main_fam)
        _eprint(f"  {n} samples x {est_snps} SNPs (for factor estimation)")
    subset = None
    if a.est_rate is not None:
        if not (0.0 <= a.est_rate <= 1.0):
            raise ValueError(f"--est-rate must be in (0.0, 1.0], got {rust_display(a.est_rate)}")
        _eprint(f"  Estimation subset: thinning at rate {rust_display(a.est_rate)}")

    nmf = bool(a.nmf_impute)
    prec = A.PREC_I8 if a.precision == "mixed" else A.PREC_FP64
    cfg = A.Lfmm2Config(k=a.k, lam=a.lam, chunk_size=a.chunk_size, oversampling=a.oversampling,
                        n_power_iter=a.power_iter, seed=a.seed, n_workers=a.threads,
                        progress=sys.stderr.isatty(),
                        norm=A.EIGENSTRAT if a.norm == "eigenstrat" else A.CENTER_ONLY,
                        scale_cov=a.scale_cov, impute="nmf" if nmf else "mean", precision=prec)
    _eprint(f"Running LFMM2: K={cfg.k}, lambda={rust_exp1(cfg.lam)}, chunk_size={cfg.chunk_size}, "
            f"threads={cfg.n_workers}, oversampling={cfg.oversampling}, power_iter={cfg.n_power_iter}")
    ext = "parquet" if a.format == "parquet" else "tsv"
    results_path = f"{a.out}.{ext}"
    summary_path = f"{a.out}.summary.txt"
    _eprint(f"Output format: {'Parquet' if ext == 'parquet' else 'TSV'}")

    if a.host_panel:
        if keep is not None:
            raise ValueError("--host-panel streams the .bed as it is on disk; it cannot drop or reorder samples")
        if nmf:
            raise ValueError("NMF imputation needs the test panel in device memory (drop --host-panel)")
        y_full = A.BedFile.open_host(a.bed)
    else:
        y_full = A.BedFile.open(a.bed, sample_keep=keep)
    y_est = A.BedFile.open(a.est_bed, sample_keep=est_keep) if a.est_bed is not None else y_full
    if a.est_rate is not None:
        subset = _rate_indices(a.est_rate, y_est.n_snps)
    if nmf:
        ncfg = A.NmfParams(k=a.nmf_k if a.nmf_k is not None else a.k, n_iter=a.nmf_iter, eps=1e-16,
                           cv_rate=a.nmf_cv_rate, chunk_size=a.chunk_size, seed=a.seed)
        # mixed precision also runs the NMF update passes on the int8 matrix cores
        res = A.fit_lfmm2_nmf(y_est, subset, y_full, x, cfg, ncfg, nmf_int8=prec == A.PREC_I8)
    else:
        res = A.fit_lfmm2(y_est, subset, y_full, x, cfg)

    write_results(results_path, res, a.bed, cov_names=cov_names, fmt=ext, n_workers=max(a.threads, 1),
                  chunk_size=a.chunk_size, threads=a.threads)
    _eprint(f"GIF: {res.gif:.4f}")
    _eprint(f"Results written to {results_path}")
    write_summary(summary_path, res, n, n_snps, a.k, cov_names, a.lam, a.threads, nmf)
    _eprint(f"Done. Output: {results_path}, {summary_path}")
    return res


def main(argv: Optional[Sequence[str]] = None) -> int:
    a = build_parser().parse_args(argv)
    from ._lib import SlfmmError
    try:
        run(a)
    except (ValueError, OSError, SlfmmError) as e:
        _eprint(f"Error: {e.msg if isinstance(e, SlfmmError) else e}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
