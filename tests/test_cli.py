"""The command line (src/main.rs:12-289) on the host: flag names, short forms
and defaults of the reference's Cli struct, the reference's number spellings
in its progress lines and summary file, and the error paths that end the run
before any genotype reaches the GPU."""
import json
import os

import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import cli

# SubsetSpec::Rate expansions from the reference's own tests
KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")))


def test_defaults_match_reference_cli():
    # src/main.rs:16-107
    a = cli.build_parser().parse_args(["-b", "x.bed", "-c", "c.tsv", "-k", "7"])
    assert (a.bed, a.cov, a.k) == ("x.bed", "c.tsv", 7)
    assert a.lam == 1e-5 and a.out == "lfmm2_out" and a.format == "tsv"
    assert (a.chunk_size, a.oversampling, a.power_iter, a.seed) == (10_000, 10, 2, 42)
    assert a.est_bed is None and a.est_rate is None and not a.intersect_samples and not a.verbose
    assert a.norm == "eigenstrat" and not a.scale_cov and not a.nmf_impute
    assert a.nmf_k is None and a.nmf_iter == 10 and a.nmf_cv_rate == 0.0005
    assert a.precision == "fp64" and not a.host_panel


def test_short_and_long_forms():
    a = cli.build_parser().parse_args(
        ["--bed", "x.bed", "--cov", "c.csv", "-k", "3", "-l", "0.5", "-o", "pre", "-f", "PARQUET", "-t", "4",
         "--chunk-size", "500", "--oversampling", "5", "--power-iter", "3", "--seed", "9", "--est-rate", "0.1",
         "--intersect-samples", "-v", "--norm", "center-only", "--scale-cov", "--nmf-impute", "--nmf-k", "6",
         "--nmf-iter", "4", "--nmf-cv-rate", "0.01"])
    assert (a.lam, a.out, a.format, a.threads) == (0.5, "pre", "parquet", 4)
    assert (a.chunk_size, a.oversampling, a.power_iter, a.seed, a.est_rate) == (500, 5, 3, 9, 0.1)
    assert a.intersect_samples and a.verbose and a.norm == "center-only" and a.scale_cov
    assert a.nmf_impute and (a.nmf_k, a.nmf_iter, a.nmf_cv_rate) == (6, 4, 0.01)


@pytest.mark.parametrize("argv", [
    ["-c", "c.tsv", "-k", "3"],                                   # --bed is required
    ["-b", "x.bed", "-c", "c.tsv"],                               # -k is required
    ["-b", "x.bed", "-c", "c.tsv", "-k", "3", "--mean-impute", "--nmf-impute"],   # ArgGroup "impute"
    ["-b", "x.bed", "-c", "c.tsv", "-k", "3", "-f", "csv"],
    ["-b", "x.bed", "-c", "c.tsv", "-k", "3", "--norm", "zscore"],
])
def test_usage_errors(argv):
    with pytest.raises(SystemExit) as e:
        cli.build_parser().parse_args(argv)
    assert e.value.code == 2


@pytest.mark.parametrize("x,want", [
    # Rust's Display for f64: shortest round-trip digits, never an exponent
    (1e-5, "0.00001"), (0.1, "0.1"), (2.0, "2"), (0.5, "0.5"), (1.5e-7, "0.00000015"),
    (1e21, "1000000000000000000000"), (123.25, "123.25"), (-0.0, "-0"), (-3.0, "-3"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (0.30000000000000004, "0.30000000000000004"),
])
def test_rust_display(x, want):
    assert cli.rust_display(x) == want


@pytest.mark.parametrize("x,want", [
    (1e-5, "1.0e-5"), (0.001, "1.0e-3"), (12345.0, "1.2e4"), (0.05, "5.0e-2"), (1.0, "1.0e0"), (0.25, "2.5e-1"),
])
def test_rust_exp1(x, want):
    assert cli.rust_exp1(x) == want


def test_rate_indices_match_subset_golden_vectors():
    for c in KAT["subset_indices"]:
        if isinstance(c["spec"], float):
            assert list(cli._rate_indices(c["spec"], c["n"])) == c["idx"]
    assert list(cli._rate_indices(0.3, 10)) == [0, 4, 8]      # ceil(1/0.3) = 4
    assert list(cli._rate_indices(0.0, 10)) == [0]            # step saturates: SNP 0 alone
    assert list(cli._rate_indices(1.0, 3)) == [0, 1, 2]


def test_summary_file(tmp_path):
    class R:
        gif, nmf_cv, nmf_cv_mean, gwas_cv = 1.0123456789, np.array([0.31, 0.2999996]), 0.4, (0.25, 0.5, 17)
    out = tmp_path / "s.txt"
    cli.write_summary(str(out), R, 100, 2000, 5, ["temp", "prec"], 1e-5, 7, True)
    assert out.read_text() == (
        "GIF: 1.012346\nn_samples: 100\nn_snps: 2000\nK: 5\nd: 2\ncovariates: temp, prec\nlambda: 0.00001\n"
        "threads: 7\nimputation: nmf\nnmf_iterations: 2\nnmf_cv_mae_iter_1: 0.310000\n"
        "nmf_cv_mae_iter_2: 0.300000\nmean_impute_cv_mae: 0.400000\nnmf_gwas_cv_mae: 0.250000\n"
        "nmf_gwas_cv_mean_mae: 0.500000\nnmf_gwas_cv_count: 17\n")

    class M:
        gif, nmf_cv, nmf_cv_mean, gwas_cv = 0.99, None, None, None
    cli.write_summary(str(out), M, 10, 20, 2, ["x"], 0.5, 1, False)
    assert out.read_text() == ("GIF: 0.990000\nn_samples: 10\nn_snps: 20\nK: 2\nd: 1\ncovariates: x\nlambda: 0.5\n"
                               "threads: 1\nimputation: mean\n")


@pytest.fixture
def panel(tmp_path):
    g = np.random.default_rng(0).integers(0, 3, size=(6, 40)).astype(np.uint8)
    O.write_plink(str(tmp_path / "p"), g)
    (tmp_path / "c.tsv").write_text("IID\ta\n" + "".join(f"ind{i}\t{i * 0.5}\n" for i in range(6)))
    return tmp_path


def _run(capsys, argv):
    rc = cli.main(argv)
    return rc, capsys.readouterr().err


def test_est_bed_and_rate_conflict(panel, capsys):
    rc, err = _run(capsys, ["-b", str(panel / "p.bed"), "-c", str(panel / "c.tsv"), "-k", "2",
                            "--est-bed", str(panel / "p.bed"), "--est-rate", "0.5", "-o", str(panel / "o")])
    assert rc == 1
    lines = err.splitlines()
    assert lines[:3] == [f"Loading BED file: {panel / 'p.bed'}", "  6 samples x 40 SNPs",
                         f"Loading covariates from: {panel / 'c.tsv'}"]
    assert "  6 samples x 1 covariates: [a]" in lines
    assert err.splitlines()[-1] == "Error: Cannot specify both --est-bed and --est-rate"


@pytest.mark.parametrize("rate,shown", [("1.5", "1.5"), ("-0.25", "-0.25"), ("nan", "NaN")])
def test_est_rate_out_of_range(panel, capsys, rate, shown):
    rc, err = _run(capsys, ["-b", str(panel / "p.bed"), "-c", str(panel / "c.tsv"), "-k", "2", "--est-rate", rate])
    assert rc == 1 and err.splitlines()[-1] == f"Error: --est-rate must be in (0.0, 1.0], got {shown}"


def test_missing_sample_and_bad_bed(panel, capsys):
    (panel / "c2.tsv").write_text("IID\ta\nind0\t1\n")
    rc, err = _run(capsys, ["-b", str(panel / "p.bed"), "-c", str(panel / "c2.tsv"), "-k", "2"])
    assert rc == 1 and "Error: Sample 'ind1' (FID='fam1') from .fam file not found in covariate file" in err
    raw = (panel / "p.bed").read_bytes()
    (panel / "p.bed").write_bytes(raw[:-1])
    rc, err = _run(capsys, ["-b", str(panel / "p.bed"), "-c", str(panel / "c.tsv"), "-k", "2"])
    assert rc == 1 and err.splitlines()[-1].startswith("Error: ") and "size" in err.lower()


def test_module_entry_point(panel):
    """python -m schnellfmm_amd: --help exits 0, a usage error exits 2, a failing run exits 1."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    run = lambda *a: subprocess.run([sys.executable, "-m", "schnellfmm_amd", *a], cwd=root,
                                    capture_output=True, text=True, timeout=120)
    r = run("--help")
    assert r.returncode == 0 and "--nmf-impute" in r.stdout and "--est-rate" in r.stdout
    assert run("-b", "x.bed").returncode == 2
    r = run("-b", str(panel / "p.bed"), "-c", str(panel / "c.tsv"), "-k", "2", "--est-rate", "2")
    assert r.returncode == 1 and r.stderr.splitlines()[-1] == "Error: --est-rate must be in (0.0, 1.0], got 2"
