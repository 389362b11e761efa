"""The reference's whole flow (src/main.rs:114-289) through the library: .bed +
covariate file -> sample matching (intersection) -> estimation panel aligned
by IID -> fit on the GPU -> results TSV.  Every step must give what the
same fit gives on genotypes and covariates that were aligned by hand: the
kept samples are the panel's sample_keep, so the fit is bit-identical."""
import numpy as np
import pytest

from oracle import output_oracle as OO
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


def _write_panel(prefix, g, iids):
    import oracle as O
    O.write_plink(prefix, g)
    with open(prefix + ".fam", "w") as f:
        for i in iids:
            f.write(f"F_{i}\t{i}\t0\t0\t0\t-9\n")


@pytest.mark.parametrize("prec", ["i8", "fp64"])
def test_cli_flow_matches_hand_aligned_fit(tmp_path, prec):
    import oracle as O
    import schnellfmm_amd as S
    n, p = 420, 6000
    pan = synth.balding_nichols(n, p, n_pops=4, miss_rate=0.02, d=2, n_causal=10, effect=1.0, seed=77)
    rng = np.random.default_rng(1)
    iids = [f"ind{i:04d}" for i in range(n)]
    main = str(tmp_path / "main")
    _write_panel(main, pan.g, iids)
    # covariate file: shuffled rows, 17 panel samples absent, 3 extra samples, some NA cells
    drop = set(rng.choice(n, 17, replace=False).tolist())
    rows = [i for i in rng.permutation(n) if i not in drop]
    x = pan.x.copy()
    na = rng.choice(rows, 5, replace=False)
    with open(tmp_path / "cov.tsv", "w") as f:
        f.write("IID\ttemp\tprec\n")
        for i in rows:
            cells = [("NA" if (i in na and c == 1) else repr(float(x[i, c]))) for c in range(2)]
            f.write(f"{iids[i]}\t{cells[0]}\t{cells[1]}\n")
        for k in range(3):
            f.write(f"extra{k}\t1.0\t2.0\n")
    fam = S.read_fam(main + ".fam")
    names, xs, kept = S.load_covariates(str(tmp_path / "cov.tsv"), fam, intersect=True)
    assert names == ["temp", "prec"] and kept == [i for i in range(n) if i not in drop]
    # estimation panel: the same genotypes with samples permuted plus 5 unrelated samples
    perm = rng.permutation(n)
    extra = synth.balding_nichols(5, p, n_pops=1, miss_rate=0.0, d=1, n_causal=0, seed=3).g
    est_g = np.concatenate([pan.g[perm], extra], axis=0)
    est = str(tmp_path / "est")
    _write_panel(est, est_g, [iids[i] for i in perm] + [f"other{k}" for k in range(5)])
    idx = S.align_est_bed_samples(S.read_fam(est + ".fam"), [fam[i] for i in kept])
    assert idx is not None and len(idx) == len(kept)

    cfg = S.Lfmm2Config(k=4, norm=S.EIGENSTRAT, n_power_iter=2, seed=5,
                        precision=S.PREC_I8 if prec == "i8" else S.PREC_FP64)
    g_main = S.BedFile.open(main + ".bed", sample_keep=kept)
    g_est = S.BedFile.open(est + ".bed", sample_keep=idx)
    res = S.fit_lfmm2(g_est, 0.25, g_main, xs, cfg)

    # by hand: subset the genotype matrix and impute the covariates directly
    g_hand = S.BedFile.from_packed(O.encode_bed(pan.g[kept]), len(kept), p)
    x_hand = x[kept].copy()
    col = x_hand[:, 1]
    miss = np.isin(np.asarray(kept), na)
    col[miss] = sum(col[~miss].tolist()) / int((~miss).sum())   # sequential sum in .fam order
    ref = S.fit_lfmm2(g_hand, 0.25, g_hand, x_hand, cfg)
    assert np.allclose(xs, x_hand, rtol=0, atol=0)
    for nm in ("u_hat", "beta", "tstat", "r2", "pvalue", "gif_per_trait"):
        assert np.array_equal(getattr(res, nm), getattr(ref, nm)), nm

    out = str(tmp_path / "results.tsv")
    S.write_results(out, res, main + ".bed", cov_names=names)
    recs = OO.read_bim(main + ".bim")
    want = OO.results_tsv([r[0] for r in recs], [r[3] for r in recs], [r[1] for r in recs], names,
                          res.pvalue, res.beta, res.tstat, res.r2)
    assert open(out).read() == want


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_cli_end_to_end(tmp_path, prec, capsys):
    """python -m schnellfmm_amd (src/main.rs:114-289): the results file and the
    summary equal those of the library steps run by hand on the same inputs."""
    import schnellfmm_amd as S
    from schnellfmm_amd import cli
    n, p = 300, 5000
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=0.02, d=2, n_causal=8, effect=1.0, seed=21)
    iids = [f"s{i}" for i in range(n)]
    main = str(tmp_path / "main")
    _write_panel(main, pan.g, iids)
    keep = [i for i in range(n) if i % 7 != 3]
    with open(tmp_path / "cov.csv", "w") as f:
        f.write("id,env_a,env_b\n")
        for i in reversed(keep):
            f.write(f"{iids[i]},{float(pan.x[i, 0])!r},{float(pan.x[i, 1])!r}\n")
    out = str(tmp_path / "res")
    rc = cli.main(["-b", main + ".bed", "-c", str(tmp_path / "cov.csv"), "-k", "4", "-o", out, "-t", "3",
                   "--est-rate", "0.25", "--intersect-samples", "--power-iter", "3", "--seed", "11",
                   "--precision", prec])
    err = capsys.readouterr().err
    assert rc == 0, err
    assert "Running LFMM2: K=4, lambda=1.0e-5, chunk_size=10000, threads=3, oversampling=10, power_iter=3" in err

    cfg = S.Lfmm2Config(k=4, n_power_iter=3, seed=11, norm=S.EIGENSTRAT,
                        precision=S.PREC_I8 if prec == "mixed" else S.PREC_FP64)
    g = S.BedFile.open(main + ".bed", sample_keep=keep)
    ref = S.fit_lfmm2(g, 0.25, g, pan.x[keep], cfg)
    want = str(tmp_path / "want.tsv")
    S.write_results(want, ref, main + ".bed", cov_names=["env_a", "env_b"])
    assert open(out + ".tsv").read() == open(want).read()
    assert open(out + ".summary.txt").read() == (
        f"GIF: {ref.gif:.6f}\nn_samples: {len(keep)}\nn_snps: {p}\nK: 4\nd: 2\ncovariates: env_a, env_b\n"
        "lambda: 0.00001\nthreads: 3\nimputation: mean\n")
    assert f"GIF: {ref.gif:.4f}" in err and err.rstrip().endswith(f"Done. Output: {out}.tsv, {out}.summary.txt")


def test_cli_nmf_parquet(tmp_path, capsys):
    """--nmf-impute with Parquet output: the NMF CV lines of the summary
    (src/main.rs:262-282) and the Parquet table equal the library's."""
    import pyarrow.parquet as pq
    import schnellfmm_amd as S
    from schnellfmm_amd import cli
    n, p = 240, 3000
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=0.05, d=1, n_causal=5, effect=1.0, seed=8)
    iids = [f"s{i}" for i in range(n)]
    main = str(tmp_path / "main")
    _write_panel(main, pan.g, iids)
    with open(tmp_path / "cov.tsv", "w") as f:
        f.write("id\ttrait\n")
        for i in range(n):
            f.write(f"{iids[i]}\t{float(pan.x[i, 0])!r}\n")
    out = str(tmp_path / "res")
    rc = cli.main(["-b", main + ".bed", "-c", str(tmp_path / "cov.tsv"), "-k", "3", "-o", out, "-t", "2",
                   "-f", "parquet", "--chunk-size", "700", "--nmf-impute", "--nmf-k", "4", "--nmf-iter", "3",
                   "--nmf-cv-rate", "0.01", "--est-rate", "0.5"])
    err = capsys.readouterr().err
    assert rc == 0, err
    cfg = S.Lfmm2Config(k=3, chunk_size=700, norm=S.EIGENSTRAT, impute="nmf")
    nmf = S.NmfParams(k=4, n_iter=3, cv_rate=0.01, chunk_size=700, seed=42)
    g = S.BedFile.open(main + ".bed")
    ref = S.fit_lfmm2_nmf(g, 0.5, g, pan.x, cfg, nmf)
    want = str(tmp_path / "want.parquet")
    S.write_results(want, ref, main + ".bed", cov_names=["trait"], fmt="parquet", n_workers=2, chunk_size=700)
    assert pq.read_table(out + ".parquet").equals(pq.read_table(want))
    lines = open(out + ".summary.txt").read().splitlines()
    assert lines[8:] == (["imputation: nmf", "nmf_iterations: 3"]
                         + [f"nmf_cv_mae_iter_{i + 1}: {v:.6f}" for i, v in enumerate(ref.nmf_cv)]
                         + [f"mean_impute_cv_mae: {ref.nmf_cv_mean:.6f}", f"nmf_gwas_cv_mae: {ref.gwas_cv[0]:.6f}",
                            f"nmf_gwas_cv_mean_mae: {ref.gwas_cv[1]:.6f}", f"nmf_gwas_cv_count: {ref.gwas_cv[2]}"])
    assert ref.gwas_cv[2] > 0 and 0 < ref.nmf_cv_mean


def test_cli_est_bed_center_only_scale_cov(tmp_path, capsys):
    """--est-bed (samples permuted, extras), --norm center-only, --scale-cov,
    -l and --oversampling through the command line, against the same steps by hand."""
    import schnellfmm_amd as S
    from schnellfmm_amd import cli
    n, p = 260, 4000
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=0.03, d=2, n_causal=6, effect=1.0, seed=31)
    iids = [f"s{i}" for i in range(n)]
    main = str(tmp_path / "main")
    _write_panel(main, pan.g, iids)
    with open(tmp_path / "cov.tsv", "w") as f:
        f.write("id\ta\tb\n")
        for i in range(n):
            f.write(f"{iids[i]}\t{float(pan.x[i, 0])!r}\t{float(pan.x[i, 1])!r}\n")
    rng = np.random.default_rng(5)
    perm = rng.permutation(n)
    est_snps = np.ascontiguousarray(pan.g[:, ::3])
    extra = synth.balding_nichols(4, est_snps.shape[1], n_pops=1, miss_rate=0.0, d=1, n_causal=0, seed=9).g
    est = str(tmp_path / "est")
    _write_panel(est, np.concatenate([est_snps[perm], extra], axis=0),
                 [iids[i] for i in perm] + [f"x{k}" for k in range(4)])
    out = str(tmp_path / "res")
    rc = cli.main(["-b", main + ".bed", "-c", str(tmp_path / "cov.tsv"), "-k", "3", "-o", out, "-t", "2",
                   "--est-bed", est + ".bed", "--norm", "center-only", "--scale-cov", "-l", "0.001",
                   "--oversampling", "6", "--precision", "fp64"])
    err = capsys.readouterr().err
    assert rc == 0, err
    assert f"  {n} samples x {est_snps.shape[1]} SNPs (for factor estimation)" in err.splitlines()

    idx = S.align_est_bed_samples(S.read_fam(est + ".fam"), S.read_fam(main + ".fam"))
    cfg = S.Lfmm2Config(k=3, lam=0.001, oversampling=6, norm=S.CENTER_ONLY, scale_cov=True)
    g = S.BedFile.open(main + ".bed")
    ge = S.BedFile.open(est + ".bed", sample_keep=idx)
    ref = S.fit_lfmm2(ge, None, g, pan.x, cfg)
    want = str(tmp_path / "want.tsv")
    S.write_results(want, ref, main + ".bed", cov_names=["a", "b"])
    assert open(out + ".tsv").read() == open(want).read()
    assert "lambda: 0.001\n" in open(out + ".summary.txt").read()
