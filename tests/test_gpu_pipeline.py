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
