"""The parity workloads of the measurement plan, at (or near) their stated
shapes:

* workload 1 (CPU-sized parity): n = 1000 in 3 populations, p = 50 000, F = 0.1,
  p_anc ~ U(0.05, 0.5), 1 % missing, one covariate, 50 causal SNPs (beta 0.3);
  K = 5, q = 2, oversampling 10, chunks of 10 000, eigenstrat, RSVD on all SNPs.
  The FP64 GPU fit against the oracle's fit (same injected sketch): U_hat
  principal angles < 1e-8, t within 1e-7 relative, log10 p within 1e-6; the
  mixed-precision fit within |delta log10 p| <= 1e-3 of the oracle.
* workload 2 (inbred lines), p reduced from 1.8M to 120 000 so the host can
  generate it: n = 1135, 5 % heterozygous calls, MAF ~ Beta(0.4, 1.6) (>= 0.01),
  per-sample missingness ~ Beta(1, 19), 8 groups with Dirichlet(0.3) ancestry,
  the trait (d = 1) built from the first ancestry axis plus noise; K = 12,
  q = 2.  Mixed precision against the FP64 GPU path on every SNP
  (|delta log10 p| <= 1e-3, the headline accuracy bound), and the FP64 GPU
  association against the oracle on a 6 000-SNP slice with the same U_hat.
"""
import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


@pytest.fixture(scope="module")
def workload1():
    n, p, K = 1000, 50_000, 5
    pan = synth.balding_nichols(n, p, n_pops=3, fst=0.1, maf=(0.05, 0.5), miss_rate=0.01, d=1, n_causal=50,
                                effect=0.3, seed=1)
    packed = O.encode_bed(pan.g)
    bps = O.bytes_per_snp(n)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ synth.normal_matrix(42, n, K + 10)
    cfg = O.Config(k=K, norm=O.EIGENSTRAT, chunk_size=10_000, n_power_iter=2, seed=42)
    u_ref, ref, _, _ = O.fit_lfmm2(packed, bps, n, p, pan.x, cfg, np.arange(p), b0=b0)
    return pan, packed, b0, u_ref, ref


@pytest.mark.parametrize("prec", ["fp64", "i8"])
def test_workload1_vs_oracle(S, workload1, prec):
    pan, packed, b0, u_ref, ref = workload1
    n, p = pan.g.shape
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.Lfmm2Config(k=5, oversampling=10, n_power_iter=2, chunk_size=10_000, norm=S.EIGENSTRAT,
                        precision=S.PREC_FP64 if prec == "fp64" else S.PREC_I8)
    got = S.fit_lfmm2(g, "all", g, pan.x, cfg, b0=b0)
    dlp = np.abs(O.log10_pvalue(got.tstat, got.gif_per_trait) - O.log10_pvalue(ref.t, ref.gif_per_trait))
    if prec == "fp64":
        assert np.max(O.principal_angles(got.u_hat, u_ref)) < 1e-8
        assert np.allclose(got.tstat, ref.t, rtol=1e-7, atol=1e-8)
        assert np.max(dlp) < 1e-6
        assert np.allclose(got.gif_per_trait, ref.gif_per_trait, rtol=1e-12)
    else:
        assert np.max(dlp) <= 1e-3
    assert np.max(np.abs(ref.t)) > 5   # the causal SNPs are found


def test_workload2_inbred_mixed_vs_fp64(S):
    n, p, K = 1135, 120_000, 12
    pan = synth.balding_nichols(n, p, n_pops=8, fst=0.1, admix_alpha=0.3, inbred_het=0.05, maf_beta=(0.4, 1.6),
                                sample_miss_beta=(1.0, 19.0), d=1, n_causal=30, effect=0.5, seed=2)
    # trait: the first ancestry axis (population label, standardised) plus the generator's covariate
    axis = (pan.pops - pan.pops.mean()) / max(pan.pops.std(), 1e-12)
    x = (0.4 * axis)[:, None] + pan.x
    packed = O.encode_bed(pan.g)
    g = S.BedFile.from_packed(packed, n, p)
    base = dict(k=K, oversampling=10, n_power_iter=2, norm=S.EIGENSTRAT, seed=42)
    f64 = S.fit_lfmm2(g, "all", g, x, S.Lfmm2Config(precision=S.PREC_FP64, **base))
    mix = S.fit_lfmm2(g, "all", g, x, S.Lfmm2Config(precision=S.PREC_I8, **base))
    dlp = np.abs(O.log10_pvalue(mix.tstat, mix.gif_per_trait) - O.log10_pvalue(f64.tstat, f64.gif_per_trait))
    assert np.max(dlp) <= 1e-3
    assert np.max(O.principal_angles(mix.u_hat, f64.u_hat)[:K // 2]) < 1e-3
    # the FP64 association on a slice against the oracle, with the same U_hat
    j0, j1 = 40_000, 46_000
    xs = O.center_covariates(x)
    pre = O.precompute(xs, 1e-5)
    sl = O.encode_bed(pan.g[:, j0:j1])
    ref = O.test_associations_fused(sl, O.bytes_per_snp(n), n, j1 - j0, xs, f64.u_hat, pre, K, 2000, O.EIGENSTRAT,
                                    O.Impute())
    assert np.allclose(f64.tstat[j0:j1], ref.t, rtol=1e-8, atol=1e-9)
    assert np.allclose(f64.beta[j0:j1], ref.beta, rtol=1e-8, atol=1e-12)
