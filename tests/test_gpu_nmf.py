"""GPU parity of NMF imputation estimation (src/nmf.rs:80-455) against the
float64 oracle.  The oracle is driven by the same random streams as the
library: Omega' = synth.normal_matrix(seed, n, k + 10) and the CV hold-out
draws of synth.uniform_stream (the reference uses ChaCha8 for both, which is
parity unpinned).  Tolerances: W and H within 1e-8 relative (FP64 with
different summation orders over up to 10 multiplicative iterations), CV MAE
within 1e-9 relative, held-out counts exact."""
import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def _panel(n, p, seed, miss):
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=miss, d=1, n_causal=0, seed=seed)
    return O.encode_bed(pan.g), O.bytes_per_snp(n)


def _close(a, b, rtol):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) <= rtol * scale


@pytest.mark.parametrize("n,p,k,n_iter,rate,chunk,miss", [
    (300, 1500, 3, 5, None, 400, 0.05),
    (257, 2000, 4, 3, 0.5, 333, 0.10),     # ragged sample tile, subset, several CV chunks
])
def test_nmf_estimation_vs_oracle(S, n, p, k, n_iter, rate, chunk, miss):
    packed, bps = _panel(n, p, seed=n + p, miss=miss)
    idx = np.arange(p) if rate is None else O.subset_indices(rate, p)
    seed, cv_rate, eps = 42, 0.01, 1e-16
    omega = synth.normal_matrix(seed, n, k + 10)
    ref = O.nmf_impute_estimation(packed, bps, n, idx, k, n_iter, eps, cv_rate, chunk, seed, omega,
                                  synth.uniform_stream)
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.NmfParams(k=k, n_iter=n_iter, eps=eps, cv_rate=cv_rate, chunk_size=chunk, seed=seed)
    got = S.nmf_impute_estimation(g, rate, cfg, omega=omega)
    assert got.w.shape == ref.w.shape and got.h_full.shape == ref.h_full.shape
    assert _close(got.w, ref.w, 1e-8)
    assert _close(got.h_full, ref.h_full, 1e-8)
    assert got.total_mask_count == ref.total_mask_count > 0
    assert np.allclose(got.cv_mae_per_iter, ref.cv_mae_per_iter, rtol=1e-9, atol=0)
    assert np.isclose(got.cv_mae_mean_impute, ref.cv_mae_mean_impute, rtol=1e-12, atol=0)
    # the held-out error of the factorisation beats mean imputation on structured data
    assert got.cv_mae_per_iter[-1] < got.cv_mae_mean_impute
    # Omega' drawn on the device from the seed is the same stream (device log/cos
    # may differ from numpy's in the last bit, hence a tolerance)
    got2 = S.nmf_impute_estimation(g, rate, cfg)
    assert _close(got2.w, got.w, 1e-8) and _close(got2.h_full, got.h_full, 1e-8)


def test_nmf_rank_limit(S):
    packed, bps = _panel(40, 30, seed=1, miss=0.0)
    g = S.BedFile.from_packed(packed, 40, 30)
    with pytest.raises(S.api.SlfmmError):
        S.nmf_impute_estimation(g, None, S.NmfParams(k=17, n_iter=1))


@pytest.mark.parametrize("int8,cv_rate", [(False, 0.02), (True, 0.02), (False, 0.2)])
def test_fit_lfmm2_nmf_vs_oracle(S, int8, cv_rate):
    """fit_lfmm2 with NMF imputation (src/lib.rs:162-263): NmfInRam fill in the
    RSVD, NmfOnTheFly fill in the association, GWAS-phase CV
    (src/testing.rs:423-487).  FP64: U_hat angles < 1e-8, t / beta / r2
    within 1e-8 relative, CV errors within 1e-9, counts exact.  With the
    NMF update passes on the int8 matrix cores (W, H within about 1e-6 of
    FP64): angles < 1e-5, t / beta / r2 within 1e-5 of their largest
    entries, CV errors within 1e-5 relative, counts exact.  cv_rate 0.2
    holds out about 50 calls per SNP in the association-phase CV, past the
    kernel's 16-entry list (it draws again for those SNPs)."""
    tol, tol_cv = (1e-5, 1e-5) if int8 else (1e-8, 1e-9)
    n, p, K, d = 300, 2000, 3, 2
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=0.10, d=d, n_causal=20, effect=1.0, seed=77)
    packed, bps = O.encode_bed(pan.g), O.bytes_per_snp(n)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ np.random.default_rng(4).standard_normal((n, K + 10))
    kn, seed_nmf = 3, 7
    omega = synth.normal_matrix(seed_nmf, n, kn + 10)
    cfg_o = O.Config(k=K, norm=O.EIGENSTRAT, chunk_size=500, n_power_iter=2, seed=42)
    est = O.subset_indices(0.5, p)
    nmf_o = dict(k=kn, n_iter=3, eps=1e-16, cv_rate=cv_rate, seed=seed_nmf, omega=omega,
                 uniform=synth.uniform_stream)
    u_ref, ref, _, nres = O.fit_lfmm2(packed, bps, n, p, pan.x, cfg_o, est, b0=b0, nmf=nmf_o)
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, chunk_size=500, n_power_iter=2, seed=42)
    # the reference CLI gives NmfConfig the LFMM chunk size (src/main.rs:183, :203)
    nmf = S.NmfParams(k=kn, n_iter=3, cv_rate=cv_rate, seed=seed_nmf, chunk_size=500)
    got = S.fit_lfmm2_nmf(g, 0.5, g, pan.x, cfg, nmf, b0=b0, nmf_omega=omega, nmf_int8=int8)
    assert np.max(O.principal_angles(got.u_hat, u_ref)) < tol
    assert _close(got.tstat, ref.t, tol)
    assert _close(got.beta, ref.beta, tol)
    assert _close(got.r2, ref.r2, tol)
    assert np.allclose(got.gif_per_trait, ref.gif_per_trait, rtol=1e-12 if not int8 else 1e-5)
    assert np.allclose(got.nmf_cv, nres.cv_mae_per_iter, rtol=tol_cv, atol=0)
    assert np.isclose(got.nmf_cv_mean, nres.cv_mae_mean_impute, rtol=1e-12 if not int8 else 1e-5, atol=0)
    assert got.gwas_cv[2] == ref.cv[2] > 0
    assert np.allclose(got.gwas_cv[:2], ref.cv[:2], rtol=tol_cv, atol=0)
    assert np.max(np.abs(ref.t)) > 4   # the panel has real signal


@pytest.mark.parametrize("nmf_int8,kn", [(False, 3), (True, 3), (True, 10)])
def test_fit_lfmm2_nmf_mixed_vs_oracle(S, nmf_int8, kn):
    """The NMF-imputed fit in mixed precision: the filled sketch/backward
    passes and the filled association functionals on the int8 matrix cores
    (A, or B, in 3 int8 terms against the dosages, the products A_c W_r in 2
    against the missing-call indicator), counts, moments, forward passes and
    epilogue in FP64; optionally the NMF update passes in int8 too.  Against
    the FP64 oracle: U_hat angles < 1e-4 rad, |delta log10 p| <= 1e-3, GWAS-
    phase CV counts exact and errors within 1e-5 relative.  nmf-k = 10 is
    past the int8 fill plans: the filled passes run in FP64 there, with the
    same tolerances."""
    n, p, K, d = 300, 2000, 3, 2
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=0.10, d=d, n_causal=20, effect=1.0, seed=77)
    packed = O.encode_bed(pan.g)
    bps = O.bytes_per_snp(n)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ np.random.default_rng(4).standard_normal((n, K + 10))
    seed_nmf, cv_rate = 7, 0.02
    omega = synth.normal_matrix(seed_nmf, n, kn + 10)
    cfg_o = O.Config(k=K, norm=O.EIGENSTRAT, chunk_size=500, n_power_iter=2, seed=42)
    est = O.subset_indices(0.5, p)
    nmf_o = dict(k=kn, n_iter=3, eps=1e-16, cv_rate=cv_rate, seed=seed_nmf, omega=omega,
                 uniform=synth.uniform_stream)
    u_ref, ref, _, nres = O.fit_lfmm2(packed, bps, n, p, pan.x, cfg_o, est, b0=b0, nmf=nmf_o)
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, chunk_size=500, n_power_iter=2, seed=42, precision=S.PREC_I8)
    nmf = S.NmfParams(k=kn, n_iter=3, cv_rate=cv_rate, seed=seed_nmf, chunk_size=500)
    got = S.fit_lfmm2_nmf(g, 0.5, g, pan.x, cfg, nmf, b0=b0, nmf_omega=omega, nmf_int8=nmf_int8)
    ang = np.max(O.principal_angles(got.u_hat, u_ref))
    lp = np.abs(O.log10_pvalue(got.tstat, got.gif_per_trait) - O.log10_pvalue(ref.t, ref.gif_per_trait))
    print(f"[nmf mixed] angles {ang:.2e}, max |dlog10p| {lp.max():.2e}", flush=True)
    assert ang < 1e-4
    assert lp.max() <= 1e-3
    assert got.gwas_cv[2] == ref.cv[2] > 0
    assert np.allclose(got.gwas_cv[:2], ref.cv[:2], rtol=1e-5, atol=0)
    assert np.max(np.abs(ref.t)) > 4


@pytest.mark.parametrize("n,p,k,n_iter,rate,chunk,miss", [
    (300, 1500, 3, 5, None, 400, 0.05),
    (257, 2000, 8, 4, 0.5, 333, 0.10),     # rank 8 (4 row tiles), ragged sample tile, subset
    (1100, 3000, 6, 6, None, 1000, 0.15),
    (530, 40001, 1, 3, None, 5000, 0.08),   # rank 1, several SNP chunks, ragged last block
    (1030, 20000, 5, 3, None, 2000, 0.03),  # 3 sample workgroups of the forward pass
])
def test_nmf_int8_vs_oracle(S, n, p, k, n_iter, rate, chunk, miss):
    """Both NMF passes on the int8 matrix cores (SLFMM_NMF_INT8).  Forward:
    the dosages meet H in 3 int8 terms and the missing-call indicator the
    products H_r H_s in 2, on the sample-major copy; backward: W in 3 terms,
    W_r W_s in 2; every sum is recombined in FP64 from exact int32 sums.
    Tolerances: W and H within 1e-5 of the largest entry (about the FP32
    level the NMF needs) after every multiplicative iteration, per-iteration
    CV MAE within 1e-5 relative, held-out counts exact (the same mask
    stream)."""
    print(f"[nmf int8] n={n} p={p} k={k} iters={n_iter}", flush=True)
    packed, bps = _panel(n, p, seed=n + p + k, miss=miss)
    idx = np.arange(p) if rate is None else O.subset_indices(rate, p)
    seed, cv_rate, eps = 42, 0.01, 1e-16
    omega = synth.normal_matrix(seed, n, k + 10)
    ref = O.nmf_impute_estimation(packed, bps, n, idx, k, n_iter, eps, cv_rate, chunk, seed, omega,
                                  synth.uniform_stream)
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.NmfParams(k=k, n_iter=n_iter, eps=eps, cv_rate=cv_rate, chunk_size=chunk, seed=seed)
    got = S.nmf_impute_estimation(g, rate, cfg, omega=omega, int8=True)
    f64 = S.nmf_impute_estimation(g, rate, cfg, omega=omega)
    assert _close(got.w, ref.w, 1e-5) and _close(got.h_full, ref.h_full, 1e-5)
    assert not (np.array_equal(got.w, f64.w) and np.array_equal(got.h_full, f64.h_full))   # the int8 pass ran
    assert got.total_mask_count == ref.total_mask_count > 0
    assert np.allclose(got.cv_mae_per_iter, ref.cv_mae_per_iter, rtol=1e-5, atol=0)
    assert np.isclose(got.cv_mae_mean_impute, ref.cv_mae_mean_impute, rtol=1e-12, atol=0)


@pytest.mark.parametrize("k", [2, 4, 7])
def test_nmf_int8_remaining_ranks_vs_oracle(S, k):
    """The int8 passes at the ranks the shapes above leave out (their tile
    plans differ: forward 1+1, 1+2 and 2+4 tiles of 16, backward 1+1, 1+1
    and 1+2 tiles of 32), same tolerances."""
    n, p, iters, chunk = 333, 1200, 3, 500
    print(f"[nmf int8] n={n} p={p} k={k} iters={iters}", flush=True)
    packed, bps = _panel(n, p, seed=100 + k, miss=0.08)
    seed, cv_rate, eps = 11, 0.02, 1e-16
    omega = synth.normal_matrix(seed, n, k + 10)
    ref = O.nmf_impute_estimation(packed, bps, n, np.arange(p), k, iters, eps, cv_rate, chunk, seed, omega,
                                  synth.uniform_stream)
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.NmfParams(k=k, n_iter=iters, eps=eps, cv_rate=cv_rate, chunk_size=chunk, seed=seed)
    got = S.nmf_impute_estimation(g, None, cfg, omega=omega, int8=True)
    assert _close(got.w, ref.w, 1e-5) and _close(got.h_full, ref.h_full, 1e-5)
    assert got.total_mask_count == ref.total_mask_count > 0
    assert np.allclose(got.cv_mae_per_iter, ref.cv_mae_per_iter, rtol=1e-5, atol=0)


def test_nmf_int8_rank_above_8_runs_fp64(S):
    """Ranks above 8 take the FP64 backward pass: the int8 flag changes nothing."""
    packed, bps = _panel(200, 800, seed=9, miss=0.05)
    g = S.BedFile.from_packed(packed, 200, 800)
    cfg = S.NmfParams(k=10, n_iter=2, eps=1e-16, cv_rate=0.01, chunk_size=400, seed=3)
    om = synth.normal_matrix(3, 200, 20)
    a = S.nmf_impute_estimation(g, None, cfg, omega=om, int8=True)
    b = S.nmf_impute_estimation(g, None, cfg, omega=om)
    assert np.array_equal(a.w, b.w) and np.array_equal(a.h_full, b.h_full)
