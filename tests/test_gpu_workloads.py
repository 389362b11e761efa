"""GPU tests at the measurement plan's full workload shapes.

* headline (n = 10 000, p = 10M, K = 20, d = 4, 1 % estimation subset, q = 3,
  20 populations, 2 % missing, covariates equicorrelated rho = 0.3): the
  mixed-precision fit against the FP64 GPU fit on every SNP (|delta log10 p|
  <= 1e-3), r2 partitions, GIFs; both associations against the oracle on a
  2 000-SNP slice read back from HBM, with the FP64 fit's U_hat.
* batch (64 seeded problems of uneven shape): every problem of one batched
  call bit-identical to its single fit; the three smallest against the oracle.
* NMF (n = 5 000, p = 2M, 8 Dirichlet(0.2) components, structured 10 %
  missingness, nmf-k = 8, 10 iterations, cv-rate 5e-4): full-size invariants,
  and the estimation on a 3 000-SNP subset against the oracle (W, H within
  1e-8, CV MAE within 1e-9, held-out counts exact); with both update passes
  on the int8 matrix cores, W and H within 1e-5 of the FP64 run at full size
  and of the oracle on the subset, held-out counts exact.
* inbred lines (n = 1 135, p = 1.8M, K = 12): mixed against FP64 on every SNP
  and the FP64 association against the oracle on a 6 000-SNP slice.
Each test prints its workload before the GPU work starts."""
import math

import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def _log(msg):
    print(msg, flush=True)


def _ridge_only(xs, lam=1e-5):
    """The oracle's association reads only the ridge inverse of Precomputed."""
    d = xs.shape[1]
    return O.Precomputed(None, None, None, None, None, O.safe_inv(xs.T @ xs + lam * np.eye(d)))


def _r2_partition_ok(r2):
    s = r2.sum(axis=1)
    return np.all((np.abs(s - 1.0) < 1e-9) | (s == 0.0))


# ---------------------------------------------------------------------------
# headline workload
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def headline(S):
    n, p, K, d = 10_000, 10_000_000, 20, 4
    _log(f"[workload] headline n={n} p={p} K={K} d={d}: generating the panel in HBM")
    geno = S.BedFile.synthetic(n, p, n_pops=20, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02, seed=3)
    rng = np.random.default_rng(3)
    cov = np.full((d, d), 0.3) + 0.7 * np.eye(d)
    x = np.ascontiguousarray(rng.multivariate_normal(np.zeros(d), cov, size=n))
    base = dict(k=K, oversampling=10, n_power_iter=3, norm=S.EIGENSTRAT, seed=42)
    _log("[workload] headline: FP64 fit")
    f64 = S.fit_lfmm2(geno, 0.01, geno, x, S.Lfmm2Config(precision=S.PREC_FP64, **base))
    _log("[workload] headline: mixed-precision fit")
    mix = S.fit_lfmm2(geno, 0.01, geno, x, S.Lfmm2Config(precision=S.PREC_I8, **base))
    return geno, x, f64, mix, base


def test_headline_mixed_vs_fp64_every_snp(S, headline):
    geno, x, f64, mix, _ = headline
    lp_m = O.log10_pvalue(mix.tstat, mix.gif_per_trait)
    lp_f = O.log10_pvalue(f64.tstat, f64.gif_per_trait)
    dlp = np.abs(lp_m - lp_f)
    _log(f"[workload] headline: max |dlog10p| mixed vs FP64 = {dlp.max():.3e}")
    assert dlp.max() <= 1e-3
    assert np.all(np.isfinite(mix.tstat)) and np.all(np.isfinite(f64.tstat))
    assert _r2_partition_ok(mix.r2) and _r2_partition_ok(f64.r2)
    assert np.allclose(mix.gif_per_trait, f64.gif_per_trait, rtol=1e-3)
    # the latent directions with a clear spectral gap agree
    assert np.max(O.principal_angles(mix.u_hat, f64.u_hat)[:10]) < 1e-3


def test_headline_slice_vs_oracle(S, headline):
    geno, x, f64, mix, base = headline
    n = geno.n_samples
    j0, m = 4_321_000, 2_000
    packed = geno.read_packed(j0, m)
    xs = O.center_covariates(x)
    ref = O.test_associations_fused(packed.tobytes(), packed.shape[1], n, m, xs, f64.u_hat, _ridge_only(xs), 20, 500,
                                    O.EIGENSTRAT, O.Impute())
    assert np.allclose(f64.tstat[j0:j0 + m], ref.t, rtol=1e-8, atol=1e-10)
    assert np.allclose(f64.beta[j0:j0 + m], ref.beta, rtol=1e-8, atol=1e-12)
    assert np.allclose(f64.r2[j0:j0 + m], ref.r2, rtol=1e-8, atol=1e-12)
    # the int8 association with the same U_hat, on the whole panel
    got = S.test_associations_fused(geno, xs, f64.u_hat, S.Lfmm2Config(precision=S.PREC_I8, **base))
    dt = np.abs(got.tstat[j0:j0 + m] - ref.t)
    lp = np.abs(O.log10_pvalue(got.tstat[j0:j0 + m], got.gif_per_trait) - O.log10_pvalue(ref.t, got.gif_per_trait))
    _log(f"[workload] headline slice: int8 max |dt| {dt.max():.2e}, max |dlog10p| {lp.max():.2e}")
    assert dt.max() < 2e-4 and lp.max() <= 1e-3


# ---------------------------------------------------------------------------
# batch of 64 uneven problems
# ---------------------------------------------------------------------------

def test_batch_of_64_equals_single_fits_and_oracle(S):
    ctx = S.api.default_context()
    probs, shapes = [], []
    for i, n, p, K, d, pops, miss, x in synth.batch_problems(seed=4, count=64):
        g = S.BedFile.synthetic(n, p, n_pops=pops, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=miss,
                                seed=4000 + i, ctx=ctx)
        cfg = S.Lfmm2Config(k=K, oversampling=10, n_power_iter=2, norm=S.EIGENSTRAT, precision=S.PREC_I8, seed=42)
        probs.append(S.Lfmm2Problem(g, 0.1, g, x, cfg))
        shapes.append((n, p, K, d))
    _log(f"[workload] batch: 64 problems, {sum(s[1] for s in shapes)} SNPs")
    got = S.fit_lfmm2_batch(probs, n_streams=8)
    for i, (pr, r) in enumerate(zip(probs, got)):
        ref = S.fit_lfmm2(pr.y_est, pr.est_subset, pr.y_full, pr.x, pr.config)
        for name in ("u_hat", "beta", "tstat", "r2", "pvalue", "gif_per_trait"):
            assert np.array_equal(getattr(r, name), getattr(ref, name)), (i, name)
    # the three smallest problems against the oracle (association with the GPU's U_hat)
    for i in sorted(range(64), key=lambda k: shapes[k][0] * shapes[k][1])[:3]:
        n, p, K, d = shapes[i]
        pr, r = probs[i], got[i]
        packed = pr.y_full.read_packed(0, p)
        xs = O.center_covariates(pr.x)
        ref = O.test_associations_fused(packed.tobytes(), packed.shape[1], n, p, xs, r.u_hat, _ridge_only(xs), K,
                                        10_000, O.EIGENSTRAT, O.Impute())
        lp = np.abs(O.log10_pvalue(r.tstat, r.gif_per_trait) - O.log10_pvalue(ref.t, ref.gif_per_trait))
        _log(f"[workload] batch problem {i} (n={n} p={p} K={K} d={d}): max |dt| "
             f"{np.abs(r.tstat - ref.t).max():.2e}, max |dlog10p| {lp.max():.2e}")
        assert np.max(np.abs(r.tstat - ref.t)) < 2e-4
        assert lp.max() <= 1e-3
        assert _r2_partition_ok(r.r2)


# ---------------------------------------------------------------------------
# NMF imputation workload
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def nmf_panel(S):
    n, p = 5000, 2_000_000
    spec, _ = synth.nmf_spec(n, seed=5)
    _log(f"[workload] NMF n={n} p={p}: generating the panel in HBM")
    return S.BedFile.synthetic_panel(n, p, **spec)


def test_nmf_full_size_invariants(S, nmf_panel):
    g = nmf_panel
    cfg = S.NmfParams(k=8, n_iter=10, cv_rate=5e-4, seed=42, chunk_size=10_000)
    _log("[workload] NMF: estimation on every SNP")
    got = S.nmf_impute_estimation(g, None, cfg)
    eps = 1e-16
    assert got.w.shape == (5000, 8) and got.h_full.shape == (8, 2_000_000)
    assert got.w.min() >= eps and got.h_full.min() >= eps
    assert np.allclose(got.w.sum(axis=0), 1.0, rtol=1e-12)
    cv = np.asarray(got.cv_mae_per_iter)
    assert np.all(np.isfinite(cv)) and cv[-1] < got.cv_mae_mean_impute
    # cv_rate of the observed entries are held out each iteration (the draws
    # cover every entry, missing ones are skipped); about 10 % are missing
    obs_frac = float(np.mean(~np.isnan(g.decode_raw(np.arange(0, 2_000_000, 2000)))))
    assert 0.08 < 1.0 - obs_frac < 0.12
    expect = 5e-4 * 5000 * 2_000_000 * obs_frac
    assert abs(got.total_mask_count / 10 - expect) < 0.03 * expect
    _log(f"[workload] NMF: cv MAE {cv[-1]:.4f} vs mean-impute {got.cv_mae_mean_impute:.4f}, "
         f"{got.total_mask_count} held out")
    _log("[workload] NMF: estimation with both passes on the int8 matrix cores")
    i8 = S.nmf_impute_estimation(g, None, cfg, int8=True)
    dw = np.max(np.abs(i8.w - got.w)) / np.abs(got.w).max()
    dh = np.max(np.abs(i8.h_full - got.h_full)) / np.abs(got.h_full).max()
    _log(f"[workload] NMF int8 vs FP64: max |dW| {dw:.2e}, max |dH| {dh:.2e} of the largest entries")
    assert dw <= 1e-5 and dh <= 1e-5
    assert i8.total_mask_count == got.total_mask_count
    assert np.allclose(i8.cv_mae_per_iter, cv, rtol=1e-5, atol=0)


_NMF_SUB = dict(n=5000, m=3000, k=8, iters=10, seed=42, cv_rate=5e-4, chunk=1000)


@pytest.fixture(scope="module")
def nmf_subset_ref(nmf_panel):
    c = _NMF_SUB
    packed = nmf_panel.read_packed(0, c["m"]).tobytes()
    omega = synth.normal_matrix(c["seed"], c["n"], c["k"] + 10)
    _log(f"[workload] NMF subset: oracle on {c['m']} SNPs")
    ref = O.nmf_impute_estimation(packed, (c["n"] + 3) // 4, c["n"], np.arange(c["m"]), c["k"], c["iters"], 1e-16,
                                  c["cv_rate"], c["chunk"], c["seed"], omega, synth.uniform_stream)
    return ref, omega


@pytest.mark.parametrize("int8", [False, True])
def test_nmf_subset_vs_oracle(S, nmf_panel, nmf_subset_ref, int8):
    g = nmf_panel
    ref, omega = nmf_subset_ref
    tol, tol_cv = (1e-5, 1e-5) if int8 else (1e-8, 1e-9)
    m, k, iters, seed = _NMF_SUB["m"], _NMF_SUB["k"], _NMF_SUB["iters"], _NMF_SUB["seed"]
    cv_rate, chunk = _NMF_SUB["cv_rate"], _NMF_SUB["chunk"]
    _log(f"[workload] NMF subset: GPU estimation ({'int8 passes' if int8 else 'FP64'})")
    cfg = S.NmfParams(k=k, n_iter=iters, cv_rate=cv_rate, seed=seed, chunk_size=chunk)
    got = S.nmf_impute_estimation(g, np.arange(m), cfg, omega=omega, int8=int8)
    scale_w, scale_h = np.abs(ref.w).max(), np.abs(ref.h_full).max()
    assert np.max(np.abs(got.w - ref.w)) <= tol * scale_w
    assert np.max(np.abs(got.h_full - ref.h_full)) <= tol * scale_h
    assert got.total_mask_count == ref.total_mask_count > 0
    assert np.allclose(got.cv_mae_per_iter, ref.cv_mae_per_iter, rtol=tol_cv, atol=0)
    assert np.isclose(got.cv_mae_mean_impute, ref.cv_mae_mean_impute, rtol=1e-12 if not int8 else 1e-5, atol=0)


def test_nmf_fit_mixed_vs_fp64_full_size(S, nmf_panel):
    """The NMF-imputed fit (K = 8, d = 2, nmf-k = 8) in mixed precision (int8
    NMF passes, filled sketch/backward, forward and association passes)
    against the same fit in FP64 on every SNP: |delta log10 p| <= 1e-3, the
    same GIFs to 1e-6, r2 partitions."""
    g = nmf_panel
    n = g.n_samples
    x = np.random.default_rng(5).standard_normal((n, 2))
    cfg = S.NmfParams(k=8, n_iter=10, cv_rate=5e-4, seed=42, chunk_size=10_000)
    base = dict(k=8, norm=S.EIGENSTRAT, chunk_size=10_000, n_power_iter=2, seed=42)
    _log("[workload] NMF fit: FP64")
    f64 = S.fit_lfmm2_nmf(g, None, g, x, S.Lfmm2Config(**base), cfg)
    _log("[workload] NMF fit: mixed precision")
    mix = S.fit_lfmm2_nmf(g, None, g, x, S.Lfmm2Config(precision=S.PREC_I8, **base), cfg, nmf_int8=True)
    lp = np.abs(O.log10_pvalue(mix.tstat, mix.gif_per_trait) - O.log10_pvalue(f64.tstat, f64.gif_per_trait))
    _log(f"[workload] NMF fit mixed vs FP64: max |dlog10p| {lp.max():.2e}")
    assert lp.max() <= 1e-3
    assert np.allclose(mix.gif_per_trait, f64.gif_per_trait, rtol=1e-6)
    assert _r2_partition_ok(mix.r2)


# ---------------------------------------------------------------------------
# inbred lines at full size
# ---------------------------------------------------------------------------

def test_inbred_full_size(S):
    n, p, K = 1135, 1_800_000, 12
    spec, anc = synth.inbred_spec(n, seed=2)
    _log(f"[workload] inbred n={n} p={p} K={K}: generating the panel in HBM")
    g = S.BedFile.synthetic_panel(n, p, **spec)
    rng = np.random.default_rng(2)
    # trait: 0.4 x the first ancestry axis + 30 causal SNPs + N(0, 1)
    ac = anc - anc.mean(axis=0)
    axis = ac @ np.linalg.svd(ac, full_matrices=False)[2][0]
    axis = (axis - axis.mean()) / axis.std()
    causal = np.sort(rng.choice(p, size=30, replace=False))
    gc = np.stack([g.decode_raw([j])[:, 0] for j in causal], axis=1)
    gc = np.where(np.isnan(gc), np.nanmean(gc, axis=0), gc)
    gs = (gc - gc.mean(axis=0)) / np.maximum(gc.std(axis=0), 1e-12)
    y = 0.4 * axis + gs @ np.full(30, 0.15) + rng.standard_normal(n)
    x = y[:, None]
    base = dict(k=K, oversampling=10, n_power_iter=2, norm=S.EIGENSTRAT, seed=42)
    f64 = S.fit_lfmm2(g, "all", g, x, S.Lfmm2Config(precision=S.PREC_FP64, **base))
    mix = S.fit_lfmm2(g, "all", g, x, S.Lfmm2Config(precision=S.PREC_I8, **base))
    dlp = np.abs(O.log10_pvalue(mix.tstat, mix.gif_per_trait) - O.log10_pvalue(f64.tstat, f64.gif_per_trait))
    _log(f"[workload] inbred: max |dlog10p| mixed vs FP64 = {dlp.max():.3e}")
    assert dlp.max() <= 1e-3
    assert _r2_partition_ok(mix.r2)
    j0, m = 900_000, 6_000
    packed = g.read_packed(j0, m)
    xs = O.center_covariates(x)
    ref = O.test_associations_fused(packed.tobytes(), packed.shape[1], n, m, xs, f64.u_hat, _ridge_only(xs), K, 2000,
                                    O.EIGENSTRAT, O.Impute())
    assert np.allclose(f64.tstat[j0:j0 + m], ref.t, rtol=1e-8, atol=1e-9)
    assert np.allclose(f64.beta[j0:j0 + m], ref.beta, rtol=1e-8, atol=1e-12)
    # the generated panel is what the options ask for: mostly homozygous, ~5 % missing
    raw = g.decode_raw(np.arange(0, p, 3000))
    obs = ~np.isnan(raw)
    het = np.mean(raw[obs] == 1.0)
    assert 0.03 < het < 0.07
    assert 0.03 < 1.0 - obs.mean() < 0.07
