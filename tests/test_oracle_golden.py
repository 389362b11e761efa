"""Pin the CPU oracle against the reference's own known-answer vectors
(tests/golden/reference_kat.json) and against what the mathematics fixes
(brute force, closed forms, invariants)."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")))


def test_decode_roundtrip_eigenstrat():
    c = KAT["decode_roundtrip"]
    g = np.array(c["genotypes"], dtype=np.uint8)
    n, p = g.shape
    packed = O.encode_bed(g)
    dec = O.decode_chunk(packed, O.bytes_per_snp(n), n, np.arange(p), O.EIGENSTRAT)
    for j in range(p):
        col = g[:, j].astype(float)
        mean = col.sum() / n
        af = mean / 2
        exp = (col - mean) / math.sqrt(2 * af * (1 - af))
        assert np.max(np.abs(dec[:, j] - exp)) < c["tol"]


def test_missing_code():
    c = KAT["missing_code"]
    dec = O.decode_chunk(bytes(c["packed"]), 1, c["n"], [0], O.EIGENSTRAT)[:, 0]
    assert np.max(np.abs(dec - np.array(c["expected"]))) < c["tol"]


def test_norm_modes():
    c = KAT["norm_modes"]
    for key, norm in (("center_only", O.CENTER_ONLY), ("eigenstrat", O.EIGENSTRAT)):
        dec = O.decode_chunk(bytes(c["packed"]), 1, c["n"], [0], norm)[:, 0]
        assert np.max(np.abs(dec - np.array(c[key]))) < c["tol"]


def test_decode_bruteforce_with_keep_and_fill():
    rng = np.random.default_rng(0)
    pan = synth.balding_nichols(37, 11, miss_rate=0.2, seed=3)
    packed = O.encode_bed(pan.g)
    bps = O.bytes_per_snp(37)
    keep = rng.permutation(37)[:23]
    fill = rng.random((23, 11)) * 2
    got = O.decode_chunk(packed, bps, 37, np.arange(11), O.EIGENSTRAT, keep, fill)
    for j in range(11):
        col = [float(pan.g[i, j]) if pan.g[i, j] != 255 else float("nan") for i in keep]
        col = [fill[r, j] if math.isnan(v) else v for r, v in enumerate(col)]
        s = 0.0
        for v in col:
            s += v
        mean = s / len(col)
        pp = mean / 2.0
        tw = 2.0 * pp * (1.0 - pp)
        sc = 1.0 / math.sqrt(tw) if math.sqrt(tw) > 1e-10 else 0.0
        exp = [(v - mean) * sc for v in col]
        assert np.array_equal(got[:, j], np.array(exp))   # bit-exact


def test_subset_indices():
    for c in KAT["subset_indices"]:
        spec = c["spec"]
        got = O.subset_indices(spec, c["n"])
        assert list(got) == c["idx"]
        assert O.subset_snp_count(spec, c["n"]) == len(c["idx"])


def test_precompute_invariants():
    c = KAT["precompute"]
    x = np.array(c["x"]); lam = c["lambda"]; tol = c["tol"]
    pre = O.precompute(x, lam)
    assert np.allclose(pre.q_full.T @ pre.q_full, np.eye(4), atol=tol)
    s = np.linalg.svd(x, compute_uv=False)
    assert np.allclose(pre.d_lambda[:2] ** 2, lam / (lam + s ** 2), atol=tol)
    assert np.allclose(pre.d_lambda[2:], 1.0, atol=tol)
    assert np.allclose(pre.d_lambda * pre.d_lambda_inv, 1.0, atol=tol)
    assert np.allclose((x.T @ x + lam * np.eye(2)) @ pre.ridge_inv, np.eye(2), atol=tol)
    with pytest.raises(ValueError):
        O.precompute(x, 0.0)


def test_safe_inv():
    c = KAT["safe_inv"]
    a = np.array(c["spd"])
    assert np.allclose(a @ O.safe_inv(a), np.eye(2), atol=c["spd_tol"])
    s = np.array(c["singular"])
    inv = O.safe_inv(s)
    assert np.allclose((s + c["singular_eps"] * np.eye(2)) @ inv, np.eye(2), atol=c["singular_tol"])


def test_median_and_histogram():
    for c in KAT["median"]:
        h = O.TsqHistogram(1)
        # feed values whose t^2 lands in distinct bins: median semantics of the bin walk
        v = np.array(c["v"])
        h.add_chunk(np.sqrt(v)[:, None])
        med = h.median_for_trait(0)
        assert abs(med - c["m"]) < 2e-3
    c = KAT["histogram"]
    h = O.TsqHistogram(1)
    h.add_chunk(np.array(c["t"])[:, None])
    gif, _ = h.compute_gif()
    assert abs(gif[0] - c["gif"]) < c["tol"]


def test_histogram_median_bruteforce():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 100, 1001, 2000):
        t = rng.standard_normal((n, 2)) * 1.5
        h = O.TsqHistogram(2)
        h.add_chunk(t)
        for j in range(2):
            bins = np.sort(np.minimum((t[:, j] ** 2) / 0.001, 99999).astype(int))
            if n % 2:
                exp = (bins[n // 2] + 0.5) * 0.001
            else:
                exp = ((bins[n // 2 - 1] + 0.5) * 0.001 + (bins[n // 2] + 0.5) * 0.001) / 2
            assert h.median_for_trait(j) == pytest.approx(exp, abs=1e-15)


def test_calibrate_closed_form_and_order():
    c = KAT["calibrate_order"]
    p = O.calibrate(np.array(c["t"])[:, None], np.array([c["gif_sqrt"] ** 2]))
    assert p[0, 0] < p[1, 0]
    from scipy.stats import norm
    t = np.linspace(-8, 8, 33)[:, None]
    assert np.allclose(O.calibrate(t, np.array([1.7])), 2 * norm.cdf(-np.abs(t / np.sqrt(1.7))), rtol=1e-12)
    lp = O.log10_pvalue(t, np.array([1.7]))
    assert np.allclose(lp, np.log10(O.calibrate(t, np.array([1.7]))), atol=1e-12)


def _toy(n=60, p=40, k=3, d=2, seed=11, miss=0.05):
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=miss, d=d, seed=seed, n_causal=3, effect=1.0)
    return pan, O.encode_bed(pan.g), O.bytes_per_snp(n)


def test_association_bruteforce_lstsq():
    pan, packed, bps = _toy()
    n, p = pan.g.shape
    k, d = 3, pan.x.shape[1]
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    u = np.random.default_rng(1).standard_normal((n, k))
    res = O.test_associations_fused(packed, bps, n, p, xs, u, pre, k, 7, O.EIGENSTRAT, O.Impute())
    y = O.decode_chunk(packed, bps, n, np.arange(p), O.EIGENSTRAT)
    c = np.hstack([np.ones((n, 1)), xs, u])
    df = n - 1 - d - k
    for j in range(p):
        coef, *_ = np.linalg.lstsq(c, y[:, j], rcond=None)
        r = y[:, j] - c @ coef
        s2 = r @ r / df
        cov = np.linalg.inv(c.T @ c) * s2
        t = coef[1:1 + d] / np.sqrt(np.diag(cov)[1:1 + d])
        assert np.allclose(res.t[j], t, rtol=1e-9, atol=1e-12)
        # ridge effect: B = (X^T X + lam I)^-1 X^T (I - P_U) y
        pu = u @ np.linalg.inv(u.T @ u) @ u.T
        b = np.linalg.solve(xs.T @ xs + 1e-5 * np.eye(d), xs.T @ (y[:, j] - pu @ y[:, j]))
        assert np.allclose(res.beta[j], b, rtol=1e-9, atol=1e-12)
        tss = y[:, j] @ y[:, j]
        if tss > 1e-300:
            assert abs(res.r2[j].sum() - 1.0) < 1e-9 or res.r2[j, 1] == 0.0
    assert np.all(np.isfinite(res.p))


def test_rsvd_matches_exact_svd_when_sketch_is_full():
    # l = n: the sketch spans everything, HMT is exact up to rounding
    pan, packed, bps = _toy(n=24, p=50, seed=4)
    n, p = pan.g.shape
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    k, os_ = 4, 20
    omega = np.random.default_rng(2).standard_normal((n, k + os_))
    u_hat = O.estimate_factors_streaming(packed, bps, n, np.arange(p), pre, k, os_, 1, 9,
                                         O.EIGENSTRAT, O.Impute(), omega=omega)
    y = O.decode_chunk(packed, bps, n, np.arange(p), O.EIGENSTRAT)
    u_ex = np.linalg.svd(pre.m @ y, full_matrices=False)[0][:, :k]
    u_ref = pre.q_full @ (pre.d_lambda_inv[:, None] * u_ex)
    ang = O.principal_angles(u_hat, u_ref)
    assert np.max(ang) < 1e-8


def test_qr_invariants():
    a = np.random.default_rng(0).random((10, 3))
    q, _ = np.linalg.qr(a)
    assert np.allclose(q.T @ q, np.eye(3), atol=1e-10)
    b = np.fromfunction(lambda i, j: i + j, (5, 2))
    q, _ = np.linalg.qr(b)
    assert np.linalg.norm(b - q @ (q.T @ b)) < 1e-10


def test_nmf_dense_reference_thresholds():
    c = KAT["nmf_dense"]
    rng = np.random.default_rng(42)
    for key, iters in (("y", c["n_iter"]), ("y_missing", c["n_iter_missing"])):
        y = np.array([[np.nan if v is None else v for v in row] for row in c[key]])
        w0 = rng.random((4, c["k"])); h0 = rng.random((c["k"], 3)) * 2
        w, h = O.nmf_dense(y, c["k"], iters, c["eps"], w0, h0)
        assert (w >= 0).all()
        m = ~np.isnan(y)
        assert float(((y - w @ h)[m] ** 2).sum()) < c["max_err"]


def test_streaming_nmf_reduces_cv_error_and_matches_dense_fill_semantics():
    pan, packed, bps = _toy(n=50, p=80, seed=9, miss=0.1)
    n, p = pan.g.shape
    k = 3
    omega = synth.normal_matrix(1_000_042, n, k + 10)
    res = O.nmf_impute_estimation(packed, bps, n, np.arange(p), k, 6, 1e-16, 0.2, 16, 42,
                                  omega, synth.uniform_stream)
    assert res.w.min() > 0 and res.h_full.min() > 0
    assert np.allclose(res.w.sum(axis=0), 1.0)
    assert res.cv_mae_per_iter[-1] < res.cv_mae_mean_impute + 0.2
    assert res.total_mask_count > 0
