"""GPU parity: the HIP library through its C ABI vs the CPU oracle on the
same seeded inputs (sizes the oracle finishes in seconds)."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu
KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")))


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def _panel(n=300, p=2000, miss=0.02, seed=1, d=2, n_causal=10, **kw):
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=miss, d=d, n_causal=n_causal, effect=0.8,
                                seed=seed, **kw)
    return pan, O.encode_bed(pan.g), O.bytes_per_snp(n)


def test_decode_reference_kat(S):
    c = KAT["missing_code"]
    g = S.BedFile.from_packed(bytes(c["packed"]), c["n"], 1)
    assert np.max(np.abs(g.decode([0], S.EIGENSTRAT)[:, 0] - np.array(c["expected"]))) < c["tol"]
    c = KAT["norm_modes"]
    g = S.BedFile.from_packed(bytes(c["packed"]), c["n"], 1)
    assert np.max(np.abs(g.decode([0], S.CENTER_ONLY)[:, 0] - np.array(c["center_only"]))) < c["tol"]
    assert np.max(np.abs(g.decode([0], S.EIGENSTRAT)[:, 0] - np.array(c["eigenstrat"]))) < c["tol"]
    c = KAT["decode_roundtrip"]
    gt = np.array(c["genotypes"], np.uint8)
    g = S.BedFile.from_packed(O.encode_bed(gt), 7, 3)
    assert np.array_equal(g.decode(range(3), S.EIGENSTRAT),
                          O.decode_chunk(O.encode_bed(gt), 2, 7, range(3), O.EIGENSTRAT))


@pytest.mark.parametrize("n", [1, 3, 4, 5, 63, 64, 65, 301])
def test_decode_bit_exact_with_keep_and_fill(S, n):
    rng = np.random.default_rng(n)
    pan, packed, bps = _panel(n=n, p=37, miss=0.2, seed=n, n_causal=0)
    idx = rng.permutation(37)[:29]
    keep = rng.permutation(n)[: max(1, n - n // 3)]
    g = S.BedFile.from_packed(packed, n, 37, sample_keep=keep)
    for norm in (S.CENTER_ONLY, S.EIGENSTRAT):
        got = g.decode(idx, norm)
        exp = O.decode_chunk(packed, bps, n, idx, norm, keep)
        assert np.array_equal(got, exp)
        fill = rng.random((len(keep), len(idx))) * 2
        assert np.array_equal(g.decode(idx, norm, fill), O.decode_chunk(packed, bps, n, idx, norm, keep, fill))
    raw = g.decode_raw(idx)
    exp = O.decode_raw(packed, bps, n, idx, keep)
    assert np.array_equal(np.isnan(raw), np.isnan(exp))
    assert np.array_equal(np.nan_to_num(raw, nan=-1), np.nan_to_num(exp, nan=-1))
    nm, sg, nh = g.stats(idx)
    assert np.array_equal(nm, np.isnan(exp).sum(0))
    assert np.array_equal(sg, np.nansum(exp, 0).astype(np.int64))
    assert np.array_equal(nh, (exp == 2).sum(0))


def test_safe_inv_reference_kat(S):
    c = KAT["safe_inv"]
    a = np.array(c["spd"])
    inv, reg = S.safe_inv(a)
    assert not reg and np.allclose(a @ inv, np.eye(2), atol=c["spd_tol"])
    s = np.array(c["singular"])
    inv, reg = S.safe_inv(s)
    assert reg
    assert np.allclose((s + c["singular_eps"] * np.eye(2)) @ inv, np.eye(2), atol=c["singular_tol"])
    r = np.random.default_rng(3).standard_normal((25, 25))
    inv, reg = S.safe_inv(r @ r.T + np.eye(25))
    assert not reg and np.allclose(inv, np.linalg.inv(r @ r.T + np.eye(25)), rtol=1e-10, atol=1e-12)


def test_precompute_vs_oracle(S):
    c = KAT["precompute"]
    for x, lam in ((np.array(c["x"]), c["lambda"]),
                   (O.center_covariates(np.random.default_rng(2).standard_normal((500, 4))), 1e-5)):
        pre = S.precompute(x, lam)
        ref = O.precompute(x, lam)
        d = x.shape[1]
        assert np.allclose(pre.sigma, ref.sigma[:d], rtol=1e-12)
        assert np.allclose(pre.d_lambda, ref.d_lambda[:d], rtol=1e-10)
        assert np.allclose(pre.ridge_inv, ref.ridge_inv, rtol=1e-10, atol=1e-14)
        assert np.allclose(pre.q_thin.T @ pre.q_thin, np.eye(d), atol=1e-10)
        # same projector as the oracle's leading singular vectors
        assert np.allclose(pre.q_thin @ pre.q_thin.T, ref.q_full[:, :d] @ ref.q_full[:, :d].T, atol=1e-10)
    with pytest.raises(Exception) as ei:
        S.precompute(np.ones((5, 1)), 0.0)
    assert "positive" in str(ei.value)


def _assoc_case(S, n, p, K, d, miss, seed, norm, chunk=97):
    pan, packed, bps = _panel(n=n, p=p, miss=miss, seed=seed, d=d)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    # latent factors: the oracle's own RSVD estimate, as in the pipeline
    omega = np.random.default_rng(seed).standard_normal((n, K + 10))
    u = O.estimate_factors_streaming(packed, bps, n, np.arange(p), pre, K, 10, 2, chunk, norm,
                                     O.Impute(), omega=omega)
    ref = O.test_associations_fused(packed, bps, n, p, xs, u, pre, K, chunk, norm, O.Impute())
    g = S.BedFile.from_packed(packed, n, p)
    cfg = S.Lfmm2Config(k=K, norm=norm)
    got = S.test_associations_fused(g, xs, u, cfg, with_hist=True)
    return ref, got


@pytest.mark.parametrize("norm", [0, 1])
def test_association_fp64_vs_oracle(S, norm):
    ref, got = _assoc_case(S, n=300, p=3000, K=5, d=2, miss=0.02, seed=7, norm=norm)
    assert np.allclose(got.beta, ref.beta, rtol=1e-9, atol=1e-12)
    assert np.allclose(got.tstat, ref.t, rtol=1e-9, atol=1e-11)
    assert np.allclose(got.r2, ref.r2, rtol=1e-9, atol=1e-12)
    # histogram: the GPU's counts are exactly the binning of its own t values;
    # every t whose t^2 is not within rounding of a bin edge lands in the
    # oracle's bin; and the oracle's median walk on the GPU's counts gives the
    # GPU's GIF exactly
    d = got.tstat.shape[1]
    own = O.TsqHistogram(d)
    own.add_chunk(got.tstat)
    assert np.array_equal(got.hist.astype(np.uint64), own.counts)
    q_ref = ref.t ** 2 / O.BIN_WIDTH
    safe = np.abs(q_ref - np.rint(q_ref)) > 1e-6 * np.maximum(q_ref, 1.0)
    b_got = np.minimum(got.tstat ** 2 / O.BIN_WIDTH, O.N_BINS - 1).astype(np.int64)
    b_ref = np.minimum(q_ref, O.N_BINS - 1).astype(np.int64)
    assert safe.mean() > 0.99 and np.array_equal(b_got[safe], b_ref[safe])
    own_gif, own_avg = own.compute_gif()
    assert np.array_equal(got.gif_per_trait, own_gif)
    if np.array_equal(got.hist.astype(np.uint64), ref.hist.counts):
        assert np.array_equal(got.gif_per_trait, ref.gif_per_trait)
    lp_got = O.log10_pvalue(got.tstat, got.gif_per_trait)
    lp_ref = O.log10_pvalue(ref.t, ref.gif_per_trait)
    assert np.max(np.abs(lp_got - lp_ref)) <= 1e-9
    assert np.allclose(got.pvalue, ref.p, rtol=1e-9, atol=1e-300)
    assert got.gif == pytest.approx(ref.gif, rel=1e-12)


def test_association_monomorphic_and_all_missing(S):
    n, p = 40, 6
    g = np.zeros((n, p), np.uint8)
    g[:, 1] = 2                       # monomorphic hom-alt
    g[:, 2] = 255                     # all missing
    g[:, 3] = np.arange(n) % 3        # informative
    g[:, 4] = np.where(np.arange(n) % 2 == 0, 255, 1)
    g[:, 5] = (np.arange(n) % 5 == 0).astype(np.uint8)
    packed = O.encode_bed(g)
    rng = np.random.default_rng(1)
    xs = O.center_covariates(rng.standard_normal((n, 2)))
    u = rng.standard_normal((n, 3))
    pre = O.precompute(xs, 1e-5)
    for norm in (0, 1):
        ref = O.test_associations_fused(packed, 10, n, p, xs, u, pre, 3, 4, norm, O.Impute())
        G = S.BedFile.from_packed(packed, n, p)
        got = S.test_associations_fused(G, xs, u, S.Lfmm2Config(k=3, norm=norm))
        assert np.allclose(got.tstat, ref.t, rtol=1e-9, atol=1e-9)
        assert np.allclose(got.r2, ref.r2, rtol=1e-9, atol=1e-9)
        assert np.all(got.tstat[[0, 1, 2, 4]] == 0.0)


def _integer_singular_design(n, seed):
    """Zero-mean integer covariates and integer factors with the last factor
    a copy of the first: every entry of C^T C is an exact integer whatever
    the summation order, so both sides meet an exactly zero pivot."""
    rng = np.random.default_rng(seed)
    xs = rng.integers(-3, 4, size=(n, 2)).astype(np.float64)
    xs[-1] -= xs.sum(axis=0)
    u = rng.integers(-3, 4, size=(n, 3)).astype(np.float64)
    u[:, 2] = u[:, 0]
    return xs, u


def test_association_singular_design_falls_back(S, capfd):
    """Duplicated latent column: C^T C and U^T U get the safe_inv epsilon
    (src/testing.rs:996-1024) and the basis falls back to [U_c|X] with the
    RSS quadratic form; the warnings go to stderr."""
    n, p = 300, 400
    pan, packed, bps = _panel(n=n, p=p, seed=17, d=2)
    xs, u = _integer_singular_design(n, 3)
    pre = O.precompute(xs, 1e-5)
    ref = O.test_associations_fused(packed, bps, n, p, xs, u, pre, 3, 100, 1, O.Impute())
    g = S.BedFile.from_packed(packed, n, p)
    got = S.test_associations_fused(g, xs, u, S.Lfmm2Config(k=3, norm=1))
    err = capfd.readouterr().err
    assert "C^T C  where C = [1 | X | U_hat] is singular" in err
    assert "U_hat^T U_hat is singular" in err
    assert np.allclose(got.tstat, ref.t, rtol=1e-6, atol=1e-7)
    assert np.allclose(got.r2, ref.r2, rtol=1e-6, atol=1e-7)
    assert np.allclose(got.beta, ref.beta, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("n,K,os_", [
    (250, 4, 10),
    (700, 8, 10),   # l = 18: two 16-column blocks of the forward pass, a partial 512-sample tile
])
def test_rsvd_fp64_vs_oracle_injected_sketch(S, n, K, os_):
    p = 1500
    pan, packed, bps = _panel(n=n, p=p, seed=11, d=2)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    omega = np.random.default_rng(5).standard_normal((n, K + os_))
    b0 = pre.m.T @ omega
    for norm, rate in ((1, None), (0, 0.5)):
        idx = np.arange(p) if rate is None else O.subset_indices(rate, p)
        ref = O.estimate_factors_streaming(packed, bps, n, idx, pre, K, os_, 2, 300, norm, O.Impute(), b0=b0)
        g = S.BedFile.from_packed(packed, n, p)
        cfg = S.Lfmm2Config(k=K, oversampling=os_, n_power_iter=2, norm=norm)
        got = S.estimate_factors_streaming(g, rate, xs, cfg, b0=b0)
        assert np.max(O.principal_angles(got, ref)) < 1e-8
        # same columns up to sign
        sgn = np.sign(np.sum(got * ref, axis=0))
        assert np.allclose(got * sgn, ref, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("case", ["l_equals_n", "few_polymorphic"])
def test_rsvd_rank_deficient_final_projection(S, case):
    """The final projection W = N Y Q_z is rank deficient when l = n (its
    columns are centred) or when fewer than l estimation SNPs are polymorphic;
    the reference's thin SVD still returns U_hat (src/rsvd.rs:160-173), and
    so must the GPU: the K leading directions against the oracle."""
    if case == "l_equals_n":
        n, p, K, os_ = 30, 400, 3, 27
        pan, packed, bps = _panel(n=n, p=p, seed=13, d=1)
        idx = np.arange(p)
    else:
        n, p, K, os_ = 120, 60, 3, 12
        pan = synth.balding_nichols(n, p, n_pops=3, fst=0.3, miss_rate=0.0, d=1, seed=14)
        g = pan.g.copy()
        g[:, 12:] = 0                       # 12 polymorphic SNPs < l = 15
        pan.g[:] = g
        packed, bps = O.encode_bed(pan.g), O.bytes_per_snp(n)
        idx = np.arange(p)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    omega = np.random.default_rng(6).standard_normal((n, K + os_))
    b0 = pre.m.T @ omega
    ref = O.estimate_factors_streaming(packed, bps, n, idx, pre, K, os_, 1, 100, 1, O.Impute(), b0=b0)
    gb = S.BedFile.from_packed(packed, n, p)
    got = S.estimate_factors_streaming(gb, None, xs, S.Lfmm2Config(k=K, oversampling=os_, n_power_iter=1, norm=1),
                                       b0=b0)
    assert np.all(np.isfinite(got))
    assert np.max(O.principal_angles(got, ref)) < 1e-6


def test_rsvd_generated_sketch_matches_oracle_with_same_stream(S):
    # The device draws Omega' (n x l, seed) and sketches with B0 = N Omega'.
    # The reference form with Omega = Q^T Omega' gives the same B0 = M^T Omega.
    n, p, K, os_ = 200, 4000, 3, 10
    pan, packed, bps = _panel(n=n, p=p, seed=12, d=1)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    g = S.BedFile.from_packed(packed, n, p)
    got = S.estimate_factors_streaming(g, None, xs, S.Lfmm2Config(k=K, n_power_iter=3, norm=1, seed=7))
    omega_dev = synth.normal_matrix(7, n, K + os_)
    ref = O.estimate_factors_streaming(packed, bps, n, np.arange(p), pre, K, os_, 3, 1000, 1, O.Impute(),
                                       omega=pre.q_full.T @ omega_dev)
    assert np.max(O.principal_angles(got, ref)) < 1e-7


def test_fit_lfmm2_end_to_end_vs_oracle(S):
    n, p, K = 320, 2500, 5
    pan, packed, bps = _panel(n=n, p=p, seed=21, d=1)
    cfg_o = O.Config(k=K, norm=O.EIGENSTRAT, chunk_size=500)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ np.random.default_rng(9).standard_normal((n, K + 10))
    est = O.subset_indices(0.5, p)
    u_ref, ref, _, _ = O.fit_lfmm2(packed, bps, n, p, pan.x, cfg_o, est, b0=b0)
    g = S.BedFile.from_packed(packed, n, p)
    got = S.fit_lfmm2(g, 0.5, g, pan.x, S.Lfmm2Config(k=K, norm=S.EIGENSTRAT), b0=b0)
    assert np.max(O.principal_angles(got.u_hat, u_ref)) < 1e-8
    assert np.allclose(got.tstat, ref.t, rtol=1e-7, atol=1e-8)
    lp = np.abs(O.log10_pvalue(got.tstat, got.gif_per_trait) - O.log10_pvalue(ref.t, ref.gif_per_trait))
    assert np.max(lp) < 1e-6


def test_bed_file_errors(S, tmp_path):
    g = np.array([[0, 1], [2, 255], [1, 1]], np.uint8)
    prefix = str(tmp_path / "t")
    O.write_plink(prefix, g)
    n, p = S.BedFile.validate(prefix + ".bed")
    assert (n, p) == (3, 2)
    bf = S.BedFile.open(prefix + ".bed")
    assert np.array_equal(np.nan_to_num(bf.decode_raw([0, 1]), nan=-1),
                          np.nan_to_num(O.decode_raw(O.encode_bed(g), 1, 3, [0, 1]), nan=-1))
    bad = tmp_path / "bad.bed"
    bad.write_bytes(b"\x00\x00\x00" + O.encode_bed(g))
    for ext in ("bim", "fam"):
        (tmp_path / f"bad.{ext}").write_text((tmp_path / f"t.{ext}").read_text())
    with pytest.raises(Exception) as ei:
        S.BedFile.validate(str(bad))
    assert "bad magic" in str(ei.value)
    trunc = tmp_path / "tr.bed"
    trunc.write_bytes(O.BED_MAGIC + O.encode_bed(g)[:-1])
    for ext in ("bim", "fam"):
        (tmp_path / f"tr.{ext}").write_text((tmp_path / f"t.{ext}").read_text())
    with pytest.raises(Exception) as ei:
        S.BedFile.validate(str(trunc))
    assert "size mismatch" in str(ei.value)
    (tmp_path / "t.fam").write_text("a b c\n")
    with pytest.raises(Exception) as ei:
        S.BedFile.validate(prefix + ".bed")
    assert "fields, expected 6" in str(ei.value)


def test_degrees_of_freedom_and_lambda_errors(S):
    pan, packed, bps = _panel(n=8, p=20, seed=3, d=2)
    g = S.BedFile.from_packed(packed, 8, 20)
    xs = O.center_covariates(pan.x)
    with pytest.raises(Exception) as ei:
        S.test_associations_fused(g, xs, np.ones((8, 6)), S.Lfmm2Config(k=6))
    assert "Insufficient degrees of freedom" in str(ei.value)
    with pytest.raises(Exception) as ei:
        S.test_associations_fused(g, xs, np.random.default_rng(0).random((8, 2)), S.Lfmm2Config(k=2, lam=0.0))
    assert "positive" in str(ei.value)
