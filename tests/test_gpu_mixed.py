"""GPU parity of the mixed-precision path (int8 MFMA on exact genotypes,
multi-term int8 dense operands, exact int32 accumulation, FP64 epilogue)
against the CPU oracle.  Tolerances: |delta log10 p| <= 1e-3 (the metric's
bound); t within 2e-4 absolute; U_hat: separated directions within 2e-5 rad,
captured variance within 1e-6 (noise-band directions are bounded by the
quantisation error over the spectral gap, see the RSVD test)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def test_mfma_i8_lane_layout(S):
    """One v_mfma_i32_16x16x64_i8 with the kernels' fragment mapping equals A @ B exactly."""
    from schnellfmm_amd._lib import lib
    rng = np.random.default_rng(0)
    ctx = S.api.default_context()
    for _ in range(3):
        a = rng.integers(-127, 128, size=(16, 64), dtype=np.int8)
        b = rng.integers(-127, 128, size=(64, 16), dtype=np.int8)
        d = np.zeros((16, 16), np.int32)
        ctx.check(lib().slfmm_debug_mfma_i8(ctx.h, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                            d.ctypes.data_as(C.c_void_p)))
        assert np.array_equal(d, a.astype(np.int64) @ b.astype(np.int64))


def _panel(n, p, seed, d=2, miss=0.02, n_causal=10, effect=0.8, **kw):
    pan = synth.balding_nichols(n, p, n_pops=4, miss_rate=miss, d=d, n_causal=n_causal, effect=effect,
                                seed=seed, **kw)
    return pan, O.encode_bed(pan.g), O.bytes_per_snp(n)


@pytest.mark.parametrize("n,p,K,d,norm,miss", [
    (300, 3000, 5, 2, 1, 0.02),
    (257, 1037, 4, 1, 0, 0.10),     # ragged sample and SNP tails, heavy missingness
    (1100, 2000, 12, 1, 1, 0.05),
    (640, 1500, 20, 4, 1, 0.02),    # K + d = 24 (headline shape)
    (400, 1200, 22, 6, 1, 0.03),    # K + d = 28: 8 row tiles, 32-functional epilogue
    (400, 1100, 24, 8, 0, 0.03),    # K + d = 32: 9 row tiles, the unpipelined kernel
])
def test_i8_association_vs_oracle(S, n, p, K, d, norm, miss):
    pan, packed, bps = _panel(n, p, seed=n + p, d=d, miss=miss)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    omega = np.random.default_rng(1).standard_normal((n, K + 10))
    u = O.estimate_factors_streaming(packed, bps, n, np.arange(p), pre, K, 10, 2, 500, norm, O.Impute(),
                                     omega=omega)
    ref = O.test_associations_fused(packed, bps, n, p, xs, u, pre, K, 500, norm, O.Impute())
    g = S.BedFile.from_packed(packed, n, p)
    got = S.test_associations_fused(g, xs, u, S.Lfmm2Config(k=K, norm=norm, precision=S.PREC_I8))
    assert np.max(np.abs(got.tstat - ref.t)) < 2e-4
    assert np.allclose(got.beta, ref.beta, rtol=1e-3, atol=1e-4 * np.max(np.abs(ref.beta)))
    assert np.max(np.abs(got.r2 - ref.r2)) < 1e-4
    dlp = np.abs(O.log10_pvalue(got.tstat, got.gif_per_trait) - O.log10_pvalue(ref.t, ref.gif_per_trait))
    assert np.max(dlp) <= 1e-3


def test_i8_association_edge_snps(S):
    n, p = 130, 9
    g = np.random.default_rng(2).integers(0, 3, size=(n, p)).astype(np.uint8)
    g[:, 1] = 2
    g[:, 2] = 255
    g[:, 3] = 0
    g[::2, 4] = 255
    g[:, 5] = 1
    g[0, 5] = 255
    packed = O.encode_bed(g)
    rng = np.random.default_rng(1)
    xs = O.center_covariates(rng.standard_normal((n, 2)))
    u = rng.standard_normal((n, 3))
    pre = O.precompute(xs, 1e-5)
    for norm in (0, 1):
        ref = O.test_associations_fused(packed, O.bytes_per_snp(n), n, p, xs, u, pre, 3, 4, norm, O.Impute())
        G = S.BedFile.from_packed(packed, n, p)
        got = S.test_associations_fused(G, xs, u, S.Lfmm2Config(k=3, norm=norm, precision=S.PREC_I8))
        assert np.max(np.abs(got.tstat - ref.t)) < 2e-4
        assert np.all(got.tstat[[1, 2, 3, 5]] == 0.0)
        assert np.max(np.abs(got.r2 - ref.r2)) < 1e-4


def test_i8_association_singular_design(S):
    """Duplicated latent column (integer design, exactly singular C^T C): the
    fallback basis [U_c|X] is quantised column by column, so the duplicate
    columns stay identical and the quadratic-form RSS cancels the null
    direction exactly."""
    n, p = 300, 400
    pan, packed, bps = _panel(n, p, seed=17, d=2)
    rng = np.random.default_rng(3)
    xs = rng.integers(-3, 4, size=(n, 2)).astype(np.float64)
    xs[-1] -= xs.sum(axis=0)
    u = rng.integers(-3, 4, size=(n, 3)).astype(np.float64)
    u[:, 2] = u[:, 0]
    pre = O.precompute(xs, 1e-5)
    ref = O.test_associations_fused(packed, bps, n, p, xs, u, pre, 3, 100, 1, O.Impute())
    g = S.BedFile.from_packed(packed, n, p)
    got = S.test_associations_fused(g, xs, u, S.Lfmm2Config(k=3, norm=1, precision=S.PREC_I8))
    assert np.max(np.abs(got.tstat - ref.t)) < 2e-4
    assert np.max(np.abs(got.r2 - ref.r2)) < 1e-4


def _captured(packed, bps, n, idx, norm, u):
    """Variance of the standardised estimation panel captured by span(u)."""
    y = O.decode_chunk(packed, bps, n, idx, norm)
    q, _ = np.linalg.qr(u)
    return float(np.sum((q.T @ y) ** 2))


def test_i8_rsvd_vs_oracle(S):
    """Mixed-precision RSVD against the oracle driven by the same sketch.

    The panel has three axes of population structure (singular values ~240)
    above a flat noise band (~71, gaps ~0.2); K = 5 asks for two directions
    inside that band.  Those are defined only up to ~eps / gap: two FP64 runs
    that differ only in the random sketch disagree on them by ~1.4 rad and on
    the captured variance by ~1.6e-3.  The int8 operand error (2 terms, ~2e-5
    relative) moves them by < 1e-3 rad (CPU emulation of the quantised
    pipeline: 8e-5 and 3.3e-4), so the bounds are: separated directions
    < 2e-5 rad, every direction < 1e-3 rad, captured variance within 1e-6."""
    n, p, K, os_ = 400, 3000, 5, 10
    pan, packed, bps = _panel(n, p, seed=31, d=2)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ np.random.default_rng(5).standard_normal((n, K + os_))
    for norm, rate in ((1, None), (0, 0.5)):
        idx = np.arange(p) if rate is None else O.subset_indices(rate, p)
        ref = O.estimate_factors_streaming(packed, bps, n, idx, pre, K, os_, 2, 700, norm, O.Impute(), b0=b0)
        g = S.BedFile.from_packed(packed, n, p)
        got = S.estimate_factors_streaming(g, rate, xs, S.Lfmm2Config(k=K, oversampling=os_, norm=norm,
                                                                      precision=S.PREC_I8), b0=b0)
        ang = np.sort(O.principal_angles(got, ref))
        assert np.max(ang[:3]) < 2e-5          # the three separated structure axes
        assert np.max(ang) < 1e-3              # noise-band directions
        e_got, e_ref = _captured(packed, bps, n, idx, norm, got), _captured(packed, bps, n, idx, norm, ref)
        assert abs(e_got / e_ref - 1.0) < 1e-6


def test_i8_fit_end_to_end_log10p_bound(S):
    """Whole fit in mixed precision: |delta log10 p| <= 1e-3 against the
    oracle (the headline accuracy bound), captured variance within 1e-6."""
    n, p, K = 700, 6000, 8
    pan, packed, bps = _panel(n, p, seed=41, d=2, n_causal=30, effect=1.2)
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ np.random.default_rng(9).standard_normal((n, K + 10))
    cfg_o = O.Config(k=K, norm=O.EIGENSTRAT, chunk_size=1000, n_power_iter=2)
    idx = O.subset_indices(0.5, p)
    u_ref, ref, _, _ = O.fit_lfmm2(packed, bps, n, p, pan.x, cfg_o, idx, b0=b0)
    g = S.BedFile.from_packed(packed, n, p)
    got = S.fit_lfmm2(g, 0.5, g, pan.x, S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, precision=S.PREC_I8), b0=b0)
    e_got = _captured(packed, bps, n, idx, O.EIGENSTRAT, got.u_hat)
    e_ref = _captured(packed, bps, n, idx, O.EIGENSTRAT, u_ref)
    assert abs(e_got / e_ref - 1.0) < 1e-6
    dlp = np.abs(O.log10_pvalue(got.tstat, got.gif_per_trait) - O.log10_pvalue(ref.t, ref.gif_per_trait))
    assert np.max(dlp) <= 1e-3
    assert np.max(np.abs(ref.t)) > 5       # the panel has real signal


def test_i8_wide_shapes_fall_back_to_fp64_kernels(S, capfd):
    """Past 32 int8 columns the mixed-precision fit runs the FP64 kernels for
    that stage: K + oversampling > 32 gives the FP64 sketch (so U_hat equals
    the FP64 fit's bit for bit), and K + d > 32 the FP64 association too.
    Each such stage is reported once per context on stderr."""
    n, p = 500, 2500
    pan, packed, bps = _panel(n, p, seed=77, d=1, miss=0.03)
    g = S.BedFile.from_packed(packed, n, p, ctx=S.Context())   # a fresh context: its notes start empty
    capfd.readouterr()
    S.fit_lfmm2(g, 0.5, g, pan.x, S.Lfmm2Config(k=20, norm=S.EIGENSTRAT, precision=S.PREC_I8))
    assert "SLFMM_PREC_I8 requested" not in capfd.readouterr().err   # every stage in int8
    for K in (30, 40):
        f64 = S.fit_lfmm2(g, 0.5, g, pan.x, S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, precision=S.PREC_FP64))
        mix = S.fit_lfmm2(g, 0.5, g, pan.x, S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, precision=S.PREC_I8))
        assert np.array_equal(mix.u_hat, f64.u_hat)
        if K + 1 > 32:
            assert np.array_equal(mix.tstat, f64.tstat)
        dlp = np.abs(O.log10_pvalue(mix.tstat, mix.gif_per_trait) - O.log10_pvalue(f64.tstat, f64.gif_per_trait))
        assert np.max(dlp) <= 1e-3
        err = capfd.readouterr().err
        if K == 30:   # first wide fit: the sketch note (40 columns); the association has 31
            assert "randomized SVD runs in FP64 (K + oversampling = 40 > 32 columns)" in err
            assert "association runs in FP64" not in err
        else:         # the sketch note is not repeated; the association (41 columns) is new
            assert "randomized SVD runs in FP64" not in err
            assert "association runs in FP64 (K + d = 41 > 32 columns)" in err
