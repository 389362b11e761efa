"""Pin the oracle's streaming NMF, its random-probe initialisation, the
GWAS-phase CV and the sequential r2 decomposition against brute-force
restatements written as per-entry loops on tiny panels.

Each brute force follows the reference's order of operations as written:
  * NMF iteration (src/nmf.rs:108-232): forward fill from the pre-update
    W.H, W update only where denom > eps, clamp, normalise; backward fill
    from the post-W-update W and the normalised H, H update with a clamp on
    every entry, normalise; then the CV pass.
  * init_from_random_probe (src/nmf.rs:279-362): raw decode with mean fill,
    Z = Y^T Omega, W = |Omega Vt[:k]^T|, H = max(0, (W^T W)^-1 W^T Y).
  * compute_cv_error_masked (src/nmf.rs:379-455): one draw per entry
    (missing ones included), column then row, per-chunk seed.
  * GWAS-phase CV (src/testing.rs:423-487): draws only for observed entries,
    leave-masked-out column means, H = max(0, W^+ masked), pred = W H.
  * r2 (src/testing.rs:311-317, :390-405): the Type-I sums of squares of
    nested least-squares fits [1|X] then [1|X|U].

Every pin is paired with deliberate mutations of the brute force (the
mistakes a restatement could plausibly make); each mutation must move the
result by far more than the tolerance, so the pins can fail.
"""
import math

import numpy as np
import pytest

import oracle as O
from schnellfmm_amd import synth

EPS = 1e-16


# ---------------------------------------------------------------------------
# per-entry helpers (independent of the oracle's vectorised decode)
# ---------------------------------------------------------------------------

def _raw_matrix(packed: bytes, bps: int, n: int, idx):
    """{0,1,2,nan} list-of-lists n x len(idx), decoded one entry at a time."""
    table = (0.0, float("nan"), 1.0, 2.0)
    y = [[0.0] * len(idx) for _ in range(n)]
    for c, j in enumerate(idx):
        for i in range(n):
            code = (packed[j * bps + i // 4] >> (2 * (i % 4))) & 3
            y[i][c] = table[code]
    return y


def _is_nan(v):
    return v != v


def _col_means_observed(y, cols):
    n = len(y)
    out = []
    for c in cols:
        s, cnt = 0.0, 0
        for i in range(n):
            if not _is_nan(y[i][c]):
                s += y[i][c]
                cnt += 1
        out.append(s / cnt if cnt else 0.0)
    return out


def _matmul(a, b):
    n, m, p = len(a), len(b), len(b[0])
    return [[sum(a[i][t] * b[t][j] for t in range(m)) for j in range(p)] for i in range(n)]


def _transpose(a):
    return [list(r) for r in zip(*a)]


def _normalize(w, h):
    n, k = len(w), len(w[0])
    for j in range(k):
        s = 0.0
        for i in range(n):
            s += w[i][j]
        if s > EPS:
            for i in range(n):
                w[i][j] *= 1.0 / s
            for c in range(len(h[j])):
                h[j][c] *= s


def _panel(n=11, p=9, miss=0.15, seed=3):
    pan = synth.balding_nichols(n, p, n_pops=2, miss_rate=miss, seed=seed)
    return O.encode_bed(pan.g), O.bytes_per_snp(n)


# ---------------------------------------------------------------------------
# NMF iterations
# ---------------------------------------------------------------------------

def bf_nmf(packed, bps, n, p, w0, h0, n_iter, cv_rate, chunk, seed, mutation=None):
    """Masked multiplicative updates, dense, in the reference's order."""
    y = _raw_matrix(packed, bps, n, range(p))
    k = len(w0[0])
    w = [[max(v, EPS) for v in r] for r in w0]
    h = [[max(v, EPS) for v in r] for r in h0]
    _normalize(w, h)
    cv = []
    mean_mae, total = 0.0, 0

    col_mean = _col_means_observed(y, range(p))

    def filled(ww, hh):
        wh = _matmul(ww, hh)
        if mutation == "fill_with_column_mean":
            wh = [[col_mean[j] for j in range(p)] for _ in range(n)]
        return [[wh[i][j] if _is_nan(y[i][j]) else y[i][j] for j in range(p)] for i in range(n)]

    def w_step(y_fill):
        num = _matmul(y_fill, _transpose(h))
        hht = _matmul(h, _transpose(h))
        den = _matmul(w, hht)
        for i in range(n):
            for a in range(k):
                if mutation == "w_denominator_plus_eps":
                    w[i][a] *= num[i][a] / (den[i][a] + EPS)
                elif den[i][a] > EPS:
                    w[i][a] *= num[i][a] / den[i][a]
                w[i][a] = max(w[i][a], EPS)
        _normalize(w, h)

    def h_step(y_fill):
        num = _matmul(_transpose(w), y_fill)
        den = _matmul(_matmul(_transpose(w), w), h)
        for a in range(k):
            for j in range(p):
                if den[a][j] > EPS:
                    h[a][j] *= num[a][j] / den[a][j]
                h[a][j] = max(h[a][j], EPS)
        _normalize(w, h)

    for it in range(n_iter):
        model_before = _matmul(w, h)
        pre = filled(w, h)                      # clone of W, H before the W update
        if mutation == "h_before_w":
            h_step(pre)
            w_step(filled(w, h))
        else:
            w_step(pre)
            h_step(pre if mutation == "bwd_fill_from_pre_update" else filled(w, h))
        # CV pass: chunk by chunk, one draw per entry (missing included), col then row
        mae_sum, mean_sum, cnt = 0.0, 0.0, 0
        recon = model_before if mutation == "cv_from_pre_iteration_model" else _matmul(w, h)
        for seq, start in enumerate(range(0, p, chunk)):
            cols = list(range(start, min(p, start + chunk)))
            means = _col_means_observed(y, cols)
            draws = synth.uniform_stream(seed + it + seq, n * len(cols))
            pos = 0
            for ci, c in enumerate(cols):
                for i in range(n):
                    r = draws[pos]
                    pos += 1
                    if mutation == "cv_draws_observed_only" and _is_nan(y[i][c]):
                        pos -= 1
                        continue
                    if r < cv_rate and not _is_nan(y[i][c]):
                        mae_sum += abs(y[i][c] - recon[i][c])
                        mean_sum += abs(y[i][c] - means[ci])
                        cnt += 1
        cv.append(mae_sum / cnt if cnt else 0.0)
        mean_mae = mean_sum / cnt if cnt else 0.0
        total += cnt
    return np.array(w), np.array(h), cv, mean_mae, total


def _oracle_iterations(packed, bps, n, p, w0, h0, n_iter, cv_rate, chunk, seed):
    """The oracle's iteration loop driven from an injected (W0, H0)."""
    import unittest.mock as mock
    with mock.patch.object(O.lfmm_oracle, "init_from_random_probe",
                           lambda *a, **kw: (np.array(w0, float), np.array(h0, float))):
        return O.nmf_impute_estimation(packed, bps, n, np.arange(p), len(w0[0]), n_iter, EPS, cv_rate,
                                       chunk, seed, None, synth.uniform_stream)


def _nmf_case():
    n, p, k = 11, 9, 2
    packed, bps = _panel(n, p)
    rng = np.random.default_rng(8)
    w0 = rng.random((n, k)) + 0.1
    h0 = rng.random((k, p)) * 2 + 0.1
    return packed, bps, n, p, w0.tolist(), h0.tolist()


@pytest.mark.parametrize("chunk", [2, 4, 9])
def test_nmf_iterations_match_bruteforce_reference_order(chunk):
    packed, bps, n, p, w0, h0 = _nmf_case()
    res = _oracle_iterations(packed, bps, n, p, w0, h0, 3, 0.3, chunk, 42)
    w, h, cv, mmae, total = bf_nmf(packed, bps, n, p, w0, h0, 3, 0.3, chunk, 42)
    assert np.allclose(res.w, w, rtol=1e-12, atol=0)
    assert np.allclose(res.h_full, h, rtol=1e-12, atol=0)
    assert np.allclose(res.cv_mae_per_iter, cv, rtol=1e-12, atol=0)
    assert res.cv_mae_mean_impute == pytest.approx(mmae, rel=1e-12)
    assert res.total_mask_count == total > 0


# (Normalising W between the two updates is not among them: the H update is
# equivariant under W -> W D^-1, H -> D H, so skipping it moves nothing but
# rounding.)
@pytest.mark.parametrize("mutation", ["h_before_w", "bwd_fill_from_pre_update", "fill_with_column_mean",
                                      "w_denominator_plus_eps", "cv_draws_observed_only",
                                      "cv_from_pre_iteration_model"])
def test_nmf_pin_detects_mutation(mutation):
    packed, bps, n, p, w0, h0 = _nmf_case()
    res = _oracle_iterations(packed, bps, n, p, w0, h0, 3, 0.3, 4, 42)
    w, h, cv, mmae, total = bf_nmf(packed, bps, n, p, w0, h0, 3, 0.3, 4, 42, mutation)
    if mutation == "w_denominator_plus_eps":
        # eps = 1e-16 against denominators of order 1: the variant of the math
        # notes differs only by rounding, so it needs its own tiny-denominator case
        small = [[v * 1e-9 for v in r] for r in w0]
        res = _oracle_iterations(packed, bps, n, p, small, [[v * 1e-9 for v in r] for r in h0], 1, 0.3, 4, 42)
        w, h, *_ = bf_nmf(packed, bps, n, p, small, [[v * 1e-9 for v in r] for r in h0], 1, 0.3, 4, 42, mutation)
        w_ok, *_ = bf_nmf(packed, bps, n, p, small, [[v * 1e-9 for v in r] for r in h0], 1, 0.3, 4, 42)
        assert np.allclose(res.w, w_ok, rtol=1e-12)
    moved = max(float(np.max(np.abs(res.w - w) / np.abs(res.w))), float(np.max(np.abs(res.h_full - h) / res.h_full)),
                float(np.max(np.abs(np.array(res.cv_mae_per_iter) - np.array(cv)))) if len(cv) == len(res.cv_mae_per_iter) else 1.0,
                abs(res.total_mask_count - total))
    assert moved > 1e-6, mutation


# ---------------------------------------------------------------------------
# init_from_random_probe
# ---------------------------------------------------------------------------

def bf_init(packed, bps, n, p, k, omega, mutation=None):
    y = _raw_matrix(packed, bps, n, range(p))
    means = _col_means_observed(y, range(p))
    yf = [[means[j] if _is_nan(y[i][j]) else y[i][j] for j in range(p)] for i in range(n)]
    if mutation == "centred_decode":
        yf = [[yf[i][j] - means[j] for j in range(p)] for i in range(n)]
    ya = np.array(yf)
    z = ya.T @ omega
    # the top-k right singular vectors of Z as eigenvectors of Z^T Z (a different routine)
    ev, vec = np.linalg.eigh(z.T @ z)
    v = vec[:, np.argsort(ev)[::-1][:k]]                  # l x k
    if mutation == "transposed_vt":
        v = np.linalg.svd(z, full_matrices=False)[2][:, :k]
    wp = omega @ v
    w = wp if mutation == "no_abs" else np.abs(wp)
    # H: least squares of the mean-filled raw panel on W, clamped at 0
    hc = np.linalg.lstsq(w, ya, rcond=None)[0]
    h = hc if mutation == "no_clamp" else np.maximum(hc, 0.0)
    return w, h


@pytest.mark.parametrize("chunk", [3, 17])
def test_init_from_random_probe_matches_bruteforce(chunk):
    n, p, k = 13, 17, 2
    packed, bps = _panel(n, p, seed=5)
    omega = synth.normal_matrix(1_000_042, n, k + 10)
    w, h = O.init_from_random_probe(packed, bps, n, np.arange(p), k, chunk, omega)
    wb, hb = bf_init(packed, bps, n, p, k, omega)
    assert np.allclose(w, wb, rtol=1e-9, atol=1e-12)
    assert np.allclose(h, hb, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("mutation", ["centred_decode", "transposed_vt", "no_abs", "no_clamp"])
def test_init_pin_detects_mutation(mutation):
    n, p, k = 13, 17, 2
    packed, bps = _panel(n, p, seed=5)
    omega = synth.normal_matrix(1_000_042, n, k + 10)
    w, h = O.init_from_random_probe(packed, bps, n, np.arange(p), k, 3, omega)
    wb, hb = bf_init(packed, bps, n, p, k, omega, mutation)
    assert max(float(np.max(np.abs(w - wb))), float(np.max(np.abs(h - hb)))) > 1e-6, mutation


# ---------------------------------------------------------------------------
# GWAS-phase CV inside the association pass
# ---------------------------------------------------------------------------

def bf_gwas_cv(packed, bps, n, p, chunk, seed, cv_rate, w, w_pinv, mutation=None):
    y = _raw_matrix(packed, bps, n, range(p))
    k = len(w[0])
    nmf_err = mean_err = 0.0
    cnt = 0
    for seq, start in enumerate(range(0, p, chunk)):
        cols = list(range(start, min(p, start + chunk)))
        c = len(cols)
        m = [[y[i][cols[cc]] for cc in range(c)] for i in range(n)]
        n_obs = sum(1 for cc in range(c) for i in range(n) if not _is_nan(m[i][cc]))
        draws = synth.uniform_stream(seed + seq, n * c if mutation == "draw_every_entry" else n_obs)
        pos = 0
        positions = []
        for cc in range(c):
            for i in range(n):
                if _is_nan(m[i][cc]) and mutation != "draw_every_entry":
                    continue
                r = draws[pos]
                pos += 1
                if r < cv_rate and not _is_nan(m[i][cc]):
                    positions.append((i, cc, m[i][cc]))
        if not positions:
            continue
        masked = [row[:] for row in m]
        if mutation != "mean_before_masking":
            for i, cc, _ in positions:
                masked[i][cc] = float("nan")
        means = _col_means_observed(masked, range(c))
        if mutation == "mean_before_masking":
            for i, cc, _ in positions:
                masked[i][cc] = float("nan")
        for cc in range(c):
            for i in range(n):
                if _is_nan(masked[i][cc]):
                    masked[i][cc] = means[cc]
        hc = [[max(0.0, sum(w_pinv[a][i] * masked[i][cc] for i in range(n))) for cc in range(c)] for a in range(k)]
        for i, cc, true in positions:
            pred = sum(w[i][a] * hc[a][cc] for a in range(k))
            nmf_err += abs(true - pred)
            mean_err += abs(true - masked[i][cc])
        cnt += len(positions)
    return (nmf_err / cnt if cnt else 0.0, mean_err / cnt if cnt else 0.0, cnt)


def _gwas_case():
    n, p, k = 14, 11, 2
    pan = synth.balding_nichols(n, p, n_pops=2, miss_rate=0.15, d=1, seed=21)
    packed, bps = O.encode_bed(pan.g), O.bytes_per_snp(n)
    rng = np.random.default_rng(4)
    w = rng.random((n, k)) + 0.05
    w_pinv = O.safe_inv(w.T @ w) @ w.T
    xs = O.center_covariates(pan.x)
    u = rng.standard_normal((n, 2))
    return pan, packed, bps, n, p, w, w_pinv, xs, u


@pytest.mark.parametrize("chunk", [3, 5, 11])
def test_gwas_phase_cv_matches_bruteforce(chunk):
    pan, packed, bps, n, p, w, w_pinv, xs, u = _gwas_case()
    pre = O.precompute(xs, 1e-5)
    cfg = dict(cv_rate=0.35, seed=42, w=w, w_pinv=w_pinv, uniform=synth.uniform_stream)
    res = O.test_associations_fused(packed, bps, n, p, xs, u, pre, 2, chunk, O.EIGENSTRAT, O.Impute(), nmf_cv=cfg)
    exp = bf_gwas_cv(packed, bps, n, p, chunk, 42, 0.35, w.tolist(), w_pinv.tolist())
    assert res.cv[2] == exp[2] > 0
    assert res.cv[0] == pytest.approx(exp[0], rel=1e-12)
    assert res.cv[1] == pytest.approx(exp[1], rel=1e-12)


@pytest.mark.parametrize("mutation", ["draw_every_entry", "mean_before_masking"])
def test_gwas_cv_pin_detects_mutation(mutation):
    pan, packed, bps, n, p, w, w_pinv, xs, u = _gwas_case()
    pre = O.precompute(xs, 1e-5)
    cfg = dict(cv_rate=0.35, seed=42, w=w, w_pinv=w_pinv, uniform=synth.uniform_stream)
    res = O.test_associations_fused(packed, bps, n, p, xs, u, pre, 2, 5, O.EIGENSTRAT, O.Impute(), nmf_cv=cfg)
    mut = bf_gwas_cv(packed, bps, n, p, 5, 42, 0.35, w.tolist(), w_pinv.tolist(), mutation)
    assert abs(res.cv[0] - mut[0]) + abs(res.cv[1] - mut[1]) + abs(res.cv[2] - mut[2]) > 1e-6, mutation


# ---------------------------------------------------------------------------
# r2 partition: Type-I sums of squares of nested least-squares fits
# ---------------------------------------------------------------------------

def bf_r2(y, xs, u, mutation=None):
    n = len(y)
    one = np.ones((n, 1))
    tss = float(y @ y)
    if tss <= 1e-300:
        return np.zeros(3)

    def rss(design):
        coef = np.linalg.lstsq(design, y, rcond=None)[0]
        r = y - design @ coef
        return float(r @ r)

    rss0 = rss(one)                                      # y is centred: equals tss
    first = u if mutation == "latent_projector" else xs
    rss1 = rss(np.hstack([one, first]))
    rss2 = rss(np.hstack([one, xs, u]))
    ss_cov = rss0 - rss1
    if mutation == "ridge_projector":
        lam = 5.0
        fx = xs @ np.linalg.solve(xs.T @ xs + lam * np.eye(xs.shape[1]), xs.T @ y)
        ss_cov = float(fx @ fx)
    ss_lat = max(0.0, tss - rss2 - ss_cov)
    return np.array([ss_cov, ss_lat, rss2]) / tss


def _r2_case():
    pan = synth.balding_nichols(40, 30, n_pops=3, miss_rate=0.05, d=2, n_causal=4, effect=1.5, seed=31)
    n, p = pan.g.shape
    packed, bps = O.encode_bed(pan.g), O.bytes_per_snp(n)
    xs = O.center_covariates(pan.x)
    u = np.random.default_rng(2).standard_normal((n, 3))
    pre = O.precompute(xs, 1e-5)
    res = O.test_associations_fused(packed, bps, n, p, xs, u, pre, 3, 7, O.EIGENSTRAT, O.Impute())
    y = O.decode_chunk(packed, bps, n, np.arange(p), O.EIGENSTRAT)
    return res, y, xs, u


def test_r2_matches_sequential_anova():
    res, y, xs, u = _r2_case()
    for j in range(y.shape[1]):
        exp = bf_r2(y[:, j], xs, u)
        assert np.allclose(res.r2[j], exp, rtol=1e-9, atol=1e-12), j
    assert np.max(res.r2[:, 0]) > 1e-2        # the covariate part is not trivially zero


@pytest.mark.parametrize("mutation", ["latent_projector", "ridge_projector"])
def test_r2_pin_detects_mutation(mutation):
    res, y, xs, u = _r2_case()
    moved = max(float(np.max(np.abs(res.r2[j] - bf_r2(y[:, j], xs, u, mutation)))) for j in range(y.shape[1]))
    assert moved > 1e-4, mutation
