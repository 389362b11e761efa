"""CPU emulation of the operand precisions a genotype-stream GEMM could use
on gfx950, measured against FP64 on the same panel (test infrastructure: a
study of numerics, not part of the library).

Every genotype-stream product of the fit is (exact genotype codes) x (a dense
FP64 operand): the association functionals f_j = Q^T y_j
(src/testing.rs:355-419) and the randomized-SVD passes Z = Y^T B, W = Y Q_z
(src/rsvd.rs:56-157).  The genotype side is exact in every format below
(codes 0/1/2 and the missing indicator); what differs is how the dense
operand is split into MFMA terms and how the sums accumulate:

  int8 t-term   q ~ sc (t0 + t1/254 + t2/254^2), exact int32 sums, FP64
                recombination (the library's path: 2 terms for the latent
                columns and the sketch, 3 for the covariate columns, the
                per-SNP fill code on the missing side)
  bf16 hi+lo    q ~ bf16(q) + bf16(q - bf16(q)), FP32 accumulation (one
                FP32 rounding per 32-sample MFMA step)
  fp8 e4m3 xN   N e4m3 terms of the per-column-scaled residual, FP32
                accumulation (one rounding per 128-sample MFMA step)

Centring is applied as f = sum q g' - mu sum_obs q in FP64 from the
accumulated sums (an FP32 or BF16 B operand cannot hold g - mu exactly).

Dense MFMA rates on MI355X (per CU per clock, same for both accumulators):
int8 and fp8 16384 MAC per 16x16x64 / 16x16x128 instruction step of 16 / 32
cycles (equal MAC rate), bf16 half of that.  The cost column below is the
number of int8-rate MFMA passes per genotype fragment.

Usage:  python -m oracle.precision_study [--n 10000] [--p 3000] [--p-est 5000]
Prints one table row per format (max and 99.9th percentile of |dt| and
|dlog10 p| of the covariate t statistics, and the RSVD subspace angles).
"""
from __future__ import annotations

import argparse
import math

import numpy as np
from scipy.special import erfc

# ---------------------------------------------------------------------------
# number formats
# ---------------------------------------------------------------------------


def round_bf16(x):
    """Round to bfloat16 (8 significant bits, nearest even), FP32 range."""
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    return np.ldexp(np.round(m * 256.0) / 256.0, e)


def round_e4m3(x):
    """Round to FP8 e4m3 (OCP): 4 significant bits for |x| >= 2^-6,
    subnormal quantum 2^-9, saturating at 448."""
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)                       # x = m 2^e, |m| in [0.5, 1)
    normal = np.ldexp(np.round(m * 16.0) / 16.0, e)
    sub = np.round(x / 2.0 ** -9) * 2.0 ** -9
    y = np.where(np.abs(x) >= 2.0 ** -6, normal, sub)
    return np.clip(y, -448.0, 448.0)


def int8_terms(q, terms):
    """q ~ sc (t0 + t1/254 + ...) per column, t_i integers in [-127, 127]."""
    sc = np.max(np.abs(q), axis=0)
    sc = np.where(sc > 0, sc / 127.0, 1.0)
    r = q / sc
    out = np.zeros_like(q)
    w = 1.0
    for _ in range(terms):
        t = np.clip(np.round(r), -127, 127)
        out += w * t
        r = (r - t) * 254.0
        w /= 254.0
    return out * sc


def float_terms(q, terms, rnd, top):
    """terms of a floating-point format: each term is the rounded residual
    of the previous ones, scaled per column so the first fills the range."""
    sc = np.max(np.abs(q), axis=0)
    sc = np.where(sc > 0, sc / top, 1.0)
    r = q / sc
    parts = []
    for _ in range(terms):
        t = rnd(r)
        parts.append(t * sc)
        r = r - t
    return parts


# ---------------------------------------------------------------------------
# products: genotype codes (exact) x dense operand
# ---------------------------------------------------------------------------


def fp32_blocked(a_blocks):
    """sum over axis 0 of per-step partial sums with FP32 accumulation: each
    step's partial (one MFMA's dot products, exact) is rounded to FP32 and
    added to an FP32 accumulator."""
    acc = np.zeros(a_blocks.shape[1:], dtype=np.float32)
    for b in range(a_blocks.shape[0]):
        acc = (acc + a_blocks[b].astype(np.float32)).astype(np.float32)
    return acc.astype(np.float64)


def product(G, Bparts, step):
    """G (r x n, small integers) times sum of Bparts (n x c each): exact when
    step is None, else FP32 accumulation per `step` products along n."""
    if step is None:
        return G @ sum(Bparts)
    n = G.shape[1]
    nb = math.ceil(n / step)
    pad = nb * step - n
    Gp = np.pad(G, ((0, 0), (0, pad))).reshape(G.shape[0], nb, step)
    out = 0.0
    for B in Bparts:
        Bp = np.pad(B, ((0, pad), (0, 0))).reshape(nb, step, B.shape[1])
        blocks = np.einsum("rbs,bsc->brc", Gp, Bp, optimize=True)
        out = out + fp32_blocked(blocks)
    return out


FORMATS = {
    # name: (operand terms for latent / covariate / sketch columns, accumulation step, int8-rate passes per fragment)
    "fp64 (reference)": None,
    "int8 2/3 terms (library)": ("int8", 2, 3, 2, None),
    "bf16 hi+lo, fp32 acc": ("bf16", 2, 2, 2, 32),
    "bf16 1 term, fp32 acc": ("bf16", 1, 1, 1, 32),
    "fp8 e4m3 x2, fp32 acc": ("e4m3", 2, 2, 2, 128),
    "fp8 e4m3 x3, fp32 acc": ("e4m3", 3, 3, 3, 128),
}


def operand_parts(q, kind, terms):
    if kind == "int8":
        return [int8_terms(q, terms)]
    if kind == "bf16":
        return float_terms(q, terms, round_bf16, 1.0)
    return float_terms(q, terms, round_e4m3, 448.0)


# ---------------------------------------------------------------------------
# workload
# ---------------------------------------------------------------------------


def panel(n, p, seed, n_pops=20, miss=0.02):
    rng = np.random.default_rng(seed)
    p_anc = rng.uniform(0.05, 0.5, size=p)
    fst = rng.uniform(0.02, 0.15, size=n_pops)
    freqs = np.stack([np.clip(rng.beta(p_anc * (1 - f) / f, (1 - p_anc) * (1 - f) / f), 1e-6, 1 - 1e-6)
                      for f in fst])
    pops = np.repeat(np.arange(n_pops), math.ceil(n / n_pops))[:n]
    pi = freqs[pops]
    g = (rng.random((n, p)) < pi).astype(np.int8) + (rng.random((n, p)) < pi).astype(np.int8)
    obs = rng.random((n, p)) >= miss
    return np.where(obs, g, 0).astype(np.float64), obs.astype(np.float64), pops


def standardise(g, obs):
    nobs = obs.sum(0)
    mu = g.sum(0) / np.maximum(nobs, 1)
    pp = mu / 2.0
    sc = np.where(2 * pp * (1 - pp) > 1e-20, 1.0 / np.sqrt(np.maximum(2 * pp * (1 - pp), 1e-300)), 0.0)
    return mu, sc


def assoc_t(f, tss, T, m, d, n):
    """covariate t statistics from the functionals of an orthonormal basis
    Q = C T^-1 of C = [U | X] (last d columns X)."""
    Ti = np.linalg.inv(T)
    b = f @ Ti.T
    cov = Ti @ Ti.T
    df = n - m - 1
    rss = np.maximum(tss - np.sum(f * f, axis=1), 0.0)
    se = np.sqrt(np.outer(rss / df, np.diag(cov)[m - d:]))
    return b[:, m - d:] / se


def log10p(t):
    return np.log10(np.maximum(erfc(np.abs(t) / math.sqrt(2.0)), 1e-300))


def assoc_functionals(gT, oT, mu, sc, Q, K, fmt):
    """f_j = sc_j (q^T g'_j - mu_j q^T obs_j) for every SNP row of gT (p x n)."""
    if fmt is None:
        return sc[:, None] * (gT @ Q - mu[:, None] * (oT @ Q))
    kind, tl, tx, _, step = fmt
    if kind == "int8":
        # library: exact sums; the g' side carries the per-SNP fill code c
        # (missing -> c/S), the miss side 2 terms for every column
        S = 4
        qf = np.concatenate([int8_terms(Q[:, :K], tl), int8_terms(Q[:, K:], tx)], axis=1)
        qc = int8_terms(Q, 2)
        c = np.clip(np.round(S * (gT[:, :256].sum(1) / np.maximum(oT[:, :256].sum(1), 1))), 0, 2 * S)
        miss = 1.0 - oT
        fobs = gT @ qf - mu[:, None] * (oT @ qf)
        return sc[:, None] * (fobs + (mu - c / S)[:, None] * (miss @ (qc - qf)))
    parts_l = operand_parts(Q[:, :K], kind, tl)
    parts_x = operand_parts(Q[:, K:], kind, tx)
    parts = [np.concatenate([a, b], axis=1) for a, b in zip(parts_l, parts_x)]
    return sc[:, None] * (product(gT, parts, step) - mu[:, None] * product(oT, parts, step))


def rsvd_u(gT, oT, mu, sc, K, l, q, seed, fmt):
    """randomized SVD of the standardised panel (rows = SNPs) in sample space;
    every genotype product goes through the format."""
    n = gT.shape[1]
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, l))

    def project(B):   # Z = Y^T B: per SNP, sums over samples
        if fmt is None:
            return sc[:, None] * (gT @ B - mu[:, None] * (oT @ B))
        kind, _, _, ts, step = fmt
        parts = operand_parts(B, kind, ts)
        if kind == "int8":
            return sc[:, None] * (gT @ parts[0] - mu[:, None] * (oT @ parts[0]))
        return sc[:, None] * (product(gT, parts, step) - mu[:, None] * product(oT, parts, step))

    def forward(Qz):   # W = Y Qz: per sample, sums over SNPs
        Qt = sc[:, None] * Qz
        if fmt is None:
            return gT.T @ Qt - oT.T @ (mu[:, None] * Qt)
        kind, _, _, ts, step = fmt
        p1 = operand_parts(Qt, kind, ts)
        p2 = operand_parts(mu[:, None] * Qt, kind, ts)
        if kind == "int8":
            return gT.T @ p1[0] - oT.T @ p2[0]
        return product(gT.T, p1, step) - product(oT.T, p2, step)

    Z = project(B)
    Qz, _ = np.linalg.qr(Z)
    for _ in range(q):
        W = forward(Qz)
        Qw, _ = np.linalg.qr(W)
        Z = project(Qw)
        Qz, _ = np.linalg.qr(Z)
    W = forward(Qz)
    Uw, s, _ = np.linalg.svd(W, full_matrices=False)
    return Uw[:, :K]


def angles(a, b):
    s = np.linalg.svd(a.T @ b, compute_uv=False)
    return np.arccos(np.clip(s, -1.0, 1.0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--p", type=int, default=3000)
    ap.add_argument("--p-est", type=int, default=5000)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--d", type=int, default=4)
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args()
    n, K, d = args.n, args.k, args.d
    rng = np.random.default_rng(args.seed)
    # estimation panel -> U (FP64), then the association panel with signal
    ge, oe, pops = panel(n, args.p_est, args.seed)
    mue, sce = standardise(ge, oe)
    Ye = (sce * (ge - mue * oe))
    Ue, _, _ = np.linalg.svd(Ye, full_matrices=False)
    U = Ue[:, :K]
    cov = np.full((d, d), 0.3) + 0.7 * np.eye(d)
    X = rng.multivariate_normal(np.zeros(d), cov, size=n)
    X -= X.mean(0)
    g, o, _ = panel(n, args.p, args.seed + 1)
    # causal signal on the first covariate for a tenth of the SNPs (dosage shifted by x)
    causal = rng.choice(args.p, size=args.p // 10, replace=False)
    shift = (rng.random((n, causal.size)) < 1.0 / (1.0 + np.exp(-0.6 * X[:, [0]]))).astype(np.float64)
    g[:, causal] = np.where(o[:, causal] > 0, np.clip(g[:, causal] + shift * (g[:, causal] < 2), 0, 2), 0)
    mu, sc = standardise(g, o)
    C = np.concatenate([U - U.mean(0), X], axis=1)
    Q, T = np.linalg.qr(C)
    m = K + d
    gT, oT = g.T.copy(), o.T.copy()
    tss = np.sum((sc * (g - mu * o)) ** 2, axis=0)
    geT, oeT = ge.T.copy(), oe.T.copy()
    t_ref = lp_ref = u_ref = None
    print(f"n={n} p={args.p} (association) p_est={args.p_est} (RSVD, K={K}, l={K + 10}, q=3) d={d}")
    print("| format | int8-rate passes / fragment | max dt | 99.9% dt | max dlog10p | 99.9% dlog10p "
          "| RSVD max angle (rad) | RSVD mean angle |")
    print("|---|---|---|---|---|---|---|---|")
    for name, fmt in FORMATS.items():
        f = assoc_functionals(gT, oT, mu, sc, Q, K, fmt)
        t = assoc_t(f, tss, T, m, d, n)
        lp = log10p(t)
        u = rsvd_u(geT, oeT, mue, sce, K, K + 10, 3, args.seed, fmt)
        if fmt is None:
            t_ref, lp_ref, u_ref = t, lp, u
            print(f"| {name} | - | 0 | 0 | 0 | 0 | 0 | 0 |")
            continue
        kind, tl, tx, ts, _ = fmt
        # MFMA passes per genotype fragment at the int8 rate: the terms of the
        # g' operand plus those of the missing/observed operand (needed for the
        # centring); bf16 runs at half the rate
        if kind == "int8":
            cost = (tl * K + tx * d) / m + 2
        elif kind == "bf16":
            cost = 2 * (tl + tl)
        else:
            cost = tl + tl
        dt = np.abs(t - t_ref).ravel()
        dl = np.abs(lp - lp_ref).ravel()
        a = angles(u, u_ref)
        print(f"| {name} | {cost:.1f} | {dt.max():.2e} | {np.quantile(dt, 0.999):.2e} | {dl.max():.2e} | "
              f"{np.quantile(dl, 0.999):.2e} | {a.max():.2e} | {a.mean():.2e} |", flush=True)


if __name__ == "__main__":
    main()
