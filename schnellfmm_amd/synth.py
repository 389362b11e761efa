"""Seeded synthetic genotype panels (host side, numpy) for tests and the bench.

Balding–Nichols structure: per SNP an ancestral frequency p_anc, per
population p_k ~ Beta(p_anc (1-F)/F, (1-p_anc)(1-F)/F), genotype
g ~ Bin(2, p_{pop(i)}) (or the admixed p_i = sum_k q_ik p_k), then missing
calls.  This is the recipe of the reference's simulator
(examples/sim/simulate_gwas.py:357-384; causal logit shift as in
src/simulate.rs:112-131).  Large panels are generated on the GPU by the
library (slfmm_synth_bn); this module covers test-sized panels.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

MISSING = 255


@dataclass
class Panel:
    g: np.ndarray            # n x p uint8, {0,1,2} or 255 missing
    x: np.ndarray            # n x d covariates (uncentred)
    pops: np.ndarray         # n population labels (argmax ancestry when admixed)
    causal: np.ndarray       # indices of causal SNPs


def _pop_freqs(rng, p, n_pops, fst, maf_lo, maf_hi, maf_beta=None):
    if maf_beta is not None:
        a, b = maf_beta
        p_anc = rng.beta(a, b, size=p)
        p_anc = np.clip(p_anc, 0.01, 0.99)
    else:
        p_anc = rng.uniform(maf_lo, maf_hi, size=p)
    fst = np.broadcast_to(np.asarray(fst, dtype=np.float64), (n_pops,))
    out = np.empty((n_pops, p))
    for k in range(n_pops):
        f = fst[k]
        a = p_anc * (1.0 - f) / f
        b = (1.0 - p_anc) * (1.0 - f) / f
        out[k] = rng.beta(a, b)
    return np.clip(out, 1e-6, 1 - 1e-6)


def balding_nichols(n: int, p: int, n_pops: int = 3, fst=0.1, maf=(0.05, 0.5),
                    miss_rate: float = 0.01, d: int = 1, n_causal: int = 0,
                    effect: float = 0.3, seed: int = 1, admix_alpha: Optional[float] = None,
                    inbred_het: Optional[float] = None, maf_beta=None,
                    sample_miss_beta=None, cov_rho: float = 0.0) -> Panel:
    rng = np.random.default_rng(seed)
    freqs = _pop_freqs(rng, p, n_pops, fst, maf[0], maf[1], maf_beta)
    if admix_alpha is not None:
        q = rng.dirichlet(np.full(n_pops, admix_alpha), size=n)
        pi = q @ freqs                                    # n x p
        pops = np.argmax(q, axis=1)
    else:
        pops = np.repeat(np.arange(n_pops), int(np.ceil(n / n_pops)))[:n]
        pi = freqs[pops]
    cov = np.full((d, d), cov_rho) + (1.0 - cov_rho) * np.eye(d)
    x = rng.multivariate_normal(np.zeros(d), cov, size=n)
    causal = rng.choice(p, size=min(n_causal, p), replace=False) if n_causal else np.zeros(0, np.int64)
    if n_causal:
        lg = np.log(pi[:, causal] / (1 - pi[:, causal])) + effect * x[:, [0]]
        pi = pi.copy()
        pi[:, causal] = np.clip(1.0 / (1.0 + np.exp(-lg)), 0.01, 0.99)
    if inbred_het is not None:
        # mostly homozygous lines: draw a haplotype allele, then a het with prob inbred_het
        hom = (rng.random((n, p)) < pi).astype(np.uint8) * 2
        het = rng.random((n, p)) < inbred_het
        g = np.where(het, 1, hom).astype(np.uint8)
    else:
        g = ((rng.random((n, p)) < pi).astype(np.uint8) + (rng.random((n, p)) < pi).astype(np.uint8))
    if sample_miss_beta is not None:
        rate = rng.beta(sample_miss_beta[0], sample_miss_beta[1], size=n)[:, None]
        miss = rng.random((n, p)) < rate
    else:
        miss = rng.random((n, p)) < miss_rate
    g = g.astype(np.uint8)
    g[miss] = MISSING
    return Panel(g, x, pops, causal)


def uniform_stream(seed: int, count: int) -> np.ndarray:
    """Counter-based U[0,1) stream used by the library's CV masks (SplitMix64)."""
    i = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal_matrix(seed: int, n: int, l: int) -> np.ndarray:
    """Row-major n x l N(0,1) matrix from the same stream (Box–Muller pairs)."""
    u = uniform_stream(seed, 2 * n * l).reshape(-1, 2)
    u1 = 1.0 - u[:, 0]                                   # (0,1]
    r = np.sqrt(-2.0 * np.log(u1))
    return (r * np.cos(2.0 * np.pi * u[:, 1])).reshape(n, l)


# ---------------------------------------------------------------------------
# Workload shapes shared by the tests and the benchmark
# ---------------------------------------------------------------------------

def batch_problems(seed: int = 4, count: int = 64):
    """The seeded batch of uneven LFMM2 problems: per problem (in draw order)
    n ~ logU[200, 5000], p ~ logU[2e4, 2e6], K ~ U{3..20}, d ~ U{1..8},
    2-10 populations, 0-5 % missing, then the n x d covariates.  Yields
    (i, n, p, K, d, n_pops, miss, x); panel i uses generator seed 1000 seed + i."""
    import math
    rng = np.random.default_rng(seed)
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(200), math.log(5000)))))
        p = int(round(math.exp(rng.uniform(math.log(2e4), math.log(2e6)))))
        K, d = int(rng.integers(3, 21)), int(rng.integers(1, 9))
        pops, miss = int(rng.integers(2, 11)), float(rng.uniform(0.0, 0.05))
        x = np.ascontiguousarray(rng.standard_normal((n, d)))
        yield i, n, p, K, d, pops, miss, x


def inbred_spec(n: int, seed: int = 2):
    """Inbred-panel options (8 admixed groups, Dirichlet(0.3) ancestry, 5 % het
    calls, MAF ~ Beta(0.4, 1.6) truncated at 0.01, per-sample missingness
    ~ Beta(1, 19)) for BedFile.synthetic_panel, plus the ancestry matrix."""
    rng = np.random.default_rng(seed)
    anc = rng.dirichlet(np.full(8, 0.3), size=n)
    smiss = rng.beta(1.0, 19.0, size=n)
    return dict(n_pops=8, fst=(0.05, 0.15), maf=(0.4, 1.6), maf_beta=True, maf_min=0.01, ancestry=anc,
                het_rate=0.05, sample_miss=smiss, seed=seed), anc


def nmf_spec(n: int, seed: int = 5):
    """NMF-workload options (8 components, Dirichlet(0.2) ancestry, missing
    P_ij = 1 - (1 - a_i)(1 - b_j), a_i ~ Beta(2, 18), b_j ~ Beta(1, 30),
    rescaled to 10 % overall) for BedFile.synthetic_panel."""
    rng = np.random.default_rng(seed)
    anc = rng.dirichlet(np.full(8, 0.2), size=n)
    a = rng.beta(2.0, 18.0, size=n)
    raw = 1.0 - (1.0 - a.mean()) * (1.0 - 1.0 / 31.0)      # E[P_ij] before rescaling
    return dict(n_pops=8, fst=(0.02, 0.15), maf=(0.05, 0.5), ancestry=anc, sample_miss=a,
                snp_miss_beta=(1.0, 30.0), miss_scale=0.10 / raw, seed=seed), anc
