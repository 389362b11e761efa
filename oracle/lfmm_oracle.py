"""Float64 CPU restatement of schnelLFMM's streaming LFMM2 hot path.

TEST INFRASTRUCTURE ONLY: this module is the checker for the HIP library in
``schnellfmm_amd``.  It follows the reference algorithm *as written* (dense
n x n operators, per-chunk streaming order, the same guards and the same
histogram) so that it can be compared against the GPU path on the same
inputs.  Each function cites the reference file:line it restates.

Pinning (what checks this oracle, see tests/test_oracle_golden.py):
  * decode / compute_scale / encode  -- reference known-answer tests
    src/bed.rs:511-611 (round trip, missing code, norm modes) + brute force.
  * subset_indices                   -- tests/integration.rs:754-778 vectors.
  * precompute                       -- src/precompute.rs:97-157 invariants.
  * safe_inv                         -- src/testing.rs:1078-1113 vectors.
  * TsqHistogram / GIF               -- src/testing.rs:1071-1126 vectors plus
                                        brute-force median on random data.
  * calibrate                        -- closed form erfc identity and the
                                        ordering test src/testing.rs:1129-1159.
  * rsvd                             -- exact SVD of M.Y on tiny inputs (HMT
                                        bound) and QR invariants src/rsvd.rs:185-233.
  * association                      -- brute-force per-SNP lstsq on [1|X|U],
                                        closed-form ridge; r2 against the
                                        Type-I sums of squares of nested lstsq
                                        fits (tests/test_oracle_nmf_pin.py).
  * NMF (nmf_impute_estimation,      -- per-entry brute-force loops in the
    init_from_random_probe,             reference's order (src/nmf.rs:108-232,
    compute_cv_error_masked,            :279-362, :379-455; src/testing.rs:
    _gwas_cv_chunk)                     423-487), each paired with mutations
                                        that must move the result
                                        (tests/test_oracle_nmf_pin.py); dense
                                        NMF thresholds src/nmf.rs:530-585.
Parity unpinned (no reference vector exists in this environment):
  * the exact ChaCha8/rand_distr random streams (src/rsvd.rs:44-46,
    src/nmf.rs:291-293, :433-434, src/testing.rs:437-438).  The oracle takes
    Omega / uniform streams as injected inputs instead.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
from scipy.special import erfc

BED_MAGIC = bytes([0x6C, 0x1B, 0x01])          # src/bed.rs:9
DECODE_TABLE = np.array([0.0, np.nan, 1.0, 2.0])  # src/bed.rs:16
CENTER_ONLY = 0                                  # src/bed.rs:20-26
EIGENSTRAT = 1
N_BINS = 100_000                                 # src/testing.rs:125
BIN_WIDTH = 0.001                                # src/testing.rs:126


# ---------------------------------------------------------------------------
# PLINK .bed encode / decode  (src/bed.rs)
# ---------------------------------------------------------------------------

def bytes_per_snp(n_phys: int) -> int:
    """src/bed.rs:144-146."""
    return (n_phys + 3) // 4


def encode_bed(genotypes: np.ndarray) -> bytes:
    """Pack an n x p {0,1,2,other=missing} matrix; src/bed.rs:398-428 (no magic)."""
    g = np.asarray(genotypes)
    n, p = g.shape
    bps = bytes_per_snp(n)
    code = np.full(g.shape, 1, dtype=np.uint8)       # missing -> 0b01
    code[g == 0] = 0
    code[g == 1] = 2
    code[g == 2] = 3
    pad = np.zeros((bps * 4, p), dtype=np.uint8)
    pad[:n] = code
    pad = pad.reshape(bps, 4, p)
    packed = (pad[:, 0] | (pad[:, 1] << 2) | (pad[:, 2] << 4) | (pad[:, 3] << 6))
    return packed.T.astype(np.uint8).tobytes()       # SNP-major


def write_plink(prefix: str, genotypes: np.ndarray, chrom: str = "1") -> None:
    """Write .bed/.bim/.fam (src/bed.rs:398-454 formats)."""
    n, p = genotypes.shape
    with open(prefix + ".bed", "wb") as f:
        f.write(BED_MAGIC)
        f.write(encode_bed(genotypes))
    with open(prefix + ".bim", "w") as f:
        for j in range(p):
            f.write(f"{chrom}\tsnp{j}\t0\t{j + 1}\tA\tG\n")
    with open(prefix + ".fam", "w") as f:
        for i in range(n):
            f.write(f"fam{i}\tind{i}\t0\t0\t0\t-9\n")


def _codes(packed: np.ndarray, bps: int, n_phys: int, snp_idx: Sequence[int],
           sample_keep: Optional[Sequence[int]]) -> np.ndarray:
    """2-bit codes, n_out x c (SNP-major packed input, src/bed.rs:268-279)."""
    pk = np.frombuffer(packed, dtype=np.uint8) if not isinstance(packed, np.ndarray) else packed
    idx = np.asarray(snp_idx, dtype=np.int64)
    rows = pk.reshape(-1, bps)[idx] if pk.size >= bps else np.zeros((0, bps), np.uint8)
    samples = np.arange(n_phys) if sample_keep is None else np.asarray(sample_keep, dtype=np.int64)
    byte = rows[:, samples // 4]                       # c x n_out
    code = (byte >> (2 * (samples % 4)).astype(np.uint8)) & 0x03
    return code.T                                      # n_out x c


# Byte -> the four decoded values of its samples (bits 2j..2j+1 -> sample j).
_BYTE_TABLE = DECODE_TABLE[(np.arange(256)[:, None] >> (2 * np.arange(4))[None, :]) & 0x03]


def decode_raw(packed, bps: int, n_phys: int, snp_idx, sample_keep=None) -> np.ndarray:
    """{0,1,2,NaN} decode; src/bed.rs:246-281."""
    if sample_keep is None:   # every physical sample in order: one table lookup per byte
        pk = np.frombuffer(packed, dtype=np.uint8) if not isinstance(packed, np.ndarray) else packed
        idx = np.asarray(snp_idx, dtype=np.int64)
        rows = pk.reshape(-1, bps)[idx] if pk.size >= bps else np.zeros((0, bps), np.uint8)
        vals = _BYTE_TABLE[rows].reshape(len(idx), 4 * bps)[:, :n_phys]
        return np.ascontiguousarray(vals.T)
    return DECODE_TABLE[_codes(packed, bps, n_phys, snp_idx, sample_keep)]


def compute_scale(mean, norm: int):
    """src/bed.rs:382-395 (vectorised, same operation order)."""
    mean = np.asarray(mean, dtype=np.float64)
    p = mean / 2.0
    twopq = 2.0 * p * (1.0 - p)
    if norm == CENTER_ONLY:
        return np.where(twopq < 1e-20, 0.0, 1.0)
    denom = np.sqrt(twopq)
    with np.errstate(divide="ignore"):
        return np.where(denom > 1e-10, 1.0 / np.where(denom > 0, denom, 1.0), 0.0)


def _seq_colsum(a: np.ndarray) -> np.ndarray:
    """Left-to-right column sums (the reference's `sum += v` loop order)."""
    if a.shape[0] == 0:
        return np.zeros(a.shape[1])
    return np.cumsum(a, axis=0)[-1]


def decode_chunk(packed, bps: int, n_phys: int, snp_idx, norm: int,
                 sample_keep=None, fill: Optional[np.ndarray] = None) -> np.ndarray:
    """Decode + fill + centre + scale; src/bed.rs:211-235 and :300-379."""
    col = decode_raw(packed, bps, n_phys, snp_idx, sample_keep)
    if fill is not None:
        nan = np.isnan(col)
        col[nan] = np.asarray(fill)[nan]
    valid = ~np.isnan(col)
    # integer dosages sum exactly in any order; filled values need the reference's
    s = _seq_colsum(np.where(valid, col, 0.0)) if fill is not None else np.where(valid, col, 0.0).sum(axis=0)
    nv = valid.sum(axis=0)
    mean = np.where(nv > 0, s / np.maximum(nv, 1), 0.0)
    scale = compute_scale(mean, norm)
    return np.where(np.isnan(col), 0.0, (col - mean) * scale)


def fill_missing_with_column_means(y: np.ndarray) -> np.ndarray:
    """src/parallel.rs:53-71."""
    valid = ~np.isnan(y)
    s = _seq_colsum(np.where(valid, y, 0.0))
    nv = valid.sum(axis=0)
    mean = np.where(nv > 0, s / np.maximum(nv, 1), 0.0)
    return np.where(valid, y, mean)


def subset_indices(spec, n_snps: int) -> np.ndarray:
    """SubsetSpec expansion; src/parallel.rs:113-122.  spec: None|'all', float rate, or index list."""
    if spec is None or (isinstance(spec, str) and spec == "all"):
        return np.arange(n_snps, dtype=np.int64)
    if isinstance(spec, float):
        step = int(math.ceil(1.0 / spec))
        return np.arange(0, n_snps, step, dtype=np.int64)
    return np.asarray(spec, dtype=np.int64)


def subset_snp_count(spec, n_snps: int) -> int:
    """src/bed.rs:162-171."""
    if spec is None or (isinstance(spec, str) and spec == "all"):
        return n_snps
    if isinstance(spec, float):
        step = int(math.ceil(1.0 / spec))
        return (n_snps + step - 1) // step
    return len(spec)


# ---------------------------------------------------------------------------
# Streaming decode with ImputeConfig  (src/parallel.rs:156-329)
# ---------------------------------------------------------------------------

@dataclass
class Impute:
    """ImputeConfig (src/parallel.rs:22-48): kind in mean|nmf_in_ram|nmf_on_the_fly|mean_raw|nmf_raw_in_ram."""
    kind: str = "mean"
    w: Optional[np.ndarray] = None
    h_full: Optional[np.ndarray] = None
    w_pinv: Optional[np.ndarray] = None


def stream_chunks(packed, bps, n_phys, idx, chunk, norm, impute: Impute, sample_keep=None):
    """Yield (seq, start, block) exactly as one worker of parallel_stream sees them."""
    for seq, start in enumerate(range(0, len(idx), chunk)):
        cidx = idx[start:start + chunk]
        c = len(cidx)
        if impute.kind == "mean":
            blk = decode_chunk(packed, bps, n_phys, cidx, norm, sample_keep)
        elif impute.kind == "nmf_in_ram":
            fill = impute.w @ impute.h_full[:, seq * chunk: seq * chunk + c]
            blk = decode_chunk(packed, bps, n_phys, cidx, norm, sample_keep, fill)
        elif impute.kind == "nmf_on_the_fly":
            raw = fill_missing_with_column_means(decode_raw(packed, bps, n_phys, cidx, sample_keep))
            h_c = np.maximum(impute.w_pinv @ raw, 0.0)
            fill = impute.w @ h_c
            blk = decode_chunk(packed, bps, n_phys, cidx, norm, sample_keep, fill)
        elif impute.kind == "mean_raw":
            blk = fill_missing_with_column_means(decode_raw(packed, bps, n_phys, cidx, sample_keep))
        elif impute.kind == "nmf_raw_in_ram":
            blk = decode_raw(packed, bps, n_phys, cidx, sample_keep)
            fill = impute.w @ impute.h_full[:, seq * chunk: seq * chunk + c]
            nan = np.isnan(blk)
            blk[nan] = fill[nan]
        else:
            raise ValueError(impute.kind)
        yield seq, seq * chunk, blk


# ---------------------------------------------------------------------------
# Covariates, Step 0  (src/lib.rs:102-117, src/precompute.rs:30-88)
# ---------------------------------------------------------------------------

def center_covariates(x: np.ndarray, scale: bool = False) -> np.ndarray:
    """src/lib.rs:102-117."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    xs = x - x.mean(axis=0)[None, :]
    if scale:
        for j in range(xs.shape[1]):
            var = xs[:, j] @ xs[:, j] / (n - 1.0)
            sd = math.sqrt(var)
            if sd > 1e-14:
                xs[:, j] = xs[:, j] / sd
    return xs


class SafeInvError(RuntimeError):
    pass


def safe_inv(a: np.ndarray, name: str = "") -> np.ndarray:
    """LU inverse with epsilon*I fallback; src/testing.rs:996-1024."""
    a = np.asarray(a, dtype=np.float64)
    try:
        return np.linalg.inv(a)
    except np.linalg.LinAlgError:
        n = a.shape[0]
        diag_max = max(float(np.max(np.abs(np.diag(a)))) if n else 0.0, 1e-300)
        eps = 1e-8 * diag_max
        try:
            return np.linalg.inv(a + eps * np.eye(n))
        except np.linalg.LinAlgError as e:
            raise SafeInvError(f"{name} inversion failed even with epsilon={eps:.2e}") from e


@dataclass
class Precomputed:
    """src/precompute.rs:10-21."""
    q_full: np.ndarray
    sigma: np.ndarray
    d_lambda: np.ndarray
    d_lambda_inv: np.ndarray
    m: np.ndarray
    ridge_inv: np.ndarray


def precompute(x: np.ndarray, lam: float) -> Precomputed:
    """Step 0: full SVD of X, D_lambda, M = D_lambda Q^T, ridge inverse; src/precompute.rs:30-88."""
    if lam <= 0.0:
        raise ValueError(f"Ridge penalty λ must be positive, got {lam}")
    n, d = x.shape
    u, s, _ = np.linalg.svd(x, full_matrices=True)
    dl = np.ones(n)
    for j in range(min(len(s), d)):
        dl[j] = math.sqrt(lam / (lam + s[j] * s[j]))
    dli = 1.0 / dl
    m = dl[:, None] * u.T
    xtx = x.T @ x
    xtx[np.diag_indices(d)] += lam
    return Precomputed(u, s, dl, dli, m, safe_inv(xtx, "X^T X + λI (ridge)"))


# ---------------------------------------------------------------------------
# Steps 1-2: streaming randomized SVD  (src/rsvd.rs:26-177)
# ---------------------------------------------------------------------------

def estimate_factors_streaming(packed, bps, n_phys, idx, pre: Precomputed, k: int,
                               oversampling: int, n_power_iter: int, chunk: int,
                               norm: int, impute: Impute, omega: Optional[np.ndarray] = None,
                               b0: Optional[np.ndarray] = None, sample_keep=None) -> np.ndarray:
    """HMT randomized SVD of M.Y_est; src/rsvd.rs:26-177.

    Omega (n x l) or b0 = M^T Omega must be injected (the reference draws
    Omega from ChaCha8Rng(seed), src/rsvd.rs:44-46, which is parity-unpinned).
    """
    n = pre.m.shape[0]
    l = k + oversampling
    p_est = len(idx)
    if b0 is None:
        b0 = pre.m.T @ omega                                   # src/rsvd.rs:51
    z = np.zeros((p_est, l))
    for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, norm, impute, sample_keep):
        z[start:start + blk.shape[1]] = blk.T @ b0             # src/rsvd.rs:61-69
    q_z, _ = np.linalg.qr(z)                                   # src/rsvd.rs:75
    for _ in range(n_power_iter):                              # src/rsvd.rs:79
        acc = np.zeros((n, l))
        for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, norm, impute, sample_keep):
            acc += pre.m @ (blk @ q_z[start:start + blk.shape[1]])   # src/rsvd.rs:88-97
        q_a, _ = np.linalg.qr(acc)                             # src/rsvd.rs:103-105
        mt_q = pre.m.T @ q_a                                   # src/rsvd.rs:106
        z = np.zeros((p_est, l))
        for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, norm, impute, sample_keep):
            z[start:start + blk.shape[1]] = blk.T @ mt_q       # src/rsvd.rs:117-125
        q_z, _ = np.linalg.qr(z)                               # src/rsvd.rs:131
    b_svd = np.zeros((n, l))
    for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, norm, impute, sample_keep):
        b_svd += pre.m @ (blk @ q_z[start:start + blk.shape[1]])     # src/rsvd.rs:144-153
    u_small = np.linalg.svd(b_svd, full_matrices=False)[0]     # src/rsvd.rs:162
    return pre.q_full @ (pre.d_lambda_inv[:, None] * u_small[:, :k])  # src/rsvd.rs:168-172


# ---------------------------------------------------------------------------
# Steps 3-4: association, histogram, GIF, calibration  (src/testing.rs)
# ---------------------------------------------------------------------------

def t_test(coef, sigma2, cjj):
    """t only (the p-value is discarded by the caller); src/testing.rs:1036-1048, :388."""
    coef = np.asarray(coef, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        se = np.sqrt(sigma2 * cjj)
        bad = (se < 1e-300) | ~np.isfinite(se)
        t = np.where(bad, 0.0, coef / np.where(bad, 1.0, se))
    return np.where(np.isfinite(t), t, 0.0)


class TsqHistogram:
    """src/testing.rs:114-225."""

    def __init__(self, d: int):
        self.counts = np.zeros((d, N_BINS), dtype=np.uint64)
        self.total = np.zeros(d, dtype=np.uint64)

    def add_chunk(self, tstats: np.ndarray) -> None:
        """src/testing.rs:136-149."""
        tsq = tstats * tstats
        for j in range(tstats.shape[1]):
            v = tsq[:, j]
            v = v[np.isfinite(v)]
            bins = np.minimum(v / BIN_WIDTH, float(N_BINS - 1)).astype(np.int64)
            np.add.at(self.counts[j], bins, 1)
            self.total[j] += np.uint64(len(v))

    def median_for_trait(self, j: int) -> float:
        """Literal restatement of the bin walk, src/testing.rs:180-224."""
        n = int(self.total[j])
        if n == 0:
            return 0.0
        is_even = n % 2 == 0
        target = n // 2 - 1 if is_even else n // 2
        counts = self.counts[j].tolist()
        cum = 0
        first = second = None
        for b in range(N_BINS):
            cum += counts[b]
            if first is None and cum > target:
                first = b
                if not is_even:
                    return (b + 0.5) * BIN_WIDTH
            if is_even and first is not None and cum > target + 1:
                second = b
                break
            if is_even and first is not None and cum == target + 1:
                second = b
        if first is not None and second is not None:
            return ((first + 0.5) * BIN_WIDTH + (second + 0.5) * BIN_WIDTH) / 2.0
        if first is not None:
            return (first + 0.5) * BIN_WIDTH
        return 0.0

    def compute_gif(self):
        """src/testing.rs:159-177 -> (gif_per_trait clamped >= 1, avg of raw)."""
        d = self.counts.shape[0]
        per, total = [], 0.0
        for j in range(d):
            med = self.median_for_trait(j)
            raw = 1.0 if med < 1e-10 else med / 0.4549
            per.append(max(raw, 1.0))
            total += raw
        return np.array(per), total / d


def calibrate(t: np.ndarray, gif_per_trait: np.ndarray) -> np.ndarray:
    """p = 2*Phi(-|t/sqrt(GIF)|) = erfc(|z|/sqrt 2); src/testing.rs:685-713."""
    gs = np.sqrt(np.asarray(gif_per_trait, dtype=np.float64))
    z = t / gs[None, :]
    return erfc(np.abs(z) / math.sqrt(2.0))


def log10_pvalue(t: np.ndarray, gif_per_trait: np.ndarray) -> np.ndarray:
    """log10 of the calibrated p without underflow (erfcx form), for error metrics."""
    from scipy.special import erfcx
    gs = np.sqrt(np.asarray(gif_per_trait, dtype=np.float64))
    x = np.abs(t / gs[None, :]) / math.sqrt(2.0)
    return np.log10(erfcx(x)) - x * x / math.log(10.0)


@dataclass
class AssocResult:
    beta: np.ndarray
    t: np.ndarray
    r2: np.ndarray
    hist: TsqHistogram
    gif_per_trait: np.ndarray
    gif: float
    p: np.ndarray
    cv: Optional[tuple] = None


def test_associations_fused(packed, bps, n_phys, p, x, u_hat, pre: Precomputed, k: int,
                            chunk: int, norm: int, impute: Impute, sample_keep=None,
                            nmf_cv: Optional[dict] = None) -> AssocResult:
    """Dense reference-form Steps 3-4; src/testing.rs:252-528."""
    n = x.shape[0]
    d = x.shape[1]
    if n <= 1 + d + k:                                          # src/testing.rs:272-278
        raise ValueError(f"Insufficient degrees of freedom: n={n} samples but 1+d+K={1 + d + k}")
    df = float(n - 1 - d - k)
    utu_inv = safe_inv(u_hat.T @ u_hat, "U_hat^T U_hat")       # src/testing.rs:284-285
    p_u = u_hat @ utu_inv @ u_hat.T
    xtr = pre.ridge_inv @ x.T                                  # src/testing.rs:289
    i_minus_pu = np.eye(n) - p_u
    c = np.zeros((n, 1 + d + k))
    c[:, 0] = 1.0
    c[:, 1:1 + d] = x
    c[:, 1 + d:] = u_hat
    ctc_inv = safe_inv(c.T @ c, "C^T C")                       # src/testing.rs:305-306
    h = ctc_inv @ c.T
    hx = x @ safe_inv(x.T @ x, "X^T X") @ x.T                  # src/testing.rs:315-317
    cdiag = np.array([ctc_inv[1 + j, 1 + j] for j in range(d)])
    beta = np.zeros((p, d)); tt = np.zeros((p, d)); r2 = np.zeros((p, 3))
    hist = TsqHistogram(d)
    cv_acc = [0.0, 0.0, 0]
    idx = np.arange(p, dtype=np.int64)
    for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, norm, impute, sample_keep):
        cc = blk.shape[1]
        residual = i_minus_pu @ blk                            # src/testing.rs:360
        beta[start:start + cc] = (xtr @ residual).T            # src/testing.rs:361-362
        coefs = h @ blk                                        # src/testing.rs:365
        res = blk - c @ coefs                                  # src/testing.rs:366-367
        fx = hx @ blk                                          # src/testing.rs:374
        rss = np.einsum("ij,ij->j", res, res)
        sigma2 = rss / df
        tl = np.stack([t_test(coefs[1 + j], sigma2, cdiag[j]) for j in range(d)], axis=1)
        tt[start:start + cc] = tl
        tss = np.einsum("ij,ij->j", blk, blk)
        ok = tss > 1e-300
        ss_cov = np.einsum("ij,ij->j", fx, fx)
        ss_lat = np.maximum(tss - rss - ss_cov, 0.0)
        safe = np.where(ok, tss, 1.0)
        r2[start:start + cc, 0] = np.where(ok, ss_cov / safe, 0.0)
        r2[start:start + cc, 1] = np.where(ok, ss_lat / safe, 0.0)
        r2[start:start + cc, 2] = np.where(ok, rss / safe, 0.0)
        hist.add_chunk(tl)                                     # src/testing.rs:408
        if nmf_cv is not None:                                 # src/testing.rs:423-487
            _gwas_cv_chunk(packed, bps, n_phys, idx[start:start + cc], seq, nmf_cv, cv_acc, sample_keep)
    gif_pt, avg = hist.compute_gif()                           # src/testing.rs:505
    cv = None
    if nmf_cv is not None:
        cnt = cv_acc[2]
        cv = (cv_acc[0] / cnt if cnt else 0.0, cv_acc[1] / cnt if cnt else 0.0, cnt)
    return AssocResult(beta, tt, r2, hist, gif_pt, avg, calibrate(tt, gif_pt), cv)


def _gwas_cv_chunk(packed, bps, n_phys, cidx, seq, cfg, acc, sample_keep):
    """GWAS-phase CV on one block; src/testing.rs:423-487 (uniform stream injected)."""
    raw = decode_raw(packed, bps, n_phys, cidx, sample_keep)
    n, c = raw.shape
    obs = ~np.isnan(raw)
    # One draw per OBSERVED entry, consumed in column-then-row order
    # (src/testing.rs:441-451).
    obs_cm = obs.T.reshape(-1)
    stream = cfg["uniform"](cfg["seed"] + seq, int(obs_cm.sum()))
    draws_cm = np.full(n * c, 2.0)
    draws_cm[obs_cm] = stream
    draws = draws_cm.reshape(c, n).T
    mask = obs & (draws < cfg["cv_rate"])
    if not mask.any():
        return
    masked = raw.copy()
    masked[mask] = np.nan
    masked = fill_missing_with_column_means(masked)
    hc = np.maximum(cfg["w_pinv"] @ masked, 0.0)
    pred = cfg["w"] @ hc
    acc[0] += float(np.abs(raw[mask] - pred[mask]).sum())
    acc[1] += float(np.abs(raw[mask] - masked[mask]).sum())
    acc[2] += int(mask.sum())


# ---------------------------------------------------------------------------
# NMF imputation  (src/nmf.rs:80-455)
# ---------------------------------------------------------------------------

def normalize_w(w: np.ndarray, h: np.ndarray, eps: float) -> None:
    """src/nmf.rs:259-271 (in place)."""
    k = w.shape[1]
    for j in range(k):
        col_norm = float(_seq_colsum(w[:, j:j + 1])[0])
        if col_norm > eps:
            w[:, j] *= 1.0 / col_norm
            if h.shape[0] == k:
                h[j, :] *= col_norm


def init_from_random_probe(packed, bps, n_phys, idx, k, chunk, omega, sample_keep=None):
    """src/nmf.rs:279-362 with Omega' injected (n x (k+10))."""
    p_est = len(idx)
    z = np.zeros((p_est, omega.shape[1]))
    for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, 0, Impute("mean_raw"), sample_keep):
        z[start:start + blk.shape[1]] = blk.T @ omega
    vt = np.linalg.svd(z, full_matrices=False)[2]
    w = np.abs(omega @ vt[:k, :].T)
    w_pinv = safe_inv(w.T @ w, "W_init^T W_init") @ w.T
    h = np.zeros((k, p_est))
    for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, 0, Impute("mean_raw"), sample_keep):
        h[:, start:start + blk.shape[1]] = np.maximum(w_pinv @ blk, 0.0)
    return w, h


def compute_cv_error_masked(packed, bps, n_phys, idx, w, h, mask_seed, chunk, cv_rate,
                            uniform: Callable, sample_keep=None):
    """src/nmf.rs:379-455; the uniform stream (seed, count) -> draws is injected."""
    nmf_err = mean_err = 0.0
    cnt = 0
    for seq, start in enumerate(range(0, len(idx), chunk)):
        cidx = idx[start:start + chunk]
        raw = decode_raw(packed, bps, n_phys, cidx, sample_keep)
        n, c = raw.shape
        valid = ~np.isnan(raw)
        s = _seq_colsum(np.where(valid, raw, 0.0))
        nv = valid.sum(axis=0)
        cm = np.where(nv > 0, s / np.maximum(nv, 1), 0.0)
        recon = w @ h[:, start:start + c]
        r = uniform(mask_seed + seq, n * c).reshape(c, n).T     # col, then row
        sel = (r < cv_rate) & valid
        nmf_err += float(np.abs(raw[sel] - recon[sel]).sum())
        mean_err += float(np.abs(raw[sel] - np.broadcast_to(cm, raw.shape)[sel]).sum())
        cnt += int(sel.sum())
    return (nmf_err / cnt if cnt else 0.0, mean_err / cnt if cnt else 0.0, cnt)


@dataclass
class NmfResult:
    w: np.ndarray
    h_full: np.ndarray
    cv_mae_per_iter: list = field(default_factory=list)
    cv_mae_mean_impute: float = 0.0
    total_mask_count: int = 0


def nmf_impute_estimation(packed, bps, n_phys, idx, k, n_iter, eps, cv_rate, chunk, seed,
                          omega, uniform: Callable, sample_keep=None) -> NmfResult:
    """Streaming masked multiplicative-update NMF; src/nmf.rs:80-249."""
    w, h = init_from_random_probe(packed, bps, n_phys, idx, k, chunk, omega, sample_keep)
    w = np.maximum(w, eps); h = np.maximum(h, eps)
    normalize_w(w, h, eps)
    res = NmfResult(w, h)
    n = w.shape[0]
    for it in range(n_iter):
        mask_seed = seed + it
        fwd = Impute("nmf_raw_in_ram", w=w.copy(), h_full=h.copy())
        num_w = np.zeros((n, k)); hh = np.zeros((k, k))
        for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, 0, fwd, sample_keep):
            hc = h[:, start:start + blk.shape[1]]
            num_w += blk @ hc.T                                 # src/nmf.rs:141
            hh += hc @ hc.T                                     # src/nmf.rs:145
        denom = w @ hh                                          # src/nmf.rs:157
        upd = denom > eps
        w[upd] *= num_w[upd] / denom[upd]                       # src/nmf.rs:160-167
        w = np.maximum(w, eps)
        normalize_w(w, h, eps)
        wtw = w.T @ w
        bwd = Impute("nmf_raw_in_ram", w=w.copy(), h_full=h.copy())
        for seq, start, blk in stream_chunks(packed, bps, n_phys, idx, chunk, 0, bwd, sample_keep):
            cc = blk.shape[1]
            num_h = w.T @ blk                                   # src/nmf.rs:198
            hc = h[:, start:start + cc]
            dh = wtw @ hc                                       # src/nmf.rs:205
            u = dh > eps
            hc[u] *= num_h[u] / dh[u]                           # src/nmf.rs:206-213
            h[:, start:start + cc] = np.maximum(hc, eps)
        normalize_w(w, h, eps)                                  # src/nmf.rs:224
        mae, mmae, cnt = compute_cv_error_masked(packed, bps, n_phys, idx, w, h, mask_seed,
                                                 chunk, cv_rate, uniform, sample_keep)
        res.cv_mae_per_iter.append(mae)
        res.cv_mae_mean_impute = mmae
        res.total_mask_count += cnt
    res.w, res.h_full = w, h
    return res


def nmf_dense(y: np.ndarray, k: int, n_iter: int, eps: float, w0, h0):
    """Test-only dense NMF of the reference's unit tests; src/nmf.rs:464-528 (init injected)."""
    w, h = w0.copy(), h0.copy()
    mask = ~np.isnan(y)
    for _ in range(n_iter):
        yf = np.where(mask, y, w @ h)
        wty = w.T @ yf; wtwh = w.T @ w @ h
        u = wtwh > eps
        h[u] *= wty[u] / wtwh[u]
        h = np.maximum(h, eps)
        yf2 = np.where(mask, y, w @ h)
        yht = yf2 @ h.T; whht = w @ (h @ h.T)
        u = whht > eps
        w[u] *= yht[u] / whht[u]
        w = np.maximum(w, eps)
    return w, h


# ---------------------------------------------------------------------------
# Orchestration  (src/lib.rs:162-263)
# ---------------------------------------------------------------------------

@dataclass
class Config:
    """Lfmm2Config defaults, src/lib.rs:81-98."""
    k: int = 5
    lam: float = 1e-5
    chunk_size: int = 10_000
    oversampling: int = 10
    n_power_iter: int = 2
    seed: int = 42
    norm: int = CENTER_ONLY
    scale_cov: bool = False


def fit_lfmm2(packed, bps, n_phys, p, x, cfg: Config, est_idx, omega=None, b0=None,
              sample_keep=None, nmf: Optional[dict] = None):
    """Mean or NMF imputation end to end; src/lib.rs:162-263 (random streams injected)."""
    xs = center_covariates(x, cfg.scale_cov)
    pre = precompute(xs, cfg.lam)
    imp_est, imp_test, cv, nmf_res = Impute("mean"), Impute("mean"), None, None
    if nmf is not None:
        nmf_res = nmf_impute_estimation(packed, bps, n_phys, est_idx, nmf["k"], nmf["n_iter"],
                                        nmf.get("eps", 1e-16), nmf["cv_rate"], cfg.chunk_size,
                                        nmf["seed"], nmf["omega"], nmf["uniform"], sample_keep)
        w = nmf_res.w
        w_pinv = safe_inv(w.T @ w, "W^T W for NMF pinv") @ w.T          # src/lib.rs:201-203
        imp_est = Impute("nmf_in_ram", w=w, h_full=nmf_res.h_full)
        imp_test = Impute("nmf_on_the_fly", w=w, w_pinv=w_pinv)
        cv = dict(cv_rate=nmf["cv_rate"], seed=cfg.seed, w=w, w_pinv=w_pinv, uniform=nmf["uniform"])
    u_hat = estimate_factors_streaming(packed, bps, n_phys, est_idx, pre, cfg.k, cfg.oversampling,
                                       cfg.n_power_iter, cfg.chunk_size, cfg.norm, imp_est,
                                       omega=omega, b0=b0, sample_keep=sample_keep)
    res = test_associations_fused(packed, bps, n_phys, p, xs, u_hat, pre, cfg.k, cfg.chunk_size,
                                  cfg.norm, imp_test, sample_keep, nmf_cv=cv)
    return u_hat, res, pre, nmf_res


def principal_angles(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Principal angles (radians) between col(a) and col(b)."""
    qa, _ = np.linalg.qr(a)
    qb, _ = np.linalg.qr(b)
    # sine form: accurate for the small angles parity tests care about
    r = qb - qa @ (qa.T @ qb)
    s = np.clip(np.linalg.svd(r, compute_uv=False), 0.0, 1.0)
    return np.sort(np.arcsin(s))
