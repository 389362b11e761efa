"""Host-side mirror of schnelLFMM's library interface over the C ABI.

Names, argument meaning and error behaviour follow the reference crate
(src/lib.rs, src/bed.rs); every numeric step runs in libslfmm.so on the GPU.
Results come back as numpy arrays instead of the reference's on-disk
fragments (the output writer, src/testing.rs:530-983, is not part of this
path).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Union

import numpy as np

from ._lib import AssocOut, Config, NmfConfig, SlfmmError, check, lib

CENTER_ONLY, EIGENSTRAT = 0, 1          # SnpNorm, src/bed.rs:19-26
PREC_FP64, PREC_I8 = 0, 1
NMF_INT8 = 4                              # slfmm_nmf_estimate flag: both NMF passes on the int8 matrix cores
N_BINS = 100_000


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class Context:
    """One device, one HIP stream (slfmm_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(None, lib().slfmm_create(device, C.byref(h)))
        self.h = h
        from . import tuning
        tuning.apply(self)   # launch shapes of the tuning table (results do not depend on them)

    def close(self):
        if self.h:
            lib().slfmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        check(self.h, rc)


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class Lfmm2Config:
    """src/lib.rs:47-98 (+ precision)."""
    k: int = 5
    lam: float = 1e-5
    chunk_size: int = 10_000
    oversampling: int = 10
    n_power_iter: int = 2
    seed: int = 42
    n_workers: int = 0          # accepted for interface parity; the GPU needs no host workers
    progress: bool = False
    norm: int = CENTER_ONLY
    scale_cov: bool = False
    impute: str = "mean"
    precision: int = PREC_FP64

    def to_c(self) -> Config:
        c = Config()
        lib().slfmm_config_default(C.byref(c))
        c.k, c.lam, c.chunk_size = self.k, self.lam, self.chunk_size
        c.oversampling, c.n_power_iter, c.seed = self.oversampling, self.n_power_iter, self.seed
        c.norm, c.scale_cov = self.norm, int(self.scale_cov)
        c.impute = 0 if self.impute == "mean" else 1
        c.precision = self.precision
        return c


def subset_indices(spec, n_snps: int) -> np.ndarray:
    """SubsetSpec expansion (src/parallel.rs:113-122): None/'all', float rate, or indices."""
    if spec is None or (isinstance(spec, str) and spec == "all"):
        return np.arange(n_snps, dtype=np.int64)
    if isinstance(spec, float):
        if not (0.0 < spec <= 1.0):
            raise ValueError("est-rate must be in (0, 1]")
        step = int(math.ceil(1.0 / spec))
        return np.arange(0, n_snps, step, dtype=np.int64)
    return np.ascontiguousarray(np.asarray(spec, dtype=np.int64))


class BedFile:
    """A PLINK genotype panel (BedFile, src/bed.rs:60-172): device-resident
    (open, from_packed, synthetic) or host-resident and streamed to the device
    in chunks (open_host, wrap_host)."""

    def __init__(self, handle, n_samples: int, n_snps: int, ctx: Context, path: Optional[str] = None,
                 host_buffer=None):
        self.h = handle
        self.n_samples = n_samples
        self.n_snps = n_snps
        self.ctx = ctx
        self.path = path
        self._host_buffer = host_buffer   # rows borrowed by a host-resident panel

    @classmethod
    def open(cls, bed_path: str, sample_keep: Optional[Sequence[int]] = None,
             ctx: Optional[Context] = None) -> "BedFile":
        """BedFile::open (+ subset_samples): validate, mmap, upload to HBM."""
        ctx = ctx or default_context()
        keep = None if sample_keep is None else np.ascontiguousarray(np.asarray(sample_keep, np.int64))
        h = C.c_void_p()
        ctx.check(lib().slfmm_geno_load_bed(ctx.h, bed_path.encode(), _ptr(keep),
                                            0 if keep is None else len(keep), C.byref(h)))
        n, p = C.c_uint64(), C.c_uint64()
        lib().slfmm_geno_info(h, C.byref(n), C.byref(p), None)
        return cls(h, n.value, p.value, ctx, bed_path)

    @staticmethod
    def validate(bed_path: str, ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        n, p = C.c_uint64(), C.c_uint64()
        ctx.check(lib().slfmm_bed_open(ctx.h, bed_path.encode(), C.byref(n), C.byref(p)))
        return n.value, p.value

    @classmethod
    def from_packed(cls, packed: bytes, n_phys: int, n_snps: int,
                    sample_keep: Optional[Sequence[int]] = None, ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        buf = np.frombuffer(packed, dtype=np.uint8)
        keep = None if sample_keep is None else np.ascontiguousarray(np.asarray(sample_keep, np.int64))
        h = C.c_void_p()
        ctx.check(lib().slfmm_geno_upload(ctx.h, _ptr(buf), n_phys, n_snps, _ptr(keep),
                                          0 if keep is None else len(keep), C.byref(h)))
        n = len(keep) if keep is not None else n_phys
        return cls(h, n, n_snps, ctx)

    @classmethod
    def open_host(cls, bed_path: str, chunk_snps: int = 0, ctx: Optional[Context] = None) -> "BedFile":
        """Host-resident panel: the .bed stays memory-mapped and its rows are
        streamed to the device chunk by chunk (src/bed.rs:78-135 mmap,
        src/parallel.rs:156-329 chunk pipeline).  chunk_snps=0 picks about
        1 GiB of rows per chunk."""
        ctx = ctx or default_context()
        h = C.c_void_p()
        ctx.check(lib().slfmm_geno_open_bed_host(ctx.h, bed_path.encode(), chunk_snps, C.byref(h)))
        n, p = C.c_uint64(), C.c_uint64()
        lib().slfmm_geno_info(h, C.byref(n), C.byref(p), None)
        return cls(h, n.value, p.value, ctx, bed_path)

    @classmethod
    def wrap_host(cls, packed, n: int, n_snps: int, chunk_snps: int = 0,
                  ctx: Optional[Context] = None) -> "BedFile":
        """Host-resident panel over packed .bed rows in host memory (borrowed;
        the panel keeps a reference to the buffer)."""
        ctx = ctx or default_context()
        buf = np.frombuffer(packed, dtype=np.uint8) if not isinstance(packed, np.ndarray) else packed
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        if buf.size < n_snps * ((n + 3) // 4):
            raise ValueError("packed buffer shorter than n_snps rows of ceil(n/4) bytes")
        h = C.c_void_p()
        ctx.check(lib().slfmm_geno_wrap_host(ctx.h, _ptr(buf), n, n_snps, chunk_snps, C.byref(h)))
        return cls(h, n, n_snps, ctx, host_buffer=buf)

    @classmethod
    def synthetic(cls, n: int, p: int, n_pops: int = 20, fst=(0.02, 0.15), maf=(0.05, 0.5),
                  miss_rate: float = 0.02, seed: int = 3, ctx: Optional[Context] = None):
        """Balding-Nichols panel generated directly in HBM (slfmm_synth_bn)."""
        ctx = ctx or default_context()
        h = C.c_void_p()
        ctx.check(lib().slfmm_synth_bn(ctx.h, n, p, n_pops, fst[0], fst[1], maf[0], maf[1],
                                       miss_rate, seed, C.byref(h)))
        return cls(h, n, p, ctx)

    @classmethod
    def synthetic_panel(cls, n: int, p: int, n_pops: int = 20, fst=(0.02, 0.15), maf=(0.05, 0.5),
                        maf_beta: bool = False, maf_min: float = 0.0, ancestry=None, het_rate: float = -1.0,
                        miss_rate: float = 0.0, sample_miss=None, snp_miss_beta=None, miss_scale: float = 0.0,
                        seed: int = 1, ctx: Optional[Context] = None):
        """Generalised Balding-Nichols panel in HBM (slfmm_synth_bn_ex): admixture
        proportions (n x n_pops), inbred lines (het_rate >= 0), Beta(maf) ancestral
        frequencies, per-sample (sample_miss) and per-SNP (Beta snp_miss_beta)
        missingness."""
        from ._lib import SynthSpec
        ctx = ctx or default_context()
        sp = SynthSpec()
        sp.n_pops, sp.fst_lo, sp.fst_hi = n_pops, fst[0], fst[1]
        sp.maf_beta, sp.maf_a, sp.maf_b, sp.maf_min = int(bool(maf_beta)), maf[0], maf[1], maf_min
        keep = []
        if ancestry is not None:
            a = np.ascontiguousarray(ancestry, dtype=np.float32)
            if a.shape != (n, n_pops):
                raise ValueError("ancestry must be n x n_pops")
            keep.append(a)
            sp.ancestry = a.ctypes.data
        sp.het_rate, sp.miss_rate = het_rate, miss_rate
        if sample_miss is not None:
            m = np.ascontiguousarray(sample_miss, dtype=np.float32)
            if m.shape != (n,):
                raise ValueError("sample_miss must have n entries")
            keep.append(m)
            sp.sample_miss = m.ctypes.data
        sp.snp_miss_a, sp.snp_miss_b = (snp_miss_beta if snp_miss_beta is not None else (0.0, 0.0))
        sp.miss_scale, sp.seed = miss_scale, seed
        h = C.c_void_p()
        ctx.check(lib().slfmm_synth_bn_ex(ctx.h, n, p, C.byref(sp), C.byref(h)))
        return cls(h, n, p, ctx)

    def read_packed(self, j0: int = 0, rows: Optional[int] = None) -> np.ndarray:
        """Rows [j0, j0 + rows) as .bed payload bytes (rows x ceil(n/4), uint8)."""
        rows = self.n_snps - j0 if rows is None else rows
        out = np.empty((rows, (self.n_samples + 3) // 4), np.uint8)
        self.ctx.check(lib().slfmm_geno_read_packed(self.ctx.h, self.h, j0, rows, _ptr(out)))
        return out

    def subset_snp_count(self, spec) -> int:
        return len(subset_indices(spec, self.n_snps))

    def free(self):
        if self.h:
            lib().slfmm_geno_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ---- decode (src/bed.rs:211-281) ----
    def decode(self, snp_idx, norm: int = CENTER_ONLY, fill: Optional[np.ndarray] = None) -> np.ndarray:
        idx = np.ascontiguousarray(np.asarray(snp_idx, np.int64))
        out = np.empty((self.n_samples, len(idx)))
        f = None if fill is None else _f64(fill)
        self.ctx.check(lib().slfmm_decode(self.ctx.h, self.h, _ptr(idx), len(idx), norm, _ptr(f), _ptr(out)))
        return out

    def decode_raw(self, snp_idx) -> np.ndarray:
        idx = np.ascontiguousarray(np.asarray(snp_idx, np.int64))
        out = np.empty((self.n_samples, len(idx)))
        self.ctx.check(lib().slfmm_decode_raw(self.ctx.h, self.h, _ptr(idx), len(idx), _ptr(out)))
        return out

    def stats(self, snp_idx):
        idx = np.ascontiguousarray(np.asarray(snp_idx, np.int64))
        a, b, c = (np.empty(len(idx), np.int32) for _ in range(3))
        self.ctx.check(lib().slfmm_decode_stats(self.ctx.h, self.h, _ptr(idx), len(idx), _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c


def center_covariates(x, scale: bool = False) -> np.ndarray:
    """src/lib.rs:102-117."""
    x = _f64(x)
    out = np.empty_like(x)
    rc = lib().slfmm_center_covariates(_ptr(x), x.shape[0], x.shape[1], int(scale), _ptr(out))
    if rc:
        raise SlfmmError(rc, "center_covariates")
    return out


def safe_inv(a, ctx: Optional[Context] = None):
    """safe_inv (src/testing.rs:996-1024) on the GPU: returns (inverse, regularized)."""
    ctx = ctx or default_context()
    a = _f64(a)
    out = np.empty_like(a)
    reg = C.c_int32()
    ctx.check(lib().slfmm_safe_inv(ctx.h, _ptr(a), a.shape[0], _ptr(out), C.byref(reg)))
    return out, bool(reg.value)


@dataclass
class Precomputed:
    sigma: np.ndarray
    d_lambda: np.ndarray
    q_thin: np.ndarray
    ridge_inv: np.ndarray


def precompute(xs, lam: float, ctx: Optional[Context] = None) -> Precomputed:
    """Thin-form precompute (src/precompute.rs:30-88)."""
    ctx = ctx or default_context()
    xs = _f64(xs)
    n, d = xs.shape
    s, dl, q, r = np.empty(d), np.empty(d), np.empty((n, d)), np.empty((d, d))
    ctx.check(lib().slfmm_precompute(ctx.h, _ptr(xs), n, d, float(lam), _ptr(s), _ptr(dl), _ptr(q), _ptr(r)))
    return Precomputed(s, dl, q, r)


@dataclass
class TestResults:
    """src/testing.rs:76-93 plus the per-SNP arrays."""
    u_hat: np.ndarray
    gif: float
    gif_per_trait: np.ndarray
    beta: np.ndarray
    tstat: np.ndarray
    r2: np.ndarray
    pvalue: np.ndarray
    hist: Optional[np.ndarray] = None
    nmf_cv: Optional[np.ndarray] = None          # estimation CV MAE per NMF iteration
    nmf_cv_mean: Optional[float] = None          # mean-fill MAE on the final iteration's mask
    gwas_cv: Optional[tuple] = None              # (NMF MAE, mean-fill MAE, held-out count)

    __test__ = False   # not a pytest class


def estimate_factors_streaming(y_est: BedFile, subset, xs, config: Lfmm2Config, b0=None) -> np.ndarray:
    """src/rsvd.rs:26-177 on centred covariates xs; b0 = M^T Omega may be injected."""
    ctx = y_est.ctx
    xs = _f64(xs)
    idx = None if subset is None or (isinstance(subset, str) and subset == "all") else subset_indices(subset, y_est.n_snps)
    cfg = config.to_c()
    u = np.empty((y_est.n_samples, config.k))
    b = None if b0 is None else _f64(b0)
    ctx.check(lib().slfmm_estimate_factors(ctx.h, y_est.h, _ptr(idx), 0 if idx is None else len(idx),
                                           _ptr(xs), xs.shape[1], C.byref(cfg), _ptr(b), _ptr(u), 0))
    return u


def estimate_factors(y_est: BedFile, subset, x, config: Lfmm2Config, b0=None) -> np.ndarray:
    """src/lib.rs:127-136."""
    return estimate_factors_streaming(y_est, subset, center_covariates(x, config.scale_cov), config, b0)


@dataclass
class NmfParams:
    """NmfConfig (src/nmf.rs:15-50)."""
    k: int = 5
    n_iter: int = 10
    eps: float = 1e-16
    cv_rate: float = 5e-4
    chunk_size: int = 10_000
    seed: int = 42

    def to_c(self) -> NmfConfig:
        c = NmfConfig()
        lib().slfmm_nmf_config_default(C.byref(c))
        c.k, c.n_iter, c.eps, c.cv_rate = self.k, self.n_iter, self.eps, self.cv_rate
        c.chunk_size, c.seed = self.chunk_size, self.seed
        return c


@dataclass
class NmfEstResult:
    """NmfEstResult (src/nmf.rs:53-66)."""
    w: np.ndarray
    h_full: np.ndarray
    cv_mae_per_iter: np.ndarray
    cv_mae_mean_impute: float
    total_mask_count: int


def nmf_impute_estimation(bed: BedFile, subset, config: NmfParams, omega=None, int8: bool = False) -> NmfEstResult:
    """src/nmf.rs:80-249 on the GPU (FP64).  omega (n x (k + 10)) may be
    injected; otherwise it is drawn from config.seed.  int8: the forward
    pass and the H update of every iteration run on the int8 matrix cores
    (k <= 8; SLFMM_NMF_INT8)."""
    ctx = bed.ctx
    idx = None if subset is None or (isinstance(subset, str) and subset == "all") else subset_indices(subset, bed.n_snps)
    pe = bed.n_snps if idx is None else len(idx)
    cfg = config.to_c()
    w = np.empty((bed.n_samples, config.k))
    h = np.empty((config.k, pe))
    cv = np.empty(max(config.n_iter, 1))
    mean = C.c_double()
    cnt = C.c_uint64()
    om = None if omega is None else _f64(omega)
    ctx.check(lib().slfmm_nmf_estimate(ctx.h, bed.h, _ptr(idx), 0 if idx is None else len(idx), C.byref(cfg),
                                       _ptr(om), _ptr(w), _ptr(h), _ptr(cv), C.byref(mean), C.byref(cnt),
                                       NMF_INT8 if int8 else 0))
    return NmfEstResult(w, h, cv[:config.n_iter].copy(), float(mean.value), int(cnt.value))


def _assoc_arrays(p: int, d: int, with_hist: bool):
    return (np.empty((p, d)), np.empty((p, d)), np.empty((p, 3)), np.empty((p, d)),
            np.empty((d, N_BINS), np.uint64) if with_hist else None, np.empty(d), np.empty(1))


def _out_struct(arrs) -> AssocOut:
    o = AssocOut()
    o.beta, o.tstat, o.r2, o.pvalue, o.hist, o.gif_per_trait, o.gif = (_ptr(a) for a in arrs)
    return o


def test_associations_fused(y_full: BedFile, xs, u_hat, config: Lfmm2Config, with_hist: bool = False) -> TestResults:
    """src/testing.rs:252-528 on centred covariates."""
    ctx = y_full.ctx
    xs, u_hat = _f64(xs), _f64(u_hat)
    arrs = _assoc_arrays(y_full.n_snps, xs.shape[1], with_hist)
    out = _out_struct(arrs)
    cfg = config.to_c()
    ctx.check(lib().slfmm_test_associations(ctx.h, y_full.h, _ptr(xs), xs.shape[1], _ptr(u_hat),
                                            C.byref(cfg), C.byref(out), 0))
    beta, t, r2, p, hist, gpt, gif = arrs
    return TestResults(u_hat, float(gif[0]), gpt, beta, t, r2, p, hist)


def test_associations(y_full: BedFile, x, u_hat, config: Lfmm2Config, with_hist: bool = False) -> TestResults:
    """src/lib.rs:144-154."""
    return test_associations_fused(y_full, center_covariates(x, config.scale_cov), u_hat, config, with_hist)


def fit_lfmm2(y_est: BedFile, est_subset, y_full: BedFile, x, config: Lfmm2Config, b0=None,
              with_hist: bool = False) -> TestResults:
    """Full pipeline, src/lib.rs:162-263 (mean imputation)."""
    ctx = y_full.ctx
    x = _f64(x)
    idx = None if est_subset is None or (isinstance(est_subset, str) and est_subset == "all") \
        else subset_indices(est_subset, y_est.n_snps)
    d = x.shape[1]
    arrs = _assoc_arrays(y_full.n_snps, d, with_hist)
    out = _out_struct(arrs)
    u = np.empty((y_full.n_samples, config.k))
    cfg = config.to_c()
    b = None if b0 is None else _f64(b0)
    ctx.check(lib().slfmm_fit_lfmm2(ctx.h, y_est.h, _ptr(idx), 0 if idx is None else len(idx), y_full.h,
                                    _ptr(x), d, C.byref(cfg), _ptr(b), _ptr(u), C.byref(out), 0))
    beta, t, r2, p, hist, gpt, gif = arrs
    return TestResults(u, float(gif[0]), gpt, beta, t, r2, p, hist)


@dataclass
class Lfmm2Problem:
    """One problem of fit_lfmm2_batch: the arguments of fit_lfmm2."""
    y_est: BedFile
    est_subset: object
    y_full: BedFile
    x: np.ndarray
    config: Lfmm2Config
    b0: Optional[np.ndarray] = None


def fit_lfmm2_batch(problems, n_streams: Optional[int] = None, with_hist: bool = False):
    """Many independent fit_lfmm2 problems in one call (slfmm_fit_lfmm2_batch):
    the problems are spread over n_streams (1..16; default: the tuning
    table's batch_streams) streams of the context, longest first, so small
    problems overlap on the GPU.  Returns one TestResults per problem, equal
    to what fit_lfmm2 returns for the same inputs."""
    from ._lib import Problem
    from . import tuning
    if n_streams is None:
        n_streams = tuning.params()["batch_streams"]
    problems = list(problems)
    if not problems:
        return []
    ctx = problems[0].y_full.ctx
    arr = (Problem * len(problems))()
    keep, results = [], []
    for i, pr in enumerate(problems):
        x = _f64(pr.x)
        idx = None if pr.est_subset is None or (isinstance(pr.est_subset, str) and pr.est_subset == "all") \
            else subset_indices(pr.est_subset, pr.y_est.n_snps)
        d = x.shape[1]
        arrs = _assoc_arrays(pr.y_full.n_snps, d, with_hist)
        u = np.empty((pr.y_full.n_samples, pr.config.k))
        b = None if pr.b0 is None else _f64(pr.b0)
        P = arr[i]
        P.est, P.subset_idx, P.n_idx = pr.y_est.h, _ptr(idx), 0 if idx is None else len(idx)
        P.g, P.x, P.d = pr.y_full.h, _ptr(x), d
        P.cfg = pr.config.to_c()
        P.b0, P.u_hat = _ptr(b), _ptr(u)
        P.out = _out_struct(arrs)
        keep.append((x, idx, b))
        results.append((u, arrs))
    ctx.check(lib().slfmm_fit_lfmm2_batch(ctx.h, arr, len(problems), n_streams, 0))
    out = []
    for u, arrs in results:
        beta, t, r2, p, hist, gpt, gif = arrs
        out.append(TestResults(u, float(gif[0]), gpt, beta, t, r2, p, hist))
    return out


def fit_lfmm2_nmf(y_est: BedFile, est_subset, y_full: BedFile, x, config: Lfmm2Config, nmf: NmfParams,
                  b0=None, nmf_omega=None, gwas_cv: bool = True, with_hist: bool = False,
                  nmf_int8: bool = False) -> TestResults:
    """Full pipeline with NMF imputation, src/lib.rs:162-263.  With
    config.precision PREC_I8 the filled sketch/backward passes and the filled
    association functionals run on the int8 matrix cores (nmf k <= 8); the
    forward passes, moments and epilogue stay FP64.  nmf_int8: the NMF
    estimation's update passes run on the int8 matrix cores too (nmf k <= 8;
    SLFMM_NMF_INT8)."""
    ctx = y_full.ctx
    x = _f64(x)
    idx = None if est_subset is None or (isinstance(est_subset, str) and est_subset == "all") \
        else subset_indices(est_subset, y_est.n_snps)
    d = x.shape[1]
    arrs = _assoc_arrays(y_full.n_snps, d, with_hist)
    out = _out_struct(arrs)
    u = np.empty((y_full.n_samples, config.k))
    cfg = config.to_c()
    ncfg = nmf.to_c()
    b = None if b0 is None else _f64(b0)
    om = None if nmf_omega is None else _f64(nmf_omega)
    cv_iter = np.empty(nmf.n_iter + 1)
    gcv = np.zeros(3) if gwas_cv else None
    ctx.check(lib().slfmm_fit_lfmm2_nmf(ctx.h, y_est.h, _ptr(idx), 0 if idx is None else len(idx), y_full.h,
                                        _ptr(x), d, C.byref(cfg), C.byref(ncfg), _ptr(b), _ptr(om), _ptr(u),
                                        C.byref(out), _ptr(cv_iter), _ptr(gcv), NMF_INT8 if nmf_int8 else 0))
    beta, t, r2, p, hist, gpt, gif = arrs
    res = TestResults(u, float(gif[0]), gpt, beta, t, r2, p, hist)
    res.nmf_cv = cv_iter[:nmf.n_iter].copy()
    res.nmf_cv_mean = float(cv_iter[nmf.n_iter])
    if gcv is not None:
        res.gwas_cv = (float(gcv[0]), float(gcv[1]), int(gcv[2]))
    return res
