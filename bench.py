"""Benchmark: end-to-end LFMM2 SNPs/s on one MI355X.

Workload (n=10k, K=20): n=10,000 samples, p=10,000,000 SNPs (packed .bed
payload ~25 GB generated directly in HBM, Balding-Nichols with 20
populations, Fst ~ U(0.02, 0.15), 2% missing), d=4 covariates with
pairwise correlation 0.3, RSVD on a 1% est-rate subset (100,000 SNPs),
K=20, oversampling 10, 3 power iterations, eigenstrat normalisation.

One step = fit_lfmm2 end to end with the genotypes resident in HBM:
thin precompute, randomized SVD, the association pass over every SNP,
t^2 histogram, GIF and calibrated p-values (outputs stay in HBM).
The headline value is the mixed-precision path; the FP64 GPU path is timed
beside it and the max |delta log10 p| between the two is reported.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000)
    ap.add_argument("--p", type=int, default=10_000_000)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--d", type=int, default=4)
    ap.add_argument("--est-rate", type=float, default=0.01)
    ap.add_argument("--power-iter", type=int, default=3)
    ap.add_argument("--precision", choices=["auto", "i8", "fp64"], default="auto")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-snps", type=int, default=20000,
                    help="association prefix of the CPU baseline (SNPs)")
    ap.add_argument("--cpu-est-snps", type=int, default=20000,
                    help="estimation SNPs of the CPU baseline's RSVD (scaled to the workload's subset)")
    ap.add_argument("--workload", choices=["lfmm2", "nmf", "batch", "host", "mfma"], default="lfmm2",
                    help="lfmm2: the headline fit (default); nmf: NMF imputation estimation and the "
                         "NMF-imputed fit at n=5000, p=2M, nmf-k=8, 10 iterations, K=8, d=2 (FP64); "
                         "batch: 64 independent fits of seeded uneven shapes, batched vs looped; "
                         "host: the headline fit with the genotypes in host memory, streamed over "
                         "PCIe chunk by chunk (PCIe-inclusive rate); "
                         "mfma: sustained int8 / FP8 / BF16 MFMA rates and the cost per genotype fragment of "
                         "each operand format of the precision comparison")
    ap.add_argument("--host-memory", choices=["pageable", "pinned"], default="pageable",
                    help="host buffer of the genotypes (--workload host)")
    ap.add_argument("--chunk-snps", type=int, default=0, help="rows per streamed chunk (--workload host)")
    ap.add_argument("--write-tsv", default="",
                    help="directory: after timing, write the results TSV of the main fit there (synthetic .bim "
                         "metadata), report the writer's time and delete the files")
    ap.add_argument("--streams", type=int, default=None,
                    help="streams of the batched fit (--workload batch; default: the tuning table's batch_streams)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB on the device three times (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter calibration) and exit")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        x.fill_(1.0)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read_per_copy": 1 << 30,
                          "bytes_written_per_copy": 1 << 30, "copies": 3}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note (SNP-range sharding + histogram all-reduce), not run here",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "nmf":
        nmf_workload(args)
        return
    if args.workload == "batch":
        batch_workload(args)
        return
    if args.workload == "mfma":
        mfma_workload(args)
        return
    if args.workload == "host":
        host_workload(args)
        return

    import torch
    import schnellfmm_amd as S
    from schnellfmm_amd._lib import AssocOut, lib

    torch.cuda.init()
    dev = torch.device("cuda:0")
    n, p, K, d = args.n, args.p, args.k, args.d
    ctx = S.Context(0)
    t0 = time.time()
    log(f"[bench] generating n={n} p={p} panel in HBM ...")
    geno = S.BedFile.synthetic(n, p, n_pops=20, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02,
                               seed=3, ctx=ctx)
    torch.cuda.synchronize()
    log(f"[bench] panel ready in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(3)
    cov = np.full((d, d), 0.3) + 0.7 * np.eye(d)
    x = np.ascontiguousarray(rng.multivariate_normal(np.zeros(d), cov, size=n))
    est_idx = S.subset_indices(args.est_rate, p)

    def alloc_out():
        t = {k: torch.empty(shape, dtype=torch.float64, device=dev) for k, shape in
             (("beta", (p, d)), ("t", (p, d)), ("r2", (p, 3)), ("p", (p, d)), ("gpt", (d,)), ("gif", (1,)))}
        t["u"] = torch.empty((n, K), dtype=torch.float64, device=dev)
        o = AssocOut()
        o.beta, o.tstat, o.r2, o.pvalue = (t[k].data_ptr() for k in ("beta", "t", "r2", "p"))
        o.hist = None
        o.gif_per_trait, o.gif = t["gpt"].data_ptr(), t["gif"].data_ptr()
        return t, o

    def make_cfg(prec):
        return S.Lfmm2Config(k=K, oversampling=10, n_power_iter=args.power_iter, norm=S.EIGENSTRAT,
                             precision=prec, seed=42).to_c()

    def run(prec, tensors, out):
        cfg = make_cfg(prec)
        xs_ptr = x.ctypes.data_as(C.c_void_p)
        idx_ptr = est_idx.ctypes.data_as(C.c_void_p)
        ctx.check(lib().slfmm_fit_lfmm2(ctx.h, geno.h, idx_ptr, len(est_idx), geno.h, xs_ptr, d,
                                        C.byref(cfg), None, C.c_void_p(tensors["u"].data_ptr()),
                                        C.byref(out), 1))

    def timed(prec, steps, warmup, label):
        """Wall time per fit (no kernel timing events), then a separate pass
        with HIP-event timing of the named kernels on the library's stream."""
        tensors, out = alloc_out()
        for i in range(warmup):
            run(prec, tensors, out)
            log(f"[bench] {label} warmup {i + 1}/{warmup} done")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(steps):
            run(prec, tensors, out)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / steps
        lib().slfmm_profile_enable(ctx.h, 1)
        for name in (b"assoc_stream", b"rsvd_project", b"rsvd_forward"):
            lib().slfmm_profile_read(ctx.h, name, None, None, 1)
        prof_steps = max(1, min(steps, 3))
        for i in range(prof_steps):
            run(prec, tensors, out)
        torch.cuda.synchronize()
        kern = {}
        for name in ("assoc_stream", "rsvd_project", "rsvd_forward"):
            ms, cnt = C.c_double(), C.c_uint64()
            lib().slfmm_profile_read(ctx.h, name.encode(), C.byref(ms), C.byref(cnt), 1)
            kern[name] = (ms.value, cnt.value, prof_steps)
        lib().slfmm_profile_enable(ctx.h, 0)
        log(f"[bench] {label}: {dt * 1e3:.2f} ms/step, kernels {kern}")
        return dt, kern, tensors

    prec_main = {"i8": S.PREC_I8, "fp64": S.PREC_FP64}.get(args.precision)
    if prec_main is None:
        prec_main = S.PREC_I8
        try:
            t_, o_ = alloc_out()
            run(S.PREC_I8, t_, o_)
            del t_, o_
        except Exception as e:
            log(f"[bench] mixed path unavailable ({e}); reporting the FP64 path")
            prec_main = S.PREC_FP64
    dt, kern, res_main = timed(prec_main, args.steps, args.warmup, "main")
    result_main = {k: v for k, v in res_main.items()}

    # FP64 GPU path beside it, and the error of the mixed path against it
    fp64 = None
    max_dlp = None
    if prec_main != S.PREC_FP64 and args.fp64_steps > 0:
        dt64, kern64, res64 = timed(S.PREC_FP64, args.fp64_steps, 1, "fp64")
        fp64 = {"ms_per_step": dt64 * 1e3, "snps_per_s": p / dt64}
        import oracle as O
        t_m = result_main["t"].cpu().numpy(); g_m = result_main["gpt"].cpu().numpy()
        t_f = res64["t"].cpu().numpy(); g_f = res64["gpt"].cpu().numpy()
        lp_m = O.log10_pvalue(t_m, g_m); lp_f = O.log10_pvalue(t_f, g_f)
        max_dlp = float(np.max(np.abs(lp_m - lp_f)))
        fp64["max_abs_dlog10p_vs_fp64"] = max_dlp
        fp64["gif_mixed"] = g_m.tolist(); fp64["gif_fp64"] = g_f.tolist()
        del res64

    # roofline of the dominant kernel (association genotype stream)
    ams, acnt, _ = kern["assoc_stream"]
    avg_ms = ams / max(acnt, 1)
    bytes_per_launch = p * math.ceil(n / 4)
    flops_per_launch = p * 2.0 * n * (2 + 2 * d + K)
    if prec_main == S.PREC_FP64:
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_func_f64", "avg_ms": avg_ms, "bytes_per_launch": bytes_per_launch}
    else:
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n, p, K, d),
                "kernel": i8_kernel_name(K, d, 0), "avg_ms": avg_ms, "bytes_per_launch": bytes_per_launch}

    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline(geno, x, result_main["u"].cpu().numpy(), args, n, K, d)
    tsv = write_tsv_measure(args.write_tsv, result_main, ctx, p, d) if args.write_tsv else None

    line = {
        "metric": "SNPs/sec end-to-end (n=10k, K=20): mixed-prec vs FP64 GPU @ max|Δlog10p|≤1e-3",
        "value": p / dt, "unit": "SNPs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i8-mfma/i32-acc+f64" if prec_main == S.PREC_I8 else "f64",
        "data": "synthetic (Balding-Nichols, generated in HBM)",
        "config": {"workload": "lfmm2_fit n=10000 p=10000000 K=20 d=4 est_rate=0.01 q=3 eigenstrat",
                   "n": n, "p": p, "K": K, "d": d, "est_snps": int(len(est_idx)),
                   "power_iter": args.power_iter, "parallelism": "single-gpu"},
        "roofline": roof,
        "kernels_ms_per_step": {k: v[0] / max(v[2], 1) for k, v in kern.items()},
        "fp64_gpu": fp64,
        "cpu_baseline": cpu,
        "output_tsv": tsv,
        "launch_shapes": __import__("schnellfmm_amd.tuning", fromlist=["params"]).params(),
    }
    print(json.dumps(line), flush=True)


def nmf_workload(args):
    """NMF imputation (src/nmf.rs:80-455) and the NMF-imputed fit
    (src/lib.rs:162-263) at n=5000, p=2M (10 % missing), nmf-k=8, 10
    multiplicative-update iterations, cv-rate 5e-4, then LFMM2 with K=8, d=2.
    Both run in FP64 on the GPU with the panel resident in HBM."""
    import torch
    import schnellfmm_amd as S

    torch.cuda.init()
    n, p = 5000, 2_000_000
    ctx = S.Context(0)
    geno = S.BedFile.synthetic(n, p, n_pops=8, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.10, seed=5,
                               ctx=ctx)
    torch.cuda.synchronize()
    cfg = S.NmfParams(k=8, n_iter=10, cv_rate=5e-4, seed=42, chunk_size=10_000)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, 2))
    lc = S.Lfmm2Config(k=8, norm=S.EIGENSTRAT, chunk_size=10_000, n_power_iter=2, seed=42)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, out

    steps, warmup = max(args.steps, 1), max(args.warmup, 0)
    dt_est, est = timed(lambda: S.nmf_impute_estimation(geno, None, cfg), steps, warmup)
    log(f"[bench] nmf estimation: {dt_est * 1e3:.1f} ms")
    dt_est8, est8 = timed(lambda: S.nmf_impute_estimation(geno, None, cfg, int8=True), steps, warmup)
    log(f"[bench] nmf estimation, both passes on the int8 matrix cores: {dt_est8 * 1e3:.1f} ms")

    def pass_ms(int8):   # device time per iteration of each NMF pass, HIP events
        import ctypes as C
        from schnellfmm_amd._lib import lib
        ctx.check(lib().slfmm_profile_enable(ctx.h, 1))
        S.nmf_impute_estimation(geno, None, cfg, int8=int8)
        out = {}
        for name in ("nmf_forward", "nmf_backward", "nmf_cv"):
            tot, cnt = C.c_double(), C.c_uint64()
            ctx.check(lib().slfmm_profile_read(ctx.h, name.encode(), C.byref(tot), C.byref(cnt), 1))
            out[name] = tot.value / max(cnt.value, 1)
        ctx.check(lib().slfmm_profile_enable(ctx.h, 0))
        return out
    t64, t8 = pass_ms(False), pass_ms(True)
    log(f"[bench] per iteration, fp64 / int8: forward {t64['nmf_forward']:.2f} / {t8['nmf_forward']:.2f} ms, "
        f"backward {t64['nmf_backward']:.2f} / {t8['nmf_backward']:.2f} ms, "
        f"cv {t64['nmf_cv']:.2f} / {t8['nmf_cv']:.2f} ms")
    dt_fit, fit = timed(lambda: S.fit_lfmm2_nmf(geno, None, geno, x, lc, cfg), max(steps // 2, 1), min(warmup, 1))
    log(f"[bench] nmf fit: {dt_fit * 1e3:.1f} ms")
    dt_fit8, fit8 = timed(lambda: S.fit_lfmm2_nmf(geno, None, geno, x, lc, cfg, nmf_int8=True), max(steps // 2, 1),
                          min(warmup, 1))
    log(f"[bench] nmf fit, NMF passes on the int8 matrix cores: {dt_fit8 * 1e3:.1f} ms")
    lc8 = S.Lfmm2Config(k=8, norm=S.EIGENSTRAT, chunk_size=10_000, n_power_iter=2, seed=42, precision=S.PREC_I8)
    dt_fitm, fitm = timed(lambda: S.fit_lfmm2_nmf(geno, None, geno, x, lc8, cfg, nmf_int8=True), max(steps // 2, 1),
                          min(warmup, 1))
    log(f"[bench] nmf fit in mixed precision (int8 NMF passes and filled functionals): {dt_fitm * 1e3:.1f} ms")
    bytes_pass = p * math.ceil(n / 4)
    passes = 2 + 3 * cfg.n_iter
    line = {
        "metric": "NMF imputation estimation SNP-iterations/s (n=5000, nmf-k=8)",
        "value": p * cfg.n_iter / dt_est, "unit": "SNP-iterations/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": dt_est * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (Balding-Nichols, 10% missing, generated in HBM)",
        "config": {"workload": "nmf_impute_estimation n=5000 p=2000000 nmf_k=8 n_iter=10 cv_rate=5e-4",
                   "n": n, "p": p, "nmf_k": 8, "n_iter": cfg.n_iter},
        "roofline": {"bound": "fp64", "achieved": 8.0 * n * 8 * p * cfg.n_iter / dt_est / 1e12,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": 8.0 * n * 8 * p * cfg.n_iter / dt_est / 1e12 / FP64_PEAK_TFLOPS,
                     "traffic": None, "kernel": "nmf iteration (forward + backward + CV)",
                     "genotype_bytes_per_pass": bytes_pass, "passes": passes},
        "cv_mae_last": float(est.cv_mae_per_iter[-1]), "cv_mae_mean_impute": float(est.cv_mae_mean_impute),
        "ms_per_iter_fp64": {k[4:]: v for k, v in t64.items()},
        "int8": {"estimation_ms": dt_est8 * 1e3, "ms_per_iter": {k[4:]: v for k, v in t8.items()},
                 "forward_gb_per_s": bytes_pass / (t8["nmf_forward"] * 1e-3) / 1e9,
                 "backward_gb_per_s": bytes_pass / (t8["nmf_backward"] * 1e-3) / 1e9,
                 "cv_mae_last": float(est8.cv_mae_per_iter[-1]),
                 "max_rel_dw_vs_fp64": float(np.max(np.abs(est8.w - est.w)) / np.max(np.abs(est.w))),
                 "max_rel_dh_vs_fp64": float(np.max(np.abs(est8.h_full - est.h_full))
                                             / np.max(np.abs(est.h_full)))},
        "fit_ms": dt_fit * 1e3, "fit_snps_per_s": p / dt_fit, "fit_gif": float(fit.gif),
        "fit_int8_nmf": {"ms": dt_fit8 * 1e3, "gif": float(fit8.gif),
                         "max_abs_dlog10p_vs_fp64": float(np.max(np.abs(np.log10(fit8.pvalue)
                                                                         - np.log10(fit.pvalue))))},
        "fit_mixed": {"ms": dt_fitm * 1e3, "gif": float(fitm.gif),
                      "max_abs_dlog10p_vs_fp64": float(np.max(np.abs(np.log10(fitm.pvalue)
                                                                      - np.log10(fit.pvalue))))},
        "gwas_cv": [float(v) for v in fit.gwas_cv],
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def batch_workload(args):
    """64 independent LFMM2 problems of uneven size (gen seed 4): n ~ logU[200, 5000],
    p ~ logU[2e4, 2e6], K ~ U{3..20}, d ~ U{1..8}, 2-10 populations, 0-5 %
    missing, RSVD on a 10 % SNP subset, mixed precision, outputs in HBM.
    Timed as one slfmm_fit_lfmm2_batch call and as a loop of slfmm_fit_lfmm2."""
    import ctypes as C
    import torch
    import schnellfmm_amd as S
    from schnellfmm_amd._lib import Problem, lib

    torch.cuda.init()
    dev = torch.device("cuda:0")
    from schnellfmm_amd import synth
    ctx = S.Context(0)
    count = 64
    arr = (Problem * count)()
    keep = []
    shapes = []
    total_snps = 0
    total_bytes = 0
    for i, n, p, K, d, pops, miss, x in synth.batch_problems(seed=4, count=count):
        g = S.BedFile.synthetic(n, p, n_pops=pops, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=miss,
                                seed=4000 + i, ctx=ctx)
        idx = np.ascontiguousarray(S.subset_indices(0.1, p), dtype=np.int64)
        t = {k: torch.empty(sh, dtype=torch.float64, device=dev) for k, sh in
             (("beta", (p, d)), ("t", (p, d)), ("r2", (p, 3)), ("p", (p, d)), ("gpt", (d,)), ("gif", (1,)),
              ("u", (n, K)))}
        cfg = S.Lfmm2Config(k=K, oversampling=10, n_power_iter=2, norm=S.EIGENSTRAT, precision=S.PREC_I8,
                            seed=42).to_c()
        P = arr[i]
        P.est, P.subset_idx, P.n_idx, P.g = g.h, idx.ctypes.data, len(idx), g.h
        P.x, P.d, P.cfg, P.b0 = x.ctypes.data, d, cfg, None
        P.u_hat = t["u"].data_ptr()
        P.out.beta, P.out.tstat, P.out.r2, P.out.pvalue = (t[k].data_ptr() for k in ("beta", "t", "r2", "p"))
        P.out.hist, P.out.gif_per_trait, P.out.gif = None, t["gpt"].data_ptr(), t["gif"].data_ptr()
        keep.append((g, x, idx, t))
        shapes.append((n, p, K, d))
        total_snps += p
        total_bytes += p * math.ceil(n / 4)
    torch.cuda.synchronize()
    log(f"[bench] batch: {count} problems, {total_snps} SNPs, {total_bytes / 1e9:.2f} GB of genotypes")

    def run_loop():
        for i in range(count):
            P = arr[i]
            ctx.check(lib().slfmm_fit_lfmm2(ctx.h, P.est, P.subset_idx, P.n_idx, P.g, P.x, P.d, C.byref(P.cfg),
                                            None, P.u_hat, C.byref(P.out), 1))

    def run_batch(ns):
        ctx.check(lib().slfmm_fit_lfmm2_batch(ctx.h, arr, count, ns, 1))

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    from schnellfmm_amd import tuning
    if args.streams is None:
        args.streams = tuning.params()["batch_streams"]
    steps, warmup = max(args.steps, 1), max(args.warmup, 1)
    dt_loop = timed(run_loop, steps, warmup)
    log(f"[bench] loop of single fits: {dt_loop * 1e3:.1f} ms")
    dt_b = {}
    for ns in sorted({1, args.streams}):
        dt_b[ns] = timed(lambda: run_batch(ns), steps, warmup)
        log(f"[bench] batched, {ns} stream(s): {dt_b[ns] * 1e3:.1f} ms")
    dt = dt_b[args.streams]
    line = {
        "metric": "LFMM2 batch SNPs/s (64 problems of uneven size, mixed precision)",
        "value": total_snps / dt, "unit": "SNPs/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i8-mfma/i32-acc+f64", "data": "synthetic (Balding-Nichols, generated in HBM)",
        "config": {"workload": "batch of 64 lfmm2 fits, gen seed 4", "problems": count, "snps": total_snps,
                   "genotype_bytes": total_bytes, "streams": args.streams, "est_rate": 0.1},
        "loop_ms": dt_loop * 1e3, "batch_ms_by_streams": {str(k): v * 1e3 for k, v in dt_b.items()},
        "speedup_vs_loop": dt_loop / dt, "roofline": None, "cpu_baseline": None,
        "launch_shapes": tuning.params(),
    }
    print(json.dumps(line), flush=True)


def write_tsv_measure(dirpath, res, ctx, p, d):
    """The output stage after the fit: slfmm_write_tsv straight from the HBM
    outputs (p-values, betas, t, r2) with a synthetic .bim of p records."""
    import shutil
    import tempfile
    from schnellfmm_amd._lib import lib
    tmp = tempfile.mkdtemp(dir=dirpath)
    try:
        bim = os.path.join(tmp, "panel.bim")
        with open(bim, "w") as f:
            for j0 in range(0, p, 1_000_000):
                f.write("".join(f"1\tsnp{j}\t0\t{j + 1}\tA\tG\n" for j in range(j0, min(p, j0 + 1_000_000))))
        out = os.path.join(tmp, "results.tsv")
        t0 = time.perf_counter()
        ctx.check(lib().slfmm_write_tsv(ctx.h, out.encode(), bim.encode(), None, None, None, p, d, None,
                                        res["p"].data_ptr(), res["beta"].data_ptr(), res["t"].data_ptr(),
                                        res["r2"].data_ptr(), 1, 0))
        dt = time.perf_counter() - t0
        size = os.path.getsize(out)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    log(f"[bench] results TSV: {size / 1e9:.2f} GB in {dt:.2f}s")
    return {"seconds": dt, "bytes": size, "snps_per_s": p / dt, "threads": min(16, os.cpu_count() or 1),
            "includes": "D2H copy of the outputs, .bim parse, formatting, file write"}


def host_workload(args):
    """The headline fit (same panel, covariates and configuration) with the
    genotypes in host memory: every step streams the 1 % estimation rows and
    then all p rows over PCIe in double-buffered chunks, the copies
    overlapping the association of the previous chunk."""
    import torch
    import schnellfmm_amd as S
    from schnellfmm_amd._lib import AssocOut, lib

    torch.cuda.init()
    dev = torch.device("cuda:0")
    n, p, K, d = args.n, args.p, args.k, args.d
    ctx = S.Context(0)
    t0 = time.time()
    gdev = S.BedFile.synthetic(n, p, n_pops=20, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02,
                               seed=3, ctx=ctx)
    bps = (n + 3) // 4
    if args.host_memory == "pinned":
        hbuf_t = torch.empty(p * bps, dtype=torch.uint8, pin_memory=True)
        hbuf = hbuf_t.numpy()
    else:
        hbuf_t = None
        hbuf = np.empty(p * bps, np.uint8)
    ctx.check(lib().slfmm_geno_read_packed(ctx.h, gdev.h, 0, p, hbuf.ctypes.data_as(C.c_void_p)))
    gdev.free()
    geno = S.BedFile.wrap_host(hbuf, n, p, chunk_snps=args.chunk_snps, ctx=ctx)
    log(f"[bench] host panel ({args.host_memory}, {p * bps / 1e9:.1f} GB) ready in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(3)
    cov = np.full((d, d), 0.3) + 0.7 * np.eye(d)
    x = np.ascontiguousarray(rng.multivariate_normal(np.zeros(d), cov, size=n))
    est_idx = S.subset_indices(args.est_rate, p)
    t = {k: torch.empty(shape, dtype=torch.float64, device=dev) for k, shape in
         (("beta", (p, d)), ("t", (p, d)), ("r2", (p, 3)), ("p", (p, d)), ("gpt", (d,)), ("gif", (1,)),
          ("u", (n, K)))}
    o = AssocOut()
    o.beta, o.tstat, o.r2, o.pvalue = (t[k].data_ptr() for k in ("beta", "t", "r2", "p"))
    o.hist, o.gif_per_trait, o.gif = None, t["gpt"].data_ptr(), t["gif"].data_ptr()
    cfg = S.Lfmm2Config(k=K, oversampling=10, n_power_iter=args.power_iter, norm=S.EIGENSTRAT,
                        precision=S.PREC_I8, seed=42).to_c()

    def run():
        ctx.check(lib().slfmm_fit_lfmm2(ctx.h, geno.h, est_idx.ctypes.data_as(C.c_void_p), len(est_idx), geno.h,
                                        x.ctypes.data_as(C.c_void_p), d, C.byref(cfg), None,
                                        C.c_void_p(t["u"].data_ptr()), C.byref(o), 1))

    for i in range(args.warmup):
        run()
        log(f"[bench] host warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(args.steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t1) / args.steps
    log(f"[bench] host panel: {dt * 1e3:.1f} ms/step")
    line = {
        "metric": "SNPs/sec end-to-end (n=10k, K=20), genotypes in host memory (PCIe-inclusive)",
        "value": p / dt, "unit": "SNPs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i8-mfma/i32-acc+f64", "data": "synthetic (Balding-Nichols, generated on the device, "
                                                 "copied to host memory before timing)",
        "config": {"workload": "lfmm2_fit n=10000 p=10000000 K=20 d=4 est_rate=0.01 q=3 eigenstrat, "
                               "host-resident genotypes", "n": n, "p": p, "K": K, "d": d,
                   "host_memory": args.host_memory, "chunk_snps": args.chunk_snps,
                   "genotype_bytes": p * bps, "pcie_gbs": p * bps / dt / 1e9},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    geno.free()
    del hbuf_t


# Operand terms per genotype fragment of each format in the precision
# comparison (the dense terms against the dosage operand plus those against
# the missing/observed operand that the centring needs), as in its cost
# model; int8 is the library's layout at K = 20, d = 4.
PRECISION_FORMATS = [
    ("int8 2/3 terms (library)", 0, (2 * 20 + 3 * 4) / 24 + 2),
    ("bf16 hi+lo, fp32 acc", 2, 2 + 2),
    ("bf16 1 term, fp32 acc", 2, 1 + 1),
    ("fp8 e4m3 x2, fp32 acc", 1, 2 + 2),
    ("fp8 e4m3 x3, fp32 acc", 1, 3 + 3),
]


def mfma_workload(args):
    """Sustained dense MFMA rate of the three formats that hold genotype codes
    exactly (slfmm_debug_mfma_rate: whole chip, 2 waves per SIMD, 4 chains per
    wave), and the time per genotype fragment (32 SNPs x 32 samples against
    one 32-row dense tile) of each format of the precision comparison:
    terms x 2*32^3 operations / measured rate."""
    import ctypes as C
    import torch
    import schnellfmm_amd as S
    from schnellfmm_amd._lib import lib
    torch.cuda.init()
    ctx = S.Context(0)
    names = {0: "int8 (v_mfma_i32_32x32x32_i8)", 1: "fp8 e4m3 (v_mfma_scale_f32_32x32x64_f8f6f4)",
             2: "bf16 (v_mfma_f32_32x32x16_bf16)"}
    rate = {}
    for kind in (0, 1, 2):
        best = 0.0
        for _ in range(max(args.steps, 1)):
            t = C.c_double()
            ctx.check(lib().slfmm_debug_mfma_rate(ctx.h, kind, 200_000, C.byref(t)))
            best = max(best, t.value)
        rate[kind] = best
        log(f"[bench] {names[kind]}: {best:.0f} TOPS sustained")
    frag_ops = 2.0 * 32 ** 3
    fmts = []
    for name, kind, terms in PRECISION_FORMATS:
        ns = terms * frag_ops / (rate[kind] * 1e12) * 1e9
        fmts.append({"format": name, "operand_terms": round(terms, 2), "mfma": names[kind],
                     "ns_per_fragment_chip": ns, "relative_to_int8": None})
    base = fmts[0]["ns_per_fragment_chip"]
    for f in fmts:
        f["relative_to_int8"] = f["ns_per_fragment_chip"] / base
    peaks = {0: 5000.0, 1: 5000.0, 2: 2500.0}
    line = {
        "metric": "sustained dense MFMA rate by operand format (TOPS)", "value": rate[0], "unit": "TOPS",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "higher_is_better": True, "vs_baseline": None,
        "dtype": "i8 / fp8 e4m3 / bf16", "data": "fixed operand bytes",
        "config": {"workload": "mfma rates: 2 waves per SIMD, 4 accumulator chains per wave, 200000 MFMAs each"},
        "rates_tops": {names[k]: rate[k] for k in rate},
        "fraction_of_dense_peak": {names[k]: rate[k] / peaks[k] for k in rate},
        "formats": fmts, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def i8_kernel_name(K, d, mode):
    """Template instance of the int8 SNP-side kernel that runs the pass
    (layout from the library: k_snp_i8<NQ, NT0, NT1, ST, MODE>)."""
    from schnellfmm_amd._lib import lib
    m = K + d
    dims = np.zeros(5, np.int32)
    lib().slfmm_debug_i8_layout(m, K if mode == 0 else m, K if mode == 0 else m,
                                dims.ctypes.data_as(C.c_void_p), None, 0)
    nq, nt0, nt1 = (int(v) for v in dims[:3])
    st = 2 if nt0 + nt1 <= 3 else 1
    return f"k_snp_i8<{nq}, {nt0}, {nt1}, {st}, {mode}>"


def measured_traffic(n, p, K, d):
    """HBM bytes per launch of the association kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (counters
    scaled by the device-copy calibration); None if no summary matches."""
    path = os.path.join(ROOT, "profiles", "k_snp_i8_traffic_n10000_p10M.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    w = rec.get("workload", {})
    if (w.get("n"), w.get("p"), w.get("K"), w.get("d")) != (n, p, K, d):
        return None
    return rec.get("hbm_bytes_per_launch")


def _host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(geno, x, u_hat, args, n, K, d):
    """CPU restatement of the reference algorithm (oracle/, numpy + OpenBLAS)
    on a bounded sample of the same workload: the randomized SVD on a
    --cpu-est-snps subset of the estimation SNPs (its cost is linear in them,
    so the time is scaled to the workload's subset) and the dense
    reference-form association on a --cpu-snps prefix (scaled to p).  The
    value is p / (scaled RSVD + scaled association) in SNPs/s."""
    import oracle as O
    try:
        from threadpoolctl import threadpool_info
        info = threadpool_info()
        threads = max((i.get("num_threads", 1) for i in info), default=1)
        blas = sorted({f"{i.get('internal_api')} {i.get('version')}" for i in info})
    except Exception:
        threads, blas = os.cpu_count() or 1, []
    p = geno.n_snps
    p_est = len(np.arange(0, p, int(math.ceil(1.0 / args.est_rate))))
    xs = O.center_covariates(x)
    # RSVD on a sample of the estimation SNPs (every 100th SNP, the first ones)
    me = min(int(args.cpu_est_snps), p_est)
    step = int(math.ceil(1.0 / args.est_rate))
    est = np.arange(0, me * step, step)
    raw = geno.decode_raw(est)
    packed_e = O.encode_bed(np.where(np.isnan(raw), 255, raw).astype(np.uint8))
    del raw
    bps = O.bytes_per_snp(n)
    omega = np.random.default_rng(1).standard_normal((n, K + 10))
    t0 = time.perf_counter()
    pre = O.precompute(xs, 1e-5)
    t1 = time.perf_counter()
    O.estimate_factors_streaming(packed_e, bps, n, np.arange(me), pre, K, 10, args.power_iter, 10_000,
                                 O.EIGENSTRAT, O.Impute(), omega=omega)
    t2 = time.perf_counter()
    # association on a prefix, with the GPU's U_hat
    m = min(int(args.cpu_snps), p)
    raw = geno.decode_raw(np.arange(m))
    packed = O.encode_bed(np.where(np.isnan(raw), 255, raw).astype(np.uint8))
    del raw
    t3 = time.perf_counter()
    O.test_associations_fused(packed, bps, n, m, xs, u_hat, pre, K, 10_000, O.EIGENSTRAT, O.Impute())
    t4 = time.perf_counter()
    rsvd_full = (t1 - t0) + (t2 - t1) * p_est / me
    assoc_full = (t4 - t3) * p / m
    return {"value": p / (rsvd_full + assoc_full), "unit": "SNPs/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle/ float64 restatement of the reference algorithm at n={n}, K={K}, d={d}: "
                      f"precompute + RSVD (q={args.power_iter}) on {me} of the {p_est} estimation SNPs, "
                      f"scaled x{p_est / me:.0f}, and the dense association on the first {m} SNPs, "
                      f"scaled x{p / m:.0f}; chunks of 10000",
            "seconds": {"precompute": t1 - t0, "rsvd_sample": t2 - t1, "assoc_sample": t4 - t3,
                        "rsvd_scaled": rsvd_full, "assoc_scaled": assoc_full},
            "host": {"cpu": _host_cpu(), "nproc": os.cpu_count(), "blas": blas, "threads": int(threads)}}


if __name__ == "__main__":
    main()
