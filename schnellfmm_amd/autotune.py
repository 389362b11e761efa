"""Offline tuning pass for the launch shapes of the HIP library (slfmm_tuning).

    python -m schnellfmm_amd.autotune [--out schnellfmm_amd/tuning.json] [--quick]

Runs on the GPU.  For each launch-shape choice it times the candidates on
synthetic panels held in HBM and writes the table that every Context of this
package applies (schnellfmm_amd.tuning), together with the timings behind it:

* snp_two_tiles_min_blocks: the int8 SNP-side kernels (association pass and
  RSVD sketch) with 1 and with 2 32-SNP tiles per wave, over panels from
  1 000 to 10 000 samples and 20 000 to 2M SNPs; 2 tiles are used from the
  smallest 2-tile grid work (workgroups x 256-sample blocks) at and above
  which 2 tiles are never slower (within 1 %).
* fwd_target_wg: the int8 RSVD forward pass at 256 .. 2048 target
  workgroups, at two estimation shapes; the target with the least total
  time.
* gram_max_blocks: the whole randomized SVD with 256 .. 2048 Gram blocks.
* batch_streams, batch_fit_ms: the batched fit of the 64 seeded problems of
  uneven shape with 4 .. 16 streams and three per-fit cost estimates.

The tile choice gives bit-identical results (the per-SNP arithmetic does not
depend on it); the workgroup targets change only how FP64 partial sums are
grouped before they are added (rounding level).  The GPU tests check both.
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

HERE = os.path.dirname(os.path.abspath(__file__))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---- decision rules (pure functions of the measurements; CPU-tested) ------------

def pick_two_tile_threshold(rows, tol=0.01):
    """rows: [(work2, ms_one_tile, ms_two_tiles)], one per measured shape, where
    work2 is the work of the 2-tile launch (workgroups x 256-sample blocks).
    Returns the smallest work w such that every shape with work2 >= w is no
    slower with 2 tiles (ms2 <= ms1 (1 + tol)); 0 when that holds for every
    shape."""
    sizes = sorted({g for g, _, _ in rows})
    for g in sizes:
        if all(m2 <= m1 * (1.0 + tol) for gg, m1, m2 in rows if gg >= g):
            return 0 if g == sizes[0] else int(g)
    return int(sizes[-1]) + 1   # one tile everywhere measured


def pick_min_total(times):
    """times: {candidate: [ms per shape]} -> the candidate with the least total
    time over the shapes (the larger problems weigh most; ties: the smaller
    candidate)."""
    cands = sorted(times)
    return min(cands, key=lambda c: (round(sum(times[c]), 9), c))


def pick_batch(times):
    """times: {(streams, fit_ms): ms} -> (streams, fit_ms) of the fastest run
    (ties within 1 %: fewer streams, then the smaller estimate)."""
    best = min(times.values())
    ok = [k for k, v in times.items() if v <= best * 1.01]
    return min(ok)


# ---- GPU measurements ----------------------------------------------------------

def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


class _Bench:
    def __init__(self):
        import torch
        import schnellfmm_amd as S
        from schnellfmm_amd import tuning
        from schnellfmm_amd._lib import lib
        torch.cuda.init()
        self.torch, self.S, self.tuning, self.lib = torch, S, tuning, lib()
        self.ctx = S.Context(0)
        self.base = dict(tuning.DEFAULTS)

    def set(self, **kw):
        p = dict(self.base)
        p.update(kw)
        self.tuning.apply(self.ctx, p)

    def kernel_ms(self, name, fn, reps):
        """Median per-launch device time of the named kernel over reps calls of fn."""
        L, ctx = self.lib, self.ctx
        ctx.check(L.slfmm_profile_enable(ctx.h, 1))
        fn()   # warm-up (buffers, code objects)
        tot, n = C.c_double(), C.c_uint64()
        ctx.check(L.slfmm_profile_read(ctx.h, name.encode(), C.byref(tot), C.byref(n), 1))
        per = []
        for _ in range(reps):
            fn()
            ctx.check(L.slfmm_profile_read(ctx.h, name.encode(), C.byref(tot), C.byref(n), 1))
            per.append(tot.value / max(n.value, 1))
        ctx.check(L.slfmm_profile_enable(ctx.h, 0))
        return _median(per)

    def wall_ms(self, fn, reps):
        torch = self.torch
        fn()
        torch.cuda.synchronize()
        per = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            per.append((time.perf_counter() - t0) * 1e3)
        return _median(per)


def measure_tiles(B, quick):
    """Both int8 SNP-side kernels with 1 and 2 tiles per wave: the
    association pass (kernel timer "assoc_stream") and the RSVD sketch
    ("rsvd_project", with its operand packing)."""
    S = B.S
    rows = []
    assoc = [(1000, 5, 2, [20_000, 50_000, 200_000, 500_000] + ([] if quick else [2_000_000])),
             (5000, 20, 4, [20_000, 50_000, 100_000, 200_000] + ([] if quick else [500_000])),
             (10_000, 20, 4, [20_000, 50_000] + ([] if quick else [100_000, 200_000]))]
    for n, K, d, ps in assoc:
        rng = np.random.default_rng(n + K)
        u = np.ascontiguousarray(rng.standard_normal((n, K)))
        x = np.ascontiguousarray(rng.standard_normal((n, d)))
        xs = S.center_covariates(x)
        cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, precision=S.PREC_I8)
        for p in ps:
            g = S.BedFile.synthetic(n, p, n_pops=4, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02,
                                    seed=p % 997, ctx=B.ctx)
            fn = (lambda g=g: S.test_associations_fused(g, xs, u, cfg))
            res = {}
            for label, thr in (("one", 2 ** 32 - 1), ("two", 0)):
                B.set(snp_two_tiles_min_blocks=thr)
                res[label] = B.kernel_ms("assoc_stream", fn, 5)
            work2 = math.ceil(p / 256) * math.ceil(n / 256)
            log(f"[autotune] tiles association n={n} m={K + d} p={p} work2={work2}: one {res['one']:.3f} ms, "
                f"two {res['two']:.3f} ms")
            rows.append((work2, res["one"], res["two"], "association", n, K + d, p))
            del g
    for n, pe in ((2000, 20_000), (10_000, 100_000)) if not quick else ((10_000, 100_000),):
        g = S.BedFile.synthetic(n, pe, n_pops=8, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02, seed=5,
                                ctx=B.ctx)
        x = np.ascontiguousarray(np.random.default_rng(6).standard_normal((n, 2)))
        cfg = S.Lfmm2Config(k=20, oversampling=10, n_power_iter=1, norm=S.EIGENSTRAT, precision=S.PREC_I8)
        fn = (lambda g=g: S.estimate_factors(g, "all", x, cfg))
        res = {}
        for label, thr in (("one", 2 ** 32 - 1), ("two", 0)):
            B.set(snp_two_tiles_min_blocks=thr)
            res[label] = B.kernel_ms("rsvd_project", fn, 3)
        work2 = math.ceil(pe / 256) * math.ceil(n / 256)
        log(f"[autotune] tiles sketch n={n} l=30 p_est={pe} work2={work2}: one {res['one']:.3f} ms, "
            f"two {res['two']:.3f} ms")
        rows.append((work2, res["one"], res["two"], "rsvd sketch", n, 30, pe))
        del g
    B.set()
    return rows


def measure_fwd(B, quick):
    S = B.S
    cands = [256, 512, 1024, 2048]
    shapes = [(2000, 20_000), (10_000, 100_000)]
    times = {c: [] for c in cands}
    for n, pe in shapes:
        g = S.BedFile.synthetic(n, pe, n_pops=8, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02, seed=7,
                                ctx=B.ctx)
        x = np.ascontiguousarray(np.random.default_rng(3).standard_normal((n, 2)))
        cfg = S.Lfmm2Config(k=20, oversampling=10, n_power_iter=3, norm=S.EIGENSTRAT, precision=S.PREC_I8)
        fn = (lambda g=g: S.estimate_factors(g, "all", x, cfg))
        for c in cands:
            B.set(fwd_target_wg=c)
            times[c].append(B.kernel_ms("rsvd_forward", fn, 3))
            log(f"[autotune] fwd n={n} p_est={pe} target={c}: {times[c][-1]:.3f} ms")
        del g
    B.set()
    return times


def measure_gram(B, quick):
    S = B.S
    cands = [256, 512, 1024, 2048]
    n, pe = 10_000, 100_000
    g = S.BedFile.synthetic(n, pe, n_pops=8, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02, seed=9, ctx=B.ctx)
    x = np.ascontiguousarray(np.random.default_rng(4).standard_normal((n, 2)))
    cfg = S.Lfmm2Config(k=20, oversampling=10, n_power_iter=3, norm=S.EIGENSTRAT, precision=S.PREC_I8)
    times = {}
    for c in cands:
        B.set(gram_max_blocks=c)
        times[c] = [B.wall_ms(lambda: S.estimate_factors(g, "all", x, cfg), 3 if quick else 5)]
        log(f"[autotune] gram blocks={c}: RSVD {times[c][0]:.3f} ms")
    B.set()
    return times


def measure_batch(B, quick):
    S, torch = B.S, B.torch
    from schnellfmm_amd import synth
    from schnellfmm_amd._lib import Problem
    count = 32 if quick else 64
    arr = (Problem * count)()
    keep = []
    dev = torch.device("cuda:0")
    for i, n, p, K, d, pops, miss, x in synth.batch_problems(seed=4, count=count):
        g = S.BedFile.synthetic(n, p, n_pops=pops, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=miss,
                                seed=4000 + i, ctx=B.ctx)
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
    torch.cuda.synchronize()
    times = {}
    for fit_ms in (0.7, 1.4, 2.8):
        for ns in (4, 8, 12, 16):
            B.set(batch_fit_ms=fit_ms)
            fn = (lambda ns=ns: B.ctx.check(B.lib.slfmm_fit_lfmm2_batch(B.ctx.h, arr, count, ns, 1)))
            times[(ns, fit_ms)] = B.wall_ms(fn, 3 if quick else 5)
            log(f"[autotune] batch of {count}: {ns} streams, fit estimate {fit_ms} ms: {times[(ns, fit_ms)]:.2f} ms")
    B.set()
    return times, count


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    ap.add_argument("--quick", action="store_true", help="fewer shapes and repeats")
    args = ap.parse_args(argv)
    os.environ["SLFMM_TUNING"] = "none"   # measure from the library defaults
    B = _Bench()
    t0 = time.perf_counter()
    tile_rows = measure_tiles(B, args.quick)
    thr = pick_two_tile_threshold([(w, m1, m2) for w, m1, m2, *_ in tile_rows])
    fwd = measure_fwd(B, args.quick)
    fwd_pick = pick_min_total(fwd)
    gram = measure_gram(B, args.quick)
    gram_pick = pick_min_total(gram)
    batch, count = measure_batch(B, args.quick)
    streams, fit_ms = pick_batch(batch)
    props = B.torch.cuda.get_device_properties(0)
    table = {
        "device": f"{props.name} ({getattr(props, 'gcnArchName', '')})",
        "command": "python -m schnellfmm_amd.autotune" + (" --quick" if args.quick else ""),
        "seconds": round(time.perf_counter() - t0, 1),
        "params": {"snp_two_tiles_min_blocks": thr, "fwd_target_wg": fwd_pick, "gram_max_blocks": gram_pick,
                   "batch_fit_ms": fit_ms, "batch_gb_per_ms": B.base["batch_gb_per_ms"], "batch_streams": streams},
        "measurements": {
            "snp_kernel_ms": [{"kernel": kname, "n": n, "m": m, "p": p, "blocks_two_tiles": w,
                               "one_tile": round(m1, 4), "two_tiles": round(m2, 4)}
                              for w, m1, m2, kname, n, m, p in tile_rows],
            "rsvd_forward_ms": {"shapes": ["n=2000 p_est=20000", "n=10000 p_est=100000"],
                                **{str(c): [round(v, 4) for v in vs] for c, vs in fwd.items()}},
            "rsvd_wall_ms_n10000_pest100000": {str(c): round(v[0], 3) for c, v in gram.items()},
            f"batch_of_{count}_wall_ms": {f"{ns} streams, fit {fm} ms": round(v, 2) for (ns, fm), v in batch.items()},
        },
    }
    from schnellfmm_amd.tuning import validate
    validate(table["params"])
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")
    log(f"[autotune] wrote {args.out}: {table['params']}")
    print(json.dumps(table["params"]))


if __name__ == "__main__":
    main()
