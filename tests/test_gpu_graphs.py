"""Graph replay of fits (slfmm_fit_lfmm2 and slfmm_fit_lfmm2_batch with
device outputs).  A fit with device-resident panels and device outputs is
captured into a HIP graph by the call after the one that sized its buffers,
and later calls with the same shapes, configuration and pointers replay it.
Every replay must equal, bit for bit, the same fit queued kernel by kernel
(SLFMM_EAGER): with eager work and synchronisations between replays, with
new input values behind the same pointers (covariates, estimation subset),
in both precisions, and for batches spread over several streams."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUT_DEVICE, EAGER = 1, 2


@pytest.fixture(scope="module")
def env():
    import schnellfmm_amd as S
    from schnellfmm_amd._lib import AssocOut, lib
    lib()   # loads the HIP runtime the library uses; the buffers below come from the same one
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return S, hip, AssocOut, lib


def _hip_ok(rc):
    assert rc == 0, f"HIP error {rc}"


class DevOut:
    """Device output buffers of one fit (hipMalloc) and the slfmm_assoc_out over them."""

    def __init__(self, env, n, p, K, d):
        S, hip, AssocOut, _ = env
        self.hip = hip
        self.shapes = {"beta": (p, d), "t": (p, d), "r2": (p, 3), "p": (p, d), "gpt": (d,), "gif": (1,), "u": (n, K)}
        self.ptr = {}
        for k, sh in self.shapes.items():
            v = C.c_void_p()
            _hip_ok(hip.hipMalloc(C.byref(v), 8 * int(np.prod(sh))))
            self.ptr[k] = v.value
        o = AssocOut()
        o.beta, o.tstat, o.r2, o.pvalue = (self.ptr[k] for k in ("beta", "t", "r2", "p"))
        o.hist = None
        o.gif_per_trait, o.gif = self.ptr["gpt"], self.ptr["gif"]
        self.o = o

    def poison(self):   # all-ones bytes: NaN in every double
        for k, sh in self.shapes.items():
            _hip_ok(self.hip.hipMemset(self.ptr[k], 0xFF, 8 * int(np.prod(sh))))
        _hip_ok(self.hip.hipDeviceSynchronize())   # done before the fit's own stream starts

    def host(self):
        out = {}
        for k, sh in self.shapes.items():
            a = np.empty(sh)
            _hip_ok(self.hip.hipMemcpy(a.ctypes.data, self.ptr[k], a.nbytes, 2))   # device to host
            out[k] = a
        return out

    def free(self):
        for v in self.ptr.values():
            self.hip.hipFree(v)


def _same(a, b, label):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), f"{label}: {k} differs"
        assert np.all(np.isfinite(a[k])), f"{label}: {k} not finite"


def _stats(lib, ctx):
    c, r = C.c_uint64(), C.c_uint64()
    ctx.check(lib.slfmm_graph_stats(ctx.h, C.byref(c), C.byref(r)))
    return c.value, r.value


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", ["i8", "fp64"])
def test_fit_graph_replays_equal_eager(env, prec):
    S, hip, _, libf = env
    lib = libf()
    n, p, K, d = 1500, 30000, 8, 3
    _log(f"[graphs] fit prec={prec} n={n} p={p} K={K} d={d}")
    ctx = S.Context(0)
    geno = S.BedFile.synthetic(n, p, n_pops=5, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02, seed=11, ctx=ctx)
    rng = np.random.default_rng(11)
    x = np.ascontiguousarray(rng.standard_normal((n, d)))
    x_new = np.ascontiguousarray(rng.standard_normal((n, d)))
    idx = np.ascontiguousarray(S.subset_indices(0.1, p), dtype=np.int64)
    idx_new = np.ascontiguousarray(np.sort(rng.choice(p, size=len(idx), replace=False)), dtype=np.int64)
    cfg = S.Lfmm2Config(k=K, oversampling=10, n_power_iter=2, norm=S.EIGENSTRAT, seed=5,
                        precision=S.PREC_I8 if prec == "i8" else S.PREC_FP64).to_c()
    out = DevOut(env, n, p, K, d)
    side = DevOut(env, n, p, K, d)

    def fit(o, flags, xx=x, ii=idx):
        o.poison()
        ctx.check(lib.slfmm_fit_lfmm2(ctx.h, geno.h, ii.ctypes.data_as(C.c_void_p), len(ii), geno.h,
                                      xx.ctypes.data_as(C.c_void_p), d, C.byref(cfg), None,
                                      C.c_void_p(o.ptr["u"]), C.byref(o.o), flags))
        return o.host()

    ref = fit(out, OUT_DEVICE | EAGER)
    c0, r0 = _stats(lib, ctx)
    for rep in range(5):
        got = fit(out, OUT_DEVICE)
        _same(got, ref, f"replay {rep}")
        # eager work of the program between replays: another fit, a memset, a sync
        fit(side, OUT_DEVICE | EAGER, xx=x_new)
        side.poison()
        _hip_ok(hip.hipDeviceSynchronize())
    c1, r1 = _stats(lib, ctx)
    _log(f"[graphs] captured {c1 - c0}, replayed {r1 - r0}")
    assert c1 - c0 >= 1 and r1 - r0 >= 3, "the fit was not replayed from a graph"
    # new values behind the same pointers
    x[:] = x_new
    got = fit(out, OUT_DEVICE)
    _same(got, fit(side, OUT_DEVICE | EAGER), "new covariates")
    idx[:] = idx_new
    got = fit(out, OUT_DEVICE)
    _same(got, fit(side, OUT_DEVICE | EAGER), "new estimation subset")
    assert _stats(lib, ctx)[1] >= r1 + 2
    out.free()
    side.free()
    geno.free()


def test_batch_graph_replays_equal_eager(env):
    S, hip, _, libf = env
    lib = libf()
    from schnellfmm_amd._lib import Problem
    shapes = [(700, 9000, 6, 2, 0.1), (1200, 5000, 12, 4, None), (333, 20001, 20, 1, 0.25), (517, 7013, 3, 3, 0.5),
              (2000, 3000, 9, 8, 0.2)]
    _log(f"[graphs] batch of {len(shapes)} problems on 3 streams")
    ctx = S.Context(0)
    arr = (Problem * len(shapes))()
    keep, outs = [], []
    for i, (n, p, K, d, rate) in enumerate(shapes):
        g = S.BedFile.synthetic(n, p, n_pops=4, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.03, seed=20 + i, ctx=ctx)
        rng = np.random.default_rng(20 + i)
        x = np.ascontiguousarray(rng.standard_normal((n, d)))
        idx = None if rate is None else np.ascontiguousarray(S.subset_indices(rate, p), dtype=np.int64)
        o = DevOut(env, n, p, K, d)
        cfg = S.Lfmm2Config(k=K, oversampling=10, n_power_iter=2, norm=S.EIGENSTRAT, seed=9 + i,
                            precision=S.PREC_I8 if i % 2 == 0 else S.PREC_FP64).to_c()
        P = arr[i]
        P.est, P.subset_idx, P.n_idx = g.h, (None if idx is None else idx.ctypes.data), 0 if idx is None else len(idx)
        P.g, P.x, P.d, P.cfg, P.b0 = g.h, x.ctypes.data, d, cfg, None
        P.u_hat = o.ptr["u"]
        P.out = o.o
        keep.append((g, x, idx))
        outs.append(o)

    def run(flags):
        for o in outs:
            o.poison()
        ctx.check(lib.slfmm_fit_lfmm2_batch(ctx.h, arr, len(shapes), 3, flags))
        return [o.host() for o in outs]

    ref = run(OUT_DEVICE | EAGER)
    c0, r0 = _stats(lib, ctx)
    for rep in range(4):
        got = run(OUT_DEVICE)
        for i, (a, b) in enumerate(zip(got, ref)):
            _same(a, b, f"replay {rep} problem {i}")
        run(OUT_DEVICE | EAGER)   # eager work between replays
        _hip_ok(hip.hipDeviceSynchronize())
    c1, r1 = _stats(lib, ctx)
    _log(f"[graphs] batch: captured {c1 - c0}, replayed {r1 - r0}")
    assert r1 - r0 >= 2 * len(shapes), "the batch was not replayed from graphs"
    for o in outs:
        o.free()
    for g, _, _ in keep:
        g.free()
