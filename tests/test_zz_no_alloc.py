"""The allocation-free association pass (slfmm_test_associations_noalloc)
through the C ABI against slfmm_test_associations on the same inputs: every
output bit-identical (the same kernels run on the same data), on two calls in
a row, the second with a smaller panel, with no change in the device's free
memory across a call; and the argument checks that reject what the static
scratch cannot serve."""
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


def _case(S, n, p, K, d, seed):
    import torch
    pan = synth.balding_nichols(n, p, n_pops=4, miss_rate=0.03, d=d, n_causal=8, effect=1.0, seed=seed)
    g = S.BedFile.from_packed(O.encode_bed(pan.g), n, p)
    xs = np.ascontiguousarray(O.center_covariates(pan.x))
    u = np.ascontiguousarray(np.random.default_rng(seed).standard_normal((n, K)))
    cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, precision=S.PREC_I8).to_c()
    return g, xs, u, cfg


def _outputs(p, d):
    import torch
    from schnellfmm_amd._lib import AssocOut
    dev = torch.device("cuda:0")
    t = {k: torch.full(s, float("nan"), dtype=torch.float64, device=dev) for k, s in
         (("beta", (p, d)), ("t", (p, d)), ("r2", (p, 3)), ("p", (p, d)), ("gpt", (d,)), ("gif", (1,)))}
    t["hist"] = torch.zeros((d, 100_000), dtype=torch.int64, device=dev)
    o = AssocOut()
    o.beta, o.tstat, o.r2, o.pvalue = (t[k].data_ptr() for k in ("beta", "t", "r2", "p"))
    o.hist, o.gif_per_trait, o.gif = t["hist"].data_ptr(), t["gpt"].data_ptr(), t["gif"].data_ptr()
    return t, o


def _run(S, fn, g, xs, u, cfg, d, flags=1, bufs=None):
    from schnellfmm_amd._lib import lib
    t, o = bufs if bufs is not None else _outputs(g.n_snps, d)
    rc = getattr(lib(), fn)(g.ctx.h, g.h, xs.ctypes.data_as(C.c_void_p), d, u.ctypes.data_as(C.c_void_p),
                            C.byref(cfg), C.byref(o), flags)
    return rc, t


def _err(g):
    from schnellfmm_amd._lib import lib
    return lib().slfmm_last_error(g.ctx.h).decode()


def test_noalloc_equals_pooled_pass_on_two_calls(S):
    import torch
    from schnellfmm_amd._lib import lib
    for (n, p, K, d, seed) in ((2000, 60_000, 10, 3, 5), (700, 9001, 5, 2, 6)):   # the second call is smaller
        print(f"[noalloc] n={n} p={p} K={K} d={d}", flush=True)
        g, xs, u, cfg = _case(S, n, p, K, d, seed)
        rc, ref = _run(S, "slfmm_test_associations", g, xs, u, cfg, d)
        assert rc == 0, _err(g)
        rc, got = _run(S, "slfmm_test_associations_noalloc", g, xs, u, cfg, d)
        assert rc == 0, _err(g)
        for k in ("beta", "t", "r2", "p", "gpt", "gif", "hist"):
            assert torch.equal(got[k], ref[k]), k
        assert torch.isfinite(got["t"]).all()
        # a repeated call into outputs allocated beforehand leaves the device's
        # free memory unchanged (the call allocates nothing)
        bufs = _outputs(p, d)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc, again = _run(S, "slfmm_test_associations_noalloc", g, xs, u, cfg, d, bufs=bufs)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert rc == 0, _err(g)
        assert free1 == free0
        for k in ("beta", "t", "r2", "p", "gpt", "gif", "hist"):
            assert torch.equal(again[k], ref[k]), k


def test_noalloc_rejects_what_it_cannot_serve(S):
    from schnellfmm_amd._lib import lib
    g, xs, u, cfg = _case(S, 300, 2000, 4, 2, 7)
    rc, _ = _run(S, "slfmm_test_associations_noalloc", g, xs, u, cfg, 2, flags=0)
    assert rc != 0 and "SLFMM_OUT_DEVICE" in _err(g)
    cfg64 = S.Lfmm2Config(k=4, norm=S.EIGENSTRAT, precision=S.PREC_FP64).to_c()
    rc, _ = _run(S, "slfmm_test_associations_noalloc", g, xs, u, cfg64, 2)
    assert rc != 0 and "SLFMM_PREC_I8" in _err(g)
    # a panel whose scratch would exceed the static array is refused before any work
    n = 400_000
    big = S.BedFile.from_packed(bytes((n + 3) // 4), n, 1)
    xb = np.zeros((n, 2))
    xb[::2, 0] = 1.0
    xb[1::2, 1] = 1.0
    ub = np.ascontiguousarray(np.random.default_rng(1).standard_normal((n, 28)))
    cfgb = S.Lfmm2Config(k=28, norm=S.EIGENSTRAT, precision=S.PREC_I8).to_c()
    rc, _ = _run(S, "slfmm_test_associations_noalloc", big, xb, ub, cfgb, 2)
    assert rc != 0 and "MiB" in _err(big)
    # the context still serves ordinary and allocation-free calls
    rc, t = _run(S, "slfmm_test_associations_noalloc", g, xs, u, cfg, 2)
    assert rc == 0
