"""Launch shapes on the GPU: the tuning table reaches the library, the tile
choice of the int8 SNP-side kernels leaves every output bit-identical, and
the workgroup targets change results at rounding level only."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def test_set_get_round_trip(S):
    from schnellfmm_amd import tuning
    ctx = S.Context(0)
    p = dict(tuning.DEFAULTS, snp_two_tiles_min_blocks=123, fwd_target_wg=777, gram_max_blocks=99, batch_fit_ms=0.5,
             batch_gb_per_ms=3.0)
    tuning.apply(ctx, p)
    got = tuning.read(ctx)
    for k in tuning.LIBRARY_KEYS:
        assert got[k] == p[k], k
    bad = dict(p, fwd_target_wg=0)
    with pytest.raises(ValueError):
        tuning.apply(ctx, bad)


@pytest.mark.parametrize("K,d", [(5, 2), (20, 4), (13, 1)])
def test_tile_choice_is_bit_identical(S, K, d):
    from schnellfmm_amd import tuning
    n, p = 1100, 70_001   # ragged last workgroup for both tile counts
    print(f"[tuning] tiles n={n} p={p} K={K} d={d}", flush=True)
    ctx = S.Context(0)
    g = S.BedFile.synthetic(n, p, n_pops=5, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.03, seed=K, ctx=ctx)
    rng = np.random.default_rng(K + d)
    x = rng.standard_normal((n, d))
    cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, precision=S.PREC_I8, n_power_iter=2)
    outs = []
    for thr in (2 ** 32 - 1, 0):   # one tile per wave everywhere, then two
        tuning.apply(ctx, dict(tuning.DEFAULTS, snp_two_tiles_min_blocks=thr))
        outs.append(S.fit_lfmm2(g, 0.05, g, x, cfg))
    a, b = outs
    for f in ("u_hat", "beta", "tstat", "r2", "pvalue", "gif_per_trait"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_workgroup_targets_change_rounding_only(S):
    from schnellfmm_amd import tuning
    n, p = 3000, 40_000
    print(f"[tuning] workgroup targets n={n} p={p}", flush=True)
    ctx = S.Context(0)
    g = S.BedFile.synthetic(n, p, n_pops=6, fst=(0.02, 0.15), maf=(0.05, 0.5), miss_rate=0.02, seed=11, ctx=ctx)
    x = np.random.default_rng(2).standard_normal((n, 3))
    outs = []
    for prec in (S.PREC_I8, S.PREC_FP64):
        cfg = S.Lfmm2Config(k=10, norm=S.EIGENSTRAT, precision=prec, n_power_iter=2)
        res = []
        for fwd, gram in ((512, 1024), (256, 64), (2048, 4096)):
            tuning.apply(ctx, dict(tuning.DEFAULTS, fwd_target_wg=fwd, gram_max_blocks=gram))
            res.append(S.fit_lfmm2(g, "all", g, x, cfg))
        for r in res[1:]:
            assert float(np.max(O.principal_angles(r.u_hat, res[0].u_hat))) < 1e-9
            assert np.allclose(r.tstat, res[0].tstat, rtol=1e-8, atol=1e-10)
        outs.append(res)
