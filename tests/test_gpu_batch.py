"""Batched fits (slfmm_fit_lfmm2_batch): many independent LFMM2 problems of
uneven shape in one call, spread over several streams.  Each problem's
outputs must equal, bit for bit, what slfmm_fit_lfmm2 returns for the same
inputs (same kernels, same device random streams), in both precisions and
with ragged SNP tails; the single-problem path is pinned against the oracle
by the parity tests.  Errors name the failing problem."""
import numpy as np
import pytest

from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


# (n, p, K, d, n_pops, miss, est_rate, precision)
SHAPES = [
    (200, 2000, 3, 1, 2, 0.00, None, "fp64"),
    (517, 7013, 12, 3, 5, 0.03, 0.5, "i8"),
    (1135, 4099, 8, 8, 8, 0.05, 0.25, "i8"),
    (333, 20001, 20, 2, 10, 0.01, 0.1, "fp64"),
    (1000, 3000, 5, 1, 3, 0.02, None, "i8"),
    (257, 1500, 4, 4, 2, 0.05, None, "fp64"),
]


def _problems(S):
    probs = []
    for i, (n, p, K, d, pops, miss, rate, prec) in enumerate(SHAPES):
        pan = synth.balding_nichols(n, p, n_pops=pops, miss_rate=miss, d=d, n_causal=10, effect=1.0, seed=100 + i)
        import oracle as O
        g = S.BedFile.from_packed(O.encode_bed(pan.g), n, p)
        cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, n_power_iter=2, seed=7 + i,
                            precision=S.PREC_I8 if prec == "i8" else S.PREC_FP64)
        probs.append(S.Lfmm2Problem(g, rate, g, pan.x, cfg))
    return probs


@pytest.mark.parametrize("n_streams", [1, 3, 8, 16])
def test_batch_equals_single_fits(S, n_streams):
    probs = _problems(S)
    got = S.fit_lfmm2_batch(probs, n_streams=n_streams)
    assert len(got) == len(probs)
    for pr, r in zip(probs, got):
        ref = S.fit_lfmm2(pr.y_est, pr.est_subset, pr.y_full, pr.x, pr.config)
        for name in ("u_hat", "beta", "tstat", "r2", "pvalue", "gif_per_trait"):
            a, b = getattr(r, name), getattr(ref, name)
            assert a.shape == b.shape
            assert np.array_equal(a, b), name
        assert r.gif == ref.gif
        assert np.all(np.isfinite(r.tstat))


def test_batch_reports_failing_problem(S):
    probs = _problems(S)[:3]
    probs[2].config = S.Lfmm2Config(k=0)
    with pytest.raises(S.api.SlfmmError, match="problem 2"):
        S.fit_lfmm2_batch(probs, n_streams=2)
    # the context stays usable
    ok = S.fit_lfmm2_batch(probs[:2], n_streams=2)
    assert len(ok) == 2


def test_batch_rejects_bad_stream_count(S):
    probs = _problems(S)[:1]
    for bad in (0, 17):
        with pytest.raises(S.api.SlfmmError, match="n_streams"):
            S.fit_lfmm2_batch(probs, n_streams=bad)
