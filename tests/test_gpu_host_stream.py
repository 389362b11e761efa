"""Host-resident (out-of-core) panels: the rows stay in host memory (a mapped
.bed or a caller buffer) and cross PCIe chunk by chunk while the previous
chunk is being tested (the reference's mmap + IO-thread chunk pipeline,
src/bed.rs:78-135, src/parallel.rs:156-329).  The kernels are the ones the
device-resident panel runs, on the same rows, so every output must equal the
device-resident panel's bit for bit; the device-resident path is pinned
against the oracle by the parity tests.  Chunk sizes are chosen to leave
ragged last chunks and chunk edges inside the kernels' 512-SNP tiles, and the
padding codes of the host rows carry garbage that the upload must clear."""
import numpy as np
import pytest

from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def _panel(n, p, seed, d=2, miss=0.03):
    import oracle as O
    pan = synth.balding_nichols(n, p, n_pops=5, miss_rate=miss, d=d, n_causal=10, effect=1.0, seed=seed)
    packed = np.frombuffer(O.encode_bed(pan.g), np.uint8).copy()
    bps = (n + 3) // 4
    if n % 4:   # garbage in the codes past the last sample (the upload clears them)
        rows = packed.reshape(p, bps)
        junk = np.random.default_rng(seed).integers(0, 256, p, dtype=np.uint8)
        keep = np.uint8((1 << (2 * (n % 4))) - 1)
        rows[:, -1] = (rows[:, -1] & keep) | (junk & ~keep)
    return pan, packed


def _same(a, b):
    for name in ("u_hat", "beta", "tstat", "r2", "pvalue", "gif_per_trait"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.shape == y.shape, name
        assert np.array_equal(x, y), name
    assert a.gif == b.gif


@pytest.mark.parametrize("prec", ["i8", "fp64"])
@pytest.mark.parametrize("chunk", [1000, 777, 0])
def test_host_panel_fit_equals_device(S, prec, chunk):
    n, p = 517, 7013
    pan, packed = _panel(n, p, seed=11)
    dev = S.BedFile.from_packed(packed.tobytes(), n, p)
    host = S.BedFile.wrap_host(packed, n, p, chunk_snps=chunk)
    cfg = S.Lfmm2Config(k=6, norm=S.EIGENSTRAT, n_power_iter=2, seed=5,
                        precision=S.PREC_I8 if prec == "i8" else S.PREC_FP64)
    ref = S.fit_lfmm2(dev, 0.3, dev, pan.x, cfg)
    got = S.fit_lfmm2(host, 0.3, host, pan.x, cfg)
    _same(got, ref)
    # mixed: device estimation panel, host test panel
    got2 = S.fit_lfmm2(dev, 0.3, host, pan.x, cfg)
    _same(got2, ref)
    assert np.max(np.abs(ref.tstat)) > 4   # the panel has real signal


def test_host_bed_file_decode_and_fit(S, tmp_path):
    import oracle as O
    n, p = 203, 2600
    pan = synth.balding_nichols(n, p, n_pops=3, miss_rate=0.05, d=1, n_causal=5, effect=1.0, seed=21)
    prefix = str(tmp_path / "h")
    O.write_plink(prefix, pan.g)
    dev = S.BedFile.open(prefix + ".bed")
    host = S.BedFile.open_host(prefix + ".bed", chunk_snps=600)
    assert (host.n_samples, host.n_snps) == (n, p)
    idx = np.random.default_rng(1).integers(0, p, 97)
    for norm in (S.CENTER_ONLY, S.EIGENSTRAT):
        assert np.array_equal(host.decode(idx, norm=norm), dev.decode(idx, norm=norm))
    assert np.array_equal(np.nan_to_num(host.decode_raw(idx), nan=-1), np.nan_to_num(dev.decode_raw(idx), nan=-1))
    for a, b in zip(host.stats(idx), dev.stats(idx)):
        assert np.array_equal(a, b)
    cfg = S.Lfmm2Config(k=3, norm=S.CENTER_ONLY, n_power_iter=1, seed=9, precision=S.PREC_I8)
    _same(S.fit_lfmm2(host, None, host, pan.x, cfg), S.fit_lfmm2(dev, None, dev, pan.x, cfg))
    host.free()
    dev.free()


def test_host_panel_nmf(S):
    n, p = 301, 1800
    pan, packed = _panel(n, p, seed=31, d=1, miss=0.05)
    dev = S.BedFile.from_packed(packed.tobytes(), n, p)
    host = S.BedFile.wrap_host(packed, n, p, chunk_snps=500)
    nmf = S.NmfParams(k=3, n_iter=2, cv_rate=0.01, seed=4, chunk_size=500)
    a = S.nmf_impute_estimation(host, 0.5, nmf)
    b = S.nmf_impute_estimation(dev, 0.5, nmf)
    assert np.array_equal(a.w, b.w) and np.array_equal(a.h_full, b.h_full)
    assert np.array_equal(a.cv_mae_per_iter, b.cv_mae_per_iter)
    cfg = S.Lfmm2Config(k=3, norm=S.EIGENSTRAT, n_power_iter=1, seed=2)
    # estimation rows may be host-resident; the NMF-imputed test pass needs the panel on the device
    _same(S.fit_lfmm2_nmf(host, 0.5, dev, pan.x, cfg, nmf), S.fit_lfmm2_nmf(dev, 0.5, dev, pan.x, cfg, nmf))
    with pytest.raises(S.api.SlfmmError, match="device-resident"):
        S.fit_lfmm2_nmf(dev, 0.5, host, pan.x, cfg, nmf)


def test_host_panel_batch(S):
    n, p = 400, 3100
    pan, packed = _panel(n, p, seed=41)
    dev = S.BedFile.from_packed(packed.tobytes(), n, p)
    host = S.BedFile.wrap_host(packed, n, p, chunk_snps=1024)
    cfg = S.Lfmm2Config(k=4, norm=S.EIGENSTRAT, n_power_iter=1, seed=3, precision=S.PREC_I8)
    probs = [S.Lfmm2Problem(host, 0.5, host, pan.x, cfg), S.Lfmm2Problem(dev, 0.5, dev, pan.x, cfg)]
    got = S.fit_lfmm2_batch(probs, n_streams=2)
    _same(got[0], got[1])


def test_read_packed_round_trip(S):
    import oracle as O
    n, p = 517, 3001
    pan, packed = _panel(n, p, seed=51)
    bps = (n + 3) // 4
    clean = np.frombuffer(O.encode_bed(pan.g), np.uint8).reshape(p, bps)
    dev = S.BedFile.from_packed(packed.tobytes(), n, p)
    host = S.BedFile.wrap_host(packed, n, p)
    assert np.array_equal(dev.read_packed(), clean)
    assert np.array_equal(host.read_packed(), clean)
    assert np.array_equal(dev.read_packed(1000, 17), clean[1000:1017])
    assert dev.read_packed(p, 0).shape == (0, bps)
    with pytest.raises(S.api.SlfmmError, match="out of range"):
        dev.read_packed(p - 5, 6)


def test_pinned_host_panel(S):
    # page-locked rows (hipHostMalloc from the HIP runtime the library loaded)
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    n, p = 333, 5000
    pan, packed = _panel(n, p, seed=61)
    ptr = C.c_void_p()
    assert hip.hipHostMalloc(C.byref(ptr), C.c_size_t(packed.size), C.c_uint(0)) == 0
    try:
        pinned = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), shape=(packed.size,))
        pinned[:] = packed
        dev = S.BedFile.from_packed(packed.tobytes(), n, p)
        host = S.BedFile.wrap_host(pinned, n, p, chunk_snps=1200)
        cfg = S.Lfmm2Config(k=5, norm=S.EIGENSTRAT, n_power_iter=2, seed=8, precision=S.PREC_I8)
        _same(S.fit_lfmm2(host, 0.5, host, pan.x, cfg), S.fit_lfmm2(dev, 0.5, dev, pan.x, cfg))
        host.free()
    finally:
        hip.hipHostFree(ptr)


def test_host_panel_fit_vs_oracle(S):
    """A host-streamed FP64 fit against the oracle directly (same injected
    sketch): U_hat angles < 1e-8, t / beta / r2 within 1e-8 relative."""
    import oracle as O
    n, p, K = 517, 7013, 6
    pan, packed = _panel(n, p, seed=23)
    clean = O.encode_bed(pan.g)          # the oracle reads the rows without padding garbage
    xs = O.center_covariates(pan.x)
    pre = O.precompute(xs, 1e-5)
    b0 = pre.m.T @ np.random.default_rng(9).standard_normal((n, K + 10))
    est = O.subset_indices(0.3, p)
    cfg_o = O.Config(k=K, norm=O.EIGENSTRAT, chunk_size=1000, n_power_iter=2, seed=5)
    u_ref, ref, _, _ = O.fit_lfmm2(clean, O.bytes_per_snp(n), n, p, pan.x, cfg_o, est, b0=b0)
    host = S.BedFile.wrap_host(packed, n, p, chunk_snps=777)
    cfg = S.Lfmm2Config(k=K, norm=S.EIGENSTRAT, n_power_iter=2, seed=5, precision=S.PREC_FP64)
    got = S.fit_lfmm2(host, 0.3, host, pan.x, cfg, b0=b0)
    assert np.max(O.principal_angles(got.u_hat, u_ref)) < 1e-8
    assert np.allclose(got.tstat, ref.t, rtol=1e-8, atol=1e-10)
    assert np.allclose(got.beta, ref.beta, rtol=1e-8, atol=1e-12)
    assert np.allclose(got.r2, ref.r2, rtol=1e-8, atol=1e-12)
