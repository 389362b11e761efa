"""Output stage on the CPU: the oracle's `{:.6e}` formatter pinned against
Rust's documented form and exact rational rounding; the library's formatter
(slfmm_format_e6) and TSV writer (slfmm_write_tsv, host arrays, no context)
against the oracle byte for byte; the Parquet writer's schema and row order
against the reference's (src/testing.rs:621-640, :861-909)."""
import ctypes as C

import numpy as np
import pytest

from oracle import output_oracle as OO


def _lib():
    from schnellfmm_amd._lib import lib
    return lib()


def _fmt_lib(xs):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    out = C.create_string_buffer(32 * len(xs) + 1)
    assert _lib().slfmm_format_e6(xs.ctypes.data, len(xs), out, 32) == 0
    raw = out.raw
    return [raw[32 * i:32 * i + 32].split(b"\0")[0].decode() for i in range(len(xs))]


def _tie_values(rng, count):
    """Doubles whose 8th significant digit onward is exactly 5 (ties), their
    neighbours, and values next to powers of ten."""
    vals = []
    for _ in range(count):
        q = int(rng.integers(10 ** 6, 10 ** 7))
        for k in range(0, 10):                       # q.5 * 2^-k stays exact for small k
            t = (q + 0.5) * 2.0 ** -k
            vals += [t, np.nextafter(t, 0), np.nextafter(t, np.inf)]
        e = int(rng.integers(-300, 300))
        b = 10.0 ** e
        vals += [b, np.nextafter(b, 0), np.nextafter(b, np.inf), 9.9999995 * b, 9.9999994999 * b]
    return np.array(vals)


# ---- oracle pinning ----------------------------------------------------------

def test_rust_e6_documented_form():
    # Rust's `{:e}` family: mantissa digits, 'e', exponent without '+' or leading zeros;
    # NaN prints "NaN", infinities "inf", and negative zero keeps its sign.
    cases = [(1.0, "1.000000e0"), (1234.5, "1.234500e3"), (1e-7, "1.000000e-7"), (0.0, "0.000000e0"),
             (-0.0, "-0.000000e0"), (-2.5, "-2.500000e0"), (1e300, "1.000000e300"), (123456789.0, "1.234568e8"),
             (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"), (0.000123, "1.230000e-4")]
    for x, want in cases:
        assert OO.rust_e6(x) == want, x
        assert OO.e6_exact(x) == want, x


def test_rust_e6_matches_exact_rounding():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2 ** 63, 4000, dtype=np.int64).view(np.float64)
    bits = bits[np.isfinite(bits)]
    vals = np.concatenate([bits, rng.standard_normal(2000) * 10.0 ** rng.integers(-20, 20, 2000),
                           _tie_values(rng, 60)])
    for x in vals:
        assert OO.rust_e6(x) == OO.e6_exact(x), repr(x)
    # ties to even on exactly representable ties
    assert OO.e6_exact(1234567.5) == "1.234568e6" and OO.e6_exact(1234568.5) == "1.234568e6"
    assert OO.e6_exact(12345665.0) == "1.234566e7"


def test_parquet_row_order_brute_force():
    for p, cs, w in [(25, 4, 3), (24, 4, 3), (10, 3, 0), (10, 3, 1), (7, 10, 4), (0, 5, 2), (40, 5, 16)]:
        n_chunks = -(-p // cs) if p else 0
        W = max(min(w, n_chunks), 1)
        files = [[] for _ in range(W)]
        for seq in range(n_chunks):                  # worker seq % W gets chunk seq, in order
            files[seq % W].extend(range(seq * cs, min(p, seq * cs + cs)))
        want = [j for f in files for j in f]
        assert OO.parquet_row_order(p, cs, w).tolist() == want


# ---- library against the oracle -------------------------------------------------

def test_format_e6_library_matches_oracle():
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2 ** 64, 200_000, dtype=np.uint64).view(np.float64)
    vals = np.concatenate([
        bits,
        rng.standard_normal(100_000),
        rng.uniform(0, 1, 100_000) ** 8,                           # p-value-like
        rng.standard_normal(50_000) * 10.0 ** rng.integers(-320, 300, 50_000),
        _tie_values(rng, 300),
        np.array([0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 5e-324, -5e-324, 2.2250738585072014e-308,
                  1.7976931348623157e308]),
    ])
    got = _fmt_lib(vals)
    for x, g in zip(vals, got):
        assert g == OO.rust_e6(x), repr(x)


def _random_results(rng, p, d):
    t = rng.standard_normal((p, d)) * 3
    t[rng.integers(0, p, 5), 0] = 0.0
    pv = rng.uniform(0, 1, (p, d)) ** 4
    be = rng.standard_normal((p, d)) * 1e-2
    r2 = rng.dirichlet([1, 1, 1], p)
    r2[:3] = 0.0                                     # monomorphic SNPs leave r2 at zero
    return pv, be, t, r2


def _write_bim(path, chrom, ids, pos):
    with open(path, "w") as f:
        for c, i, ps in zip(chrom, ids, pos):
            f.write(f"{c}\t{i}\t0.{len(i)}\t{ps}\tA\tG\n")


@pytest.mark.parametrize("threads", [1, 5])
def test_write_tsv_matches_oracle(tmp_path, threads):
    rng = np.random.default_rng(3)
    p, d = 70_001, 3                                  # three row blocks, ragged last
    pv, be, t, r2 = _random_results(rng, p, d)
    chrom = [str(1 + j // 30_000) for j in range(p)]
    ids = [f"rs{j * 7 + 1}" for j in range(p)]
    pos = (np.arange(p, dtype=np.uint64) * 13 + 5)
    bim = str(tmp_path / "x.bim")
    _write_bim(bim, chrom, ids, pos)
    names = ["temp", "prec", "alt"]
    want = OO.results_tsv(chrom, pos, ids, names, pv, be, t, r2)
    cn = (C.c_char_p * d)(*[s.encode() for s in names])
    out1 = str(tmp_path / "a.tsv")
    rc = _lib().slfmm_write_tsv(None, out1.encode(), bim.encode(), None, None, None, p, d, cn, pv.ctypes.data,
                                be.ctypes.data, t.ctypes.data, r2.ctypes.data, 0, threads)
    assert rc == 0
    assert open(out1).read() == want
    # metadata from arrays instead of the .bim
    out2 = str(tmp_path / "b.tsv")
    ch = (C.c_char_p * p)(*[s.encode() for s in chrom])
    idp = (C.c_char_p * p)(*[s.encode() for s in ids])
    rc = _lib().slfmm_write_tsv(None, out2.encode(), None, ch, pos.ctypes.data, idp, p, d, cn, pv.ctypes.data,
                                be.ctypes.data, t.ctypes.data, r2.ctypes.data, 0, threads)
    assert rc == 0
    assert open(out2).read() == want
    # the reference's reader sees the same numbers to 7 significant digits
    header, rp, rb, rt, rr = OO.parse_results_tsv(want)
    assert header[:3] == ["chr", "pos", "snp_id"] and header[3:6] == ["p_temp", "beta_temp", "t_temp"]
    assert rp.shape == (p, d)
    assert np.allclose(rt, t, rtol=5e-7, atol=0) and np.allclose(rr, r2, rtol=5e-7, atol=0)


def test_write_results_api_and_errors(tmp_path):
    import schnellfmm_amd as S
    rng = np.random.default_rng(5)
    p, d = 1000, 2
    pv, be, t, r2 = _random_results(rng, p, d)
    res = S.TestResults(u_hat=np.zeros((1, 1)), gif=1.0, gif_per_trait=np.ones(d), beta=be, tstat=t, r2=r2,
                        pvalue=pv)
    chrom, ids, pos = ["2"] * p, [f"s{j}" for j in range(p)], np.arange(p, dtype=np.uint64) + 100
    bed = str(tmp_path / "panel.bed")
    _write_bim(str(tmp_path / "panel.bim"), chrom, ids, pos)
    S.write_results(str(tmp_path / "r.tsv"), res, bed)
    assert open(tmp_path / "r.tsv").read() == OO.results_tsv(chrom, pos, ids, None, pv, be, t, r2)
    with pytest.raises(ValueError, match="unknown output format 'csv'"):
        S.write_results(str(tmp_path / "r.csv"), res, bed, fmt="csv")
    # a .bim that does not match, and one with a bad position
    with open(tmp_path / "short.bim", "w") as f:
        f.write("1\ta\t0\t5\tA\tG\n")
    with pytest.raises(S.api.SlfmmError, match="has 1 records, expected 1000"):
        S.write_results(str(tmp_path / "r2.tsv"), res, str(tmp_path / "short.bim"))
    with open(tmp_path / "bad.bim", "w") as f:
        f.write("1\ta\t0\t5\tA\tG\n1\tb\t0\tx7\tA\tG\n")
    with pytest.raises(S.api.SlfmmError, match="BIM line 2: invalid position 'x7'"):
        S.write_results(str(tmp_path / "r3.tsv"), res, str(tmp_path / "bad.bim"))
    with pytest.raises(ValueError, match="BIM line 2: invalid position 'x7'"):
        OO.read_bim(str(tmp_path / "bad.bim"))
    with open(tmp_path / "few.bim", "w") as f:
        f.write("1\ta\t0\t5\tA\n")
    with pytest.raises(S.api.SlfmmError, match="BIM line 1 has 5 fields, expected 6"):
        S.write_results(str(tmp_path / "r4.tsv"), res, str(tmp_path / "few.bim"))


@pytest.mark.parametrize("n_workers,chunk", [(0, 10_000), (3, 128), (16, 100)])
def test_write_parquet_schema_and_order(tmp_path, n_workers, chunk):
    pq = pytest.importorskip("pyarrow.parquet")
    import schnellfmm_amd as S
    rng = np.random.default_rng(9)
    p, d = 1234, 2
    pv, be, t, r2 = _random_results(rng, p, d)
    res = S.TestResults(u_hat=np.zeros((1, 1)), gif=1.0, gif_per_trait=np.ones(d), beta=be, tstat=t, r2=r2,
                        pvalue=pv)
    chrom, ids, pos = [str(j % 3 + 1) for j in range(p)], [f"s{j}" for j in range(p)], np.arange(p) * 3 + 1
    bim = str(tmp_path / "x.bim")
    _write_bim(bim, chrom, ids, pos)
    out = str(tmp_path / "r.parquet")
    S.write_results(out, res, bim, cov_names=["env_0", "env_1"], fmt="Parquet", n_workers=n_workers,
                    chunk_size=chunk)
    tab = pq.read_table(out)
    want_cols = OO.calibrated_columns(["env_0", "env_1"])
    assert [f.name for f in tab.schema] == [c for c, _ in want_cols]
    assert [str(f.type) for f in tab.schema] == [ty for _, ty in want_cols]
    assert pq.ParquetFile(out).metadata.row_group(0).column(0).compression == "ZSTD"
    order = OO.parquet_row_order(p, chunk, n_workers)
    assert tab.column("snp_id").to_pylist() == [ids[j] for j in order]
    assert tab.column("pos").to_pylist() == [int(pos[j]) for j in order]
    for j, nm in enumerate(["env_0", "env_1"]):
        assert np.array_equal(tab.column(f"p_{nm}").to_numpy(), pv[order, j])
        assert np.array_equal(tab.column(f"beta_{nm}").to_numpy(), be[order, j])
        assert np.array_equal(tab.column(f"t_{nm}").to_numpy(), t[order, j])
    assert np.array_equal(tab.column("r2_resid").to_numpy(), r2[order, 2])
