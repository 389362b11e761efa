"""Output stage after a GPU fit: the TSV written from the fit's outputs, from
host arrays and straight from device memory (SLFMM_OUT_DEVICE), equals the
oracle's restatement of coalesce_output_tsv (src/testing.rs:740-801) byte for
byte, and the Parquet file carries the same values in the reference's
calibrated schema."""
import ctypes as C

import numpy as np
import pytest

from oracle import output_oracle as OO
from schnellfmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import schnellfmm_amd as S
    return S


def test_fit_then_write_tsv(S, tmp_path):
    import oracle as O
    n, p = 300, 5000
    pan = synth.balding_nichols(n, p, n_pops=4, miss_rate=0.02, d=2, n_causal=10, effect=1.0, seed=13)
    prefix = str(tmp_path / "panel")
    O.write_plink(prefix, pan.g)
    g = S.BedFile.open(prefix + ".bed")
    cfg = S.Lfmm2Config(k=4, norm=S.EIGENSTRAT, n_power_iter=2, seed=3, precision=S.PREC_I8)
    res = S.fit_lfmm2(g, 0.2, g, pan.x, cfg)
    recs = OO.read_bim(prefix + ".bim")
    chrom, ids, pos = [r[0] for r in recs], [r[1] for r in recs], [r[3] for r in recs]
    want = OO.results_tsv(chrom, pos, ids, ["env_a", "env_b"], res.pvalue, res.beta, res.tstat, res.r2)
    out = str(tmp_path / "res.tsv")
    S.write_results(out, res, prefix + ".bed", cov_names=["env_a", "env_b"])
    assert open(out).read() == want

    # the same table straight from device copies of the outputs
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    ptrs = []
    try:
        for a in (res.pvalue, res.beta, res.tstat, res.r2):
            a = np.ascontiguousarray(a)
            ptr = C.c_void_p()
            assert hip.hipMalloc(C.byref(ptr), C.c_size_t(a.nbytes)) == 0
            ptrs.append(ptr)
            assert hip.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # H2D
        cn = (C.c_char_p * 2)(b"env_a", b"env_b")
        out2 = str(tmp_path / "res_dev.tsv")
        from schnellfmm_amd._lib import lib
        rc = lib().slfmm_write_tsv(g.ctx.h, out2.encode(), (prefix + ".bim").encode(), None, None, None, p, 2, cn,
                                   ptrs[0], ptrs[1], ptrs[2], ptrs[3], 1, 4)
        assert rc == 0, lib().slfmm_last_error(g.ctx.h)
        assert open(out2).read() == want
    finally:
        for ptr in ptrs:
            hip.hipFree(ptr)

    pq = pytest.importorskip("pyarrow.parquet")
    outp = str(tmp_path / "res.parquet")
    S.write_results(outp, res, prefix + ".bed", cov_names=["env_a", "env_b"], fmt="parquet", n_workers=4,
                    chunk_size=1000)
    tab = pq.read_table(outp)
    order = OO.parquet_row_order(p, 1000, 4)
    assert np.array_equal(tab.column("t_env_b").to_numpy(), res.tstat[order, 1])
    assert np.array_equal(tab.column("p_env_a").to_numpy(), res.pvalue[order, 0])
