"""The MFMA-rate diagnostic runs every operand format and rejects a bad kind."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_mfma_rate_probe():
    import schnellfmm_amd as S
    from schnellfmm_amd._lib import SlfmmError, lib
    ctx = S.Context(0)
    for kind in (0, 1, 2):
        print(f"[diag] mfma rate kind={kind}", flush=True)
        t = C.c_double()
        ctx.check(lib().slfmm_debug_mfma_rate(ctx.h, kind, 2000, C.byref(t)))
        assert 10.0 < t.value < 20000.0, (kind, t.value)   # TOPS: above idle, below any MI355X peak
    with pytest.raises(SlfmmError):
        ctx.check(lib().slfmm_debug_mfma_rate(ctx.h, 3, 10, C.byref(t)))
