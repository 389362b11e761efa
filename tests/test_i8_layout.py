"""CPU pin of the packed-operand row layout of the int8 SNP-side kernels
(slfmm_debug_i8_layout, host only).  The kernels use
v_mfma_i32_32x32x32_i8, whose result row 32 rt + 8 b + 4 h + v lands in
register 4 b + v of lane half h; they recombine functional i = h + 2 q from
accumulator slots fixed at compile time.  Every (column, term) row must
exist exactly once, in lane half i % 2, at those slots, and the all-ones
row exactly once."""
import ctypes as C

import numpy as np
import pytest


def layout(m, k3, km):
    from schnellfmm_amd._lib import lib
    dims = np.zeros(5, np.int32)
    rows = np.zeros(2 * 32 * 16, np.int32)
    rc = lib().slfmm_debug_i8_layout(m, k3, km, dims.ctypes.data_as(C.c_void_p), rows.ctypes.data_as(C.c_void_p),
                                     rows.size)
    assert rc == 0
    nq, nt0, nt1, g1, ones_g = (int(v) for v in dims)
    return nq, nt0, nt1, g1, ones_g, rows[: 2 * 32 * (nt0 + nt1)].reshape(-1, 2)


def _cases():
    for m in range(1, 33):
        for d in range(1, 17):
            if m - d >= 1:
                yield m, m - d, m - d      # association: 1 miss term for the latent columns
                yield m, m - d, 0          # 2 miss terms for every column
        yield m, m, m                      # RSVD sketch: 2 + 1 terms
        yield m, max(m - 4, 0), max(m - 4, 0)


def test_every_term_has_its_slot():
    for m, k3, km in _cases():
        nq, nt0, nt1, g1, ones_g, rows = layout(m, k3, km)
        assert nq == (m + 1) // 2
        seen = {}
        ones = []
        for row, (col, term) in enumerate(rows):
            rt, rr = row // 32, row % 32
            g, r = (rr // 4) % 2, 4 * (rr // 8) + rr % 4
            side = 0 if rt < nt0 else 1
            slot = 16 * (rt - (0 if side == 0 else nt0)) + r
            if col == -2:
                ones.append((side, g, slot))
                continue
            if col < 0:
                continue
            key = (col, side, term)
            assert key not in seen, (m, k3, km, key)
            seen[key] = (g, slot)
            q = col // 2
            assert g == col % 2
            if side == 0:
                assert slot == (2 * q + term if term < 2 else 3 * nq - 1 - q), (m, k3, km, key)
            else:
                assert slot == (q if term == 0 else 2 * nq - 1 - q), (m, k3, km, key)
        for i in range(m):
            for t in range(3 if i >= k3 else 2):
                assert (i, 0, t) in seen, (m, k3, km, i, t)
            for t in range(2 if i >= km else 1):
                assert (i, 1, t) in seen, (m, k3, km, i, t)
        assert len(seen) == sum((3 if i >= k3 else 2) + (2 if i >= km else 1) for i in range(m))
        assert ones == [(ones_g == 0 and 1 or 0, g1, 16 * (nt0 if ones_g else nt1) - 1)], (m, k3, km, ones)


@pytest.mark.parametrize("m,k3,km,tiles", [(24, 20, 0, 4), (24, 20, 20, 3), (30, 30, 30, 3), (15, 15, 15, 2),
                                           (22, 22, 22, 3), (13, 12, 0, 2), (32, 16, 0, 5)])
def test_tile_counts_of_the_workloads(m, k3, km, tiles):
    nq, nt0, nt1, *_ = layout(m, k3, km)
    assert nt0 + nt1 == tiles


def test_bad_shapes_rejected():
    from schnellfmm_amd._lib import lib
    dims = np.zeros(5, np.int32)
    for m, k3, km in ((0, 0, 0), (33, 30, 0), (8, 9, 0), (8, 4, 9)):
        assert lib().slfmm_debug_i8_layout(m, k3, km, dims.ctypes.data_as(C.c_void_p), None, 0) != 0


def test_every_layout_has_a_kernel_instantiation():
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "schnellfmm_amd", "csrc", "assoc_i8.hip")).read()

    def shapes(macro):
        body = re.search(r"#define " + macro + r"\(X\)(.*?)\n(?!\s)", src, re.S).group(1)
        return {tuple(int(v) for v in t) for t in re.findall(r"X\((\d+), (\d+), (\d+)\)", body)}

    assoc, rsvd = shapes("SLFMM_SNP_ASSOC"), shapes("SLFMM_SNP_RSVD")
    need_a = {layout(m, m - d, m - d)[:3] for m in range(2, 33) for d in range(1, 17) if m - d >= 1}
    need_r = {layout(l, l, l)[:3] for l in range(1, 33)}
    assert need_a == assoc
    assert need_r == rsvd
