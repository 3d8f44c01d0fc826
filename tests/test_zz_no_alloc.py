"""tsg_spgemm_count_tiles_noalloc (scratch in a fixed per-GPU device
array, no allocation) against tsg_spgemm_count_tiles on the same operands,
two calls in a row, the second on a smaller input, and against the
oracle's step-1 tile count; oversized inputs are refused before any
device work."""
import ctypes as C

import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import device_tiles, log

pytestmark = pytest.mark.gpu


def noalloc_count(ctx, ta, tb, d_out):
    ctx.count_tiles_noalloc(ta, tb, d_out)
    ctx.sync()
    v = np.zeros(1, np.int64)
    ts._check(ts.lib.tsg_d2h(ctx.h, v.ctypes.data, d_out, 8))
    return int(v[0])


def test_count_tiles_noalloc_matches_allocating_entry_point(ctx):
    d_out = C.c_void_p()
    ts._check(ts.lib.tsg_dmalloc(ctx.h, C.byref(d_out), 8))
    cases = [
        ("rmat 14 A*A^T", ts.gen_rmat(14, 16, seed=5), True),   # larger first
        ("27p 40^3 A^2", ts.gen_stencil27(40, 40, 40, periodic=True, valmode=ts.VAL_UNIFORM_PM1, seed=1), False),
        ("er 3000 A^2", ts.gen_er(3000, 9.0, seed=2), False),   # then smaller
        ("5pt 9x7 A^2", ts.gen_stencil5(9, 7), False),
    ]
    for label, a, aat in cases:
        da, ta = device_tiles(ctx, a)
        tb, dat, b = ta, None, a
        if aat:
            dat = ctx.transpose(da)
            tb = ctx.csr2tile(dat)
            ctx.build_csc(tb)
            tp, tc, tv = oracle.transpose(a)
            b = ts.Csr(a.n, a.m, tp, tc, tv)
        ref = ctx.count_tiles(ta, tb)
        for rep in range(2):
            got = noalloc_count(ctx, ta, tb, d_out)
            assert got == ref, (label, rep, got, ref)
        cptr, _ = oracle.step1_tiles(oracle.csr2tile(a), oracle.csr2tile(b))
        assert ref == int(cptr[-1]), label
        log(f"[noalloc] {label}: {got} C tiles (tile-rows {ta.tilem})")
        for o in ((tb, dat) if dat is not None else ()) + (ta, da):
            ctx.free(o)
    ctx.dfree(d_out)


def test_count_tiles_noalloc_rejects_oversized(ctx):
    """A tile matrix with more tile-rows than the scratch holds is refused
    on the host (its arrays are never read)."""
    big = ts.DTiles()
    big.m = big.n = 16 * (ts.NOALLOC_MAX_TILE_ROWS + 1)
    big.tilem = big.tilen = ts.NOALLOC_MAX_TILE_ROWS + 1
    d_out = C.c_void_p()
    ts._check(ts.lib.tsg_dmalloc(ctx.h, C.byref(d_out), 8))
    with pytest.raises(ts.TsgError, match="scratch holds at most"):
        ctx.count_tiles_noalloc(big, big, d_out)
    ctx.dfree(d_out)
