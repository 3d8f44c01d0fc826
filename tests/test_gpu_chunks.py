"""Chunked SpGEMM on the GPU: C computed in A tile-row chunks (views of A)
must equal the corresponding tile-rows of the one-shot C, array by array
and bit for bit; the chunk plan must respect its product cap and sum to
nnzCub."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import device_tiles, log

pytestmark = pytest.mark.gpu


def slice_rows(full, r0, r1):
    tp = full["tile_ptr"]
    t0, t1 = int(tp[r0]), int(tp[r1])
    n0, n1 = int(full["tile_nnz"][t0]), int(full["tile_nnz"][t1])
    return {
        "tile_ptr": tp[r0:r1 + 1] - t0,
        "tile_colidx": full["tile_colidx"][t0:t1],
        "tile_rowidx": full["tile_rowidx"][t0:t1] - r0,
        "tile_nnz": full["tile_nnz"][t0:t1 + 1] - n0,
        "csr_ptr": full["csr_ptr"][16 * t0:16 * t1],
        "mask": full["mask"][16 * t0:16 * t1],
        "csr_col": full["csr_col"][n0:n1],
        "val": full["val"][n0:n1],
    }


def check_chunks(ctx, a: ts.Csr, aat: bool, caps, label):
    log(f"[chunks] {label}")
    da, ta = device_tiles(ctx, a)
    tb, dat = ta, None
    if aat:
        dat = ctx.transpose(da)
        tb = ctx.csr2tile(dat)
        ctx.build_csc(tb)
    tc, st = ctx.spgemm(ta, tb)
    full = ctx.download_tiles(tc)
    nnzcub = ctx.nnzcub(da, dat if aat else da)
    for cap in caps:
        bounds, tot = ctx.chunk_plan(ta, tb, cap)
        assert tot == nnzcub
        assert bounds[0] == 0 and bounds[-1] == ta.tilem and all(x < y for x, y in zip(bounds, bounds[1:]))
        log(f"[chunks] cap {cap}: {len(bounds) - 1} chunks")
        for r0, r1 in zip(bounds, bounds[1:]):
            cc, cst = ctx.spgemm_tile_rows(ta, tb, r0, r1)
            got = ctx.download_tiles(cc)
            ref = slice_rows(full, r0, r1)
            for k, v in ref.items():
                assert np.array_equal(got[k], v), (cap, r0, r1, k)
            if r1 - r0 > 1:
                assert cst.nnzC <= cap, (cap, r0, r1)
            ctx.free(cc)
    for o in (tc, ta, da) + ((tb, dat) if aat else ()):
        ctx.free(o)


def test_chunks_er(ctx):
    a = ts.gen_er(5000, 12.0, seed=11)
    check_chunks(ctx, a, False, [10 ** 5, 3 * 10 ** 5, 10 ** 9], "er5000 A^2")
    check_chunks(ctx, a, True, [2 * 10 ** 5], "er5000 AA^T")


def test_chunks_rmat(ctx):
    a = ts.gen_rmat(13, 16, seed=3)
    check_chunks(ctx, a, True, [10 ** 6, 10 ** 7], "rmat13 AA^T")


def test_chunks_stencil(ctx):
    a = ts.gen_stencil27(24, 24, 24, periodic=True, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    check_chunks(ctx, a, False, [200000], "27p24 A^2")


def test_chunk_plan_single_heavy_row(ctx):
    """A cap below one tile-row's products leaves every tile-row alone."""
    a = ts.gen_banded(300, 20, seed=2)
    da, ta = device_tiles(ctx, a)
    bounds, tot = ctx.chunk_plan(ta, ta, 1)
    assert bounds == list(range(ta.tilem + 1))
    ctx.free(ta)
    ctx.free(da)
