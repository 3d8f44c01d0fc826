"""Device conversions vs the oracle: csr2tile (every array bit-exact), the
CSC tile index, CSC-ordered tile data, tile2csr round trip, the stable
transpose and nnzCub."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import log

pytestmark = pytest.mark.gpu


def inputs():
    yield "5pt64", ts.gen_stencil5(64, 64)
    yield "5pt20x7", ts.gen_stencil5(20, 7)          # m = 140, partial tile-row
    yield "27p16", ts.gen_stencil27(16, 16, 16, True)
    yield "27d13", ts.gen_stencil27(13, 11, 9, False)
    yield "er3000", ts.gen_er(3000, 9.0, seed=4)
    yield "banded777", ts.gen_banded(777, 21)
    yield "rmat11", ts.gen_rmat(11, 16, seed=2)
    yield "empty", ts.Csr(40, 33, np.zeros(41, np.int32), np.zeros(0, np.int32), np.zeros(0))
    # empty rows + a dense row + rectangular
    rp = np.zeros(51, np.int32)
    rows = [np.arange(0, 300, 3), np.array([], np.int64), np.arange(299, 300)]
    cols, ptr = [], [0]
    for i in range(50):
        r = rows[i % 3] if i % 7 else np.arange(0, 300)
        cols.append(r)
        ptr.append(ptr[-1] + len(r))
    c = np.concatenate(cols).astype(np.int32)
    yield "ragged50x300", ts.Csr(50, 300, np.array(ptr, np.int32), c, np.arange(len(c), dtype=np.float64) * 0.5)


@pytest.mark.parametrize("name,a", list(inputs()), ids=[n for n, _ in inputs()])
def test_csr2tile_matches_oracle(ctx, name, a):
    log(f"[csr2tile] {name}")
    d = ctx.upload(a)
    t = ctx.csr2tile(d)
    got = ctx.download_tiles(t)
    ref = oracle.csr2tile(a, packed=False)
    for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "csr_ptr", "csr_col", "val", "mask"):
        assert np.array_equal(got[k], ref[k]), k
    refp = oracle.csr2tile(a, packed=True)
    assert np.array_equal(ctx.packed_cols(t), refp["csr_col"])
    # CSC tile index: stable order by tile column, then tile row
    ctx.build_csc(t)
    got = ctx.download_tiles(t)
    order = np.lexsort((ref["tile_rowidx"], ref["tile_colidx"]))
    assert np.array_equal(got["csc_tileid"], order.astype(np.int32))
    assert np.array_equal(got["csc_rowidx"], ref["tile_rowidx"][order])
    assert np.array_equal(got["csc_ptr"], np.concatenate([[0], np.cumsum(np.bincount(ref["tile_colidx"], minlength=ref["tilen"]))]))
    # CSC-ordered data (the B layout of csr2tile_col_major)
    o = ctx.csc_order(t)
    og = ctx.download_tiles(o)
    sizes = np.diff(ref["tile_nnz"])[order]
    assert np.array_equal(og["tile_nnz"], np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32))
    assert np.array_equal(og["mask"].reshape(-1, 16), ref["mask"].reshape(-1, 16)[order] if len(order) else og["mask"].reshape(-1, 16))
    # and back to row-major order
    rm = ctx.from_csc_order(o)
    rg = ctx.download_tiles(rm)
    for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "csr_ptr", "csr_col", "val", "mask"):
        assert np.array_equal(rg[k], ref[k]), k
    assert np.array_equal(rg["csc_tileid"], order.astype(np.int32))
    ctx.free(rm)
    # tile2csr round trip
    back = ctx.download(ctx.tile2csr(t))
    assert np.array_equal(back.rowptr, a.rowptr) and np.array_equal(back.colidx, a.colidx)
    assert np.array_equal(back.val, a.val)
    ctx.free(o)
    ctx.free(t)
    ctx.free(d)


@pytest.mark.parametrize("name,a", list(inputs()), ids=[n for n, _ in inputs()])
def test_transpose_and_nnzcub(ctx, name, a):
    log(f"[transpose] {name}")
    d = ctx.upload(a)
    dt = ctx.transpose(d)
    t = ctx.download(dt)
    tp, tc, tv = oracle.transpose(a)
    assert np.array_equal(t.rowptr, tp) and np.array_equal(t.colidx, tc) and np.array_equal(t.val, tv)
    if a.m == a.n:
        assert ctx.nnzcub(d, d) == oracle.nnzcub(a, a)
    assert ctx.nnzcub(d, dt) == oracle.nnzcub(a, t)
    ctx.free(dt)
    ctx.free(d)
