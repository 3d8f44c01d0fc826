"""R-MAT C = A*A^T under device orderings (random, RCM, degree), as in the
R-MAT scale-20 configuration, at sizes the oracle finishes in seconds:
every array of C bit-exact against the oracle on the reordered matrix;
nnz(C) invariant under the orderings; C bit-symmetric (C(i,j) and C(j,i)
sum the same products in the same ascending-k order); chunked products
equal to the oracle's rows."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import device_tiles, log, spgemm_check

pytestmark = pytest.mark.gpu


def t(csr):
    tp, tc, tv = oracle.transpose(csr)
    return ts.Csr(csr.n, csr.m, tp, tc, tv)


def reorder(ctx, a, algo):
    d = ctx.upload(a)
    if algo == "rcm":
        old, _ = ctx.order_rcm(d)
    elif algo == "degree":
        old = ctx.order_degree(d)
    else:
        old = ctx.order_random(a.m, 1)
    new = np.empty_like(old)
    new[old] = np.arange(a.m, dtype=np.int32)
    dp = ctx.permute(d, old, new)
    ap = ctx.download(dp)
    ctx.free(dp)
    ctx.free(d)
    return ap


def test_rmat13_aat_orderings(ctx):
    a = ts.gen_rmat(13, 16, seed=1)
    nnzc = set()
    for algo in ("random", "rcm", "degree"):
        ap = reorder(ctx, a, algo)
        st = spgemm_check(ctx, ap, t(ap), label=f"rmat13 AA^T {algo}")
        nnzc.add(st.nnzC)
    assert len(nnzc) == 1


def test_rmat13_aat_symmetric_and_chunked(ctx):
    a = ts.gen_rmat(13, 16, seed=2)
    da, ta = device_tiles(ctx, a)
    dat = ctx.transpose(da)
    tb = ctx.csr2tile(dat)
    ctx.build_csc(tb)
    ctx.build_rowindex(tb)
    bounds, tot = ctx.chunk_plan(ta, tb, 2_000_000)
    log(f"[rmat] {len(bounds) - 1} chunks, {tot} products")
    assert len(bounds) > 3
    rows = []
    for r0, r1 in zip(bounds, bounds[1:]):
        cc, _ = ctx.spgemm_tile_rows(ta, tb, r0, r1)
        dc = ctx.tile2csr(cc)
        rows.append(ctx.download(dc))
        ctx.free(dc)
        ctx.free(cc)
    rp = np.concatenate([[0]] + [r.rowptr[1:] + off for r, off in
                                 zip(rows, np.cumsum([0] + [r.nnz for r in rows[:-1]]))]).astype(np.int32)
    c = ts.Csr(a.m, a.m, rp, np.concatenate([r.colidx for r in rows]), np.concatenate([r.val for r in rows]))
    rpc, cc_, vc = oracle.spgemm(a, t(a))
    assert np.array_equal(c.rowptr.astype(np.int64), rpc) and np.array_equal(c.colidx, cc_)
    assert np.array_equal(c.val.view(np.int64), vc.view(np.int64))
    ct = t(c)
    assert np.array_equal(ct.rowptr, c.rowptr) and np.array_equal(ct.colidx, c.colidx)
    assert np.array_equal(ct.val.view(np.int64), c.val.view(np.int64)), "AA^T not bit-symmetric"
    for o in (tb, dat, ta, da):
        ctx.free(o)
