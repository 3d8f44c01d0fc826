"""Batched SpGEMM on the GPU (tsg_batch): every problem's C from the fused,
graph-launched batch must be identical, array by array and bit for bit,
to a plain tsg_spgemm call on the same operands; a sample is also checked
against the oracle's CSR."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import check_tiles_equal, device_tiles, log

pytestmark = pytest.mark.gpu


def plain_tiles(ctx, a: ts.Csr, op):
    da, ta = device_tiles(ctx, a)
    if op == ts.OP_A2:
        tc, _ = ctx.spgemm(ta, ta)
        ctx.free(ta)
    else:
        dat = ctx.transpose(da)
        tb = ctx.csr2tile(dat)
        ctx.build_csc(tb)
        tc, _ = ctx.spgemm(ta, tb)
        for o in (ta, tb, dat):
            ctx.free(o)
    out = ctx.download_tiles(tc)
    ctx.free(tc)
    ctx.free(da)
    return out


def oracle_csr(a: ts.Csr, op):
    b = a
    if op == ts.OP_AAT:
        tp, tc, tv = oracle.transpose(a)
        b = ts.Csr(a.n, a.m, tp, tc, tv)
    rpc, cc = oracle.spgemm_structure(a, b)
    return rpc, cc, oracle.spgemm_values(a, b, rpc, cc)


def run_batch(ctx, mats, ops, tune=None, check_oracle=(), autotune=False):
    dps = [ctx.upload(a) for a in mats]
    b = ts.Batch(ctx, dps, ops, tune)
    try:
        if autotune:
            t, ms = b.autotune(reps=1)
            log(f"[batch] autotuned: tile_parallel={t.force_tile_parallel} dense_min={t.dense_acc_min_nnz} {ms:.3f} ms")
        for rep in range(2):
            b.run()
            ctx.sync()
            for i, (a, op) in enumerate(zip(mats, ops)):
                got = b.result_tiles(i)
                ref = plain_tiles(ctx, a, op)
                for k in ("m", "n", "tilem", "tilen", "numtile", "nnz"):
                    assert got[k] == ref[k], (i, k)
                check_tiles_equal(got, ref)
                if rep == 0 and i in check_oracle:
                    rpc, cc, vc = oracle_csr(a, op)
                    c = b.result_csr(i)
                    assert np.array_equal(c.rowptr.astype(np.int64), rpc), i
                    assert np.array_equal(c.colidx, cc), i
                    assert np.array_equal(c.val.view(np.int64), vc.view(np.int64)), i
        return b.info()
    finally:
        b.close()
        for d in dps:
            ctx.free(d)


def spec_mats(count, seed, nmax):
    specs = ts.gen_batch_spec(count, seed=seed, nmin=100, nmax=nmax)
    mats = [ts.gen_batch_problem(s) for s in specs]
    return mats, [s.op for s in specs]


def test_batch_spec_mix(ctx):
    mats, ops = spec_mats(24, 2, 3000)
    log("[batch] 24 problems n<=3000", [m.m for m in mats])
    info = run_batch(ctx, mats, ops, check_oracle=(0, 1, 5, 11))
    assert info.count == 24 and info.groups == 1


def test_batch_edge_shapes(ctx):
    """Empty matrix, n not a multiple of 16, 1x1, rectangular AA^T, a dense
    block: all in one batch."""
    er = ts.gen_er(517, 9.0, seed=4)
    rect = ts.Csr(200, er.n, er.rowptr[:201], er.colidx[:er.rowptr[200]], er.val[:er.rowptr[200]])
    empty = ts.Csr(40, 40, np.zeros(41, np.int32), np.zeros(0, np.int32), np.zeros(0))
    one = ts.Csr(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([0.5]))
    dense = ts.Csr.from_scipy(__import__("scipy.sparse", fromlist=["x"]).random(
        48, 48, density=0.9, random_state=5, format="csr"))
    mats = [er, rect, empty, one, dense, ts.gen_banded(333, 7, seed=6), rect]
    ops = [ts.OP_A2, ts.OP_AAT, ts.OP_A2, ts.OP_AAT, ts.OP_A2, ts.OP_AAT, ts.OP_AAT]
    run_batch(ctx, mats, ops, check_oracle=(1, 4))


def test_batch_single_problem(ctx):
    a = ts.gen_stencil27(20, 20, 20, periodic=True, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    run_batch(ctx, [a], [ts.OP_A2])


def test_batch_groups_split(ctx):
    """A small per-group nonzero cap forces several fused groups inside the
    one graph; results are unchanged."""
    mats, ops = spec_mats(12, 7, 2000)
    sizes = [plain_tiles(ctx, a, op)["nnz"] for a, op in zip(mats, ops)]
    tune = ts.default_tune()
    tune.batch_group_nnz_cap = max(sizes)  # every problem fits alone, not all together
    info = run_batch(ctx, mats, ops, tune=tune)
    log(f"[batch] groups {info.groups} (cap {max(sizes)}, total {sum(sizes)})")
    assert info.groups > 1 and info.nnzC == sum(sizes)


def test_batch_autotune_same_results(ctx):
    mats, ops = spec_mats(10, 9, 2500)
    run_batch(ctx, mats, ops, autotune=True)


def test_batch_rejects_bad_input(ctx):
    a = ts.gen_er(50, 3.0, seed=1)
    rect = ts.Csr(20, 50, a.rowptr[:21], a.colidx[:a.rowptr[20]], a.val[:a.rowptr[20]])
    d = ctx.upload(rect)
    with pytest.raises(ts.TsgError):
        ts.Batch(ctx, [d], [ts.OP_A2])  # A^2 of a non-square matrix
    with pytest.raises(ts.TsgError):
        ts.Batch(ctx, [d], [7])
    ctx.free(d)
