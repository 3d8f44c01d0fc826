"""Device orderings and permutations vs the oracle: degree order and RCM
bit-exact (same definition, ties by (degree, id)), random order a valid
seeded permutation, P*A*Q^T and its restore bit-exact, and SpGEMM under a
reordering giving C back after restore."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import log

pytestmark = pytest.mark.gpu


def scrambled_banded(n, bw, seed):
    b = ts.gen_banded(n, bw, seed=seed)
    rng = np.random.default_rng(seed)
    p = rng.permutation(n).astype(np.int32)
    inv = np.empty_like(p)
    inv[p] = np.arange(n, dtype=np.int32)
    return ts.Csr(n, n, *oracle.permute_csr(b, p, inv))


def graphs():
    yield "5pt", ts.gen_stencil5(30, 23)
    yield "27p12", ts.gen_stencil27(12, 12, 12, True)
    yield "banded_scr", scrambled_banded(3000, 4, 1)
    yield "er_sparse", ts.gen_er(4000, 1.5, seed=2)       # many components, isolated vertices
    yield "rmat12", ts.gen_rmat(12, 16, seed=3)
    yield "er_nonsym", ts.gen_er(2500, 6.0, seed=4)
    yield "diag", ts.Csr(50, 50, np.arange(51, dtype=np.int32), np.arange(50, dtype=np.int32), np.ones(50))
    yield "empty", ts.Csr(20, 20, np.zeros(21, np.int32), np.zeros(0, np.int32), np.zeros(0))


@pytest.mark.parametrize("name,a", list(graphs()), ids=[n for n, _ in graphs()])
def test_degree_and_rcm_match_oracle(ctx, name, a):
    log(f"[reorder] {name}")
    d = ctx.upload(a)
    assert np.array_equal(ctx.order_degree(d), oracle.degree_order(a))
    assert np.array_equal(ctx.order_degree(d, True), oracle.degree_order(a, True))
    if a.m:
        g = ctx.download(ctx.symmetrize(d))
        gp, gc = oracle.symmetrize(a)
        assert np.array_equal(g.rowptr, gp) and np.array_equal(g.colidx, gc)
        order, levels = ctx.order_rcm(d)
        ref = oracle.rcm(a)
        assert np.array_equal(order, ref), f"rcm differs at {np.argmax(order != ref)}"
        assert levels >= 1
    ctx.free(d)


def test_random_order(ctx):
    a = ctx.order_random(10000, 5)
    b = ctx.order_random(10000, 5)
    c = ctx.order_random(10000, 6)
    assert np.array_equal(np.sort(a), np.arange(10000))
    assert np.array_equal(a, b) and not np.array_equal(a, c)


@pytest.mark.parametrize("name,a", list(graphs())[:6], ids=[n for n, _ in list(graphs())[:6]])
def test_permute_and_restore(ctx, name, a):
    rng = np.random.default_rng(11)
    old = rng.permutation(a.m).astype(np.int32)
    new = np.empty_like(old)
    new[old] = np.arange(a.m, dtype=np.int32)
    d = ctx.upload(a)
    p = ctx.permute(d, old, new)
    got = ctx.download(p)
    ref = oracle.permute_csr(a, old, new)
    assert np.array_equal(got.rowptr, ref[0]) and np.array_equal(got.colidx, ref[1]) and np.array_equal(got.val, ref[2])
    back = ctx.download(ctx.permute(p, new, old))   # restore: rows by new_id, columns by old_id
    assert np.array_equal(back.rowptr, a.rowptr) and np.array_equal(back.colidx, a.colidx)
    assert np.array_equal(back.val, a.val)


def test_permute_rejects_non_permutation(ctx):
    a = ts.gen_stencil5(8, 8)
    d = ctx.upload(a)
    bad = np.arange(64, dtype=np.int32)
    bad[3] = 5
    with pytest.raises(ts.TsgError):
        ctx.permute(d, bad, np.arange(64, dtype=np.int32))


@pytest.mark.parametrize("algo", ["rcm", "degree", "random"])
def test_spgemm_under_reordering_restores_c(ctx, algo):
    """C computed on P A P^T and restored equals A^2: structure exact,
    integer-valued entries exact (the accumulation order changes with the
    inner permutation, so float values are compared to 1e-12 of sum|ab|)."""
    a = ts.set_values(scrambled_banded(4000, 6, 3), ts.VAL_MOD10)
    d = ctx.upload(a)
    if algo == "rcm":
        old, _ = ctx.order_rcm(d)
    elif algo == "degree":
        old = ctx.order_degree(d)
    else:
        old = ctx.order_random(a.m, 9)
    new = np.empty_like(old)
    new[old] = np.arange(a.m, dtype=np.int32)
    dp = ctx.permute(d, old, new)
    tp = ctx.csr2tile(dp)
    ctx.build_csc(tp)
    tc, st = ctx.spgemm(tp, tp)
    cp = ctx.tile2csr(tc)
    c = ctx.download(ctx.permute(cp, new, old))
    rpc, cc, vc = oracle.spgemm(a, a)
    assert np.array_equal(c.rowptr.astype(np.int64), rpc) and np.array_equal(c.colidx, cc)
    assert np.array_equal(c.val, vc)
    if algo == "rcm":
        t0 = ctx.csr2tile(d)
        assert tp.numtile < t0.numtile  # RCM regroups the scrambled band into far fewer tiles
