"""SpGEMM parity on the GPU: C = A^2 and C = A*A^T through the device C ABI,
every array of C compared with the oracle (tests/tsg_check.py): step-1 tile
list, step-2 masks / offsets, C's CSR structure bit-exact, values bit-exact
(stated bar 1e-12 relative to sum |a_ik b_kj|, asserted too)."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import log, spgemm_check

pytestmark = pytest.mark.gpu


def t(csr):
    tp, tc, tv = oracle.transpose(csr)
    return ts.Csr(csr.n, csr.m, tp, tc, tv)


def test_5pt_64_square_closed_forms(ctx):
    a = ts.gen_stencil5(64, 64)
    st = spgemm_check(ctx, a, label="5pt64 A^2")
    assert st.numtileC == 2652 and st.nnzC == 51972 and st.tiles_empty == 256


def test_5pt_64_mod10_values(ctx):
    a = ts.set_values(ts.gen_stencil5(64, 64), ts.VAL_MOD10)
    spgemm_check(ctx, a, label="5pt64 i%10")


@pytest.mark.parametrize("N,periodic", [(16, True), (24, True), (32, True), (20, False), (33, False)])
def test_27pt(ctx, N, periodic):
    a = ts.gen_stencil27(N, N, N, periodic=periodic, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    st = spgemm_check(ctx, a, label=f"27{'p' if periodic else 'd'}{N}")
    if periodic:
        assert st.nnzC == 125 * N ** 3


@pytest.mark.parametrize("n,mean,seed", [(1000, 4.0, 1), (5000, 16.0, 2), (3001, 32.0, 3)])
def test_er_square_and_aat(ctx, n, mean, seed):
    a = ts.gen_er(n, mean, seed=seed)
    spgemm_check(ctx, a, label=f"er{n} A^2")
    spgemm_check(ctx, a, t(a), label=f"er{n} AA^T")


@pytest.mark.parametrize("n,bw", [(1000, 2), (4097, 40), (2500, 17)])
def test_banded(ctx, n, bw):
    a = ts.gen_banded(n, bw, seed=5)
    spgemm_check(ctx, a, label=f"banded{n}/{bw}")


@pytest.mark.parametrize("scale", [8, 11, 13])
def test_rmat(ctx, scale):
    a = ts.gen_rmat(scale, 16, seed=scale)
    spgemm_check(ctx, a, label=f"rmat{scale} A^2")
    spgemm_check(ctx, a, t(a), label=f"rmat{scale} AA^T")


def test_rectangular_and_edges(ctx):
    rng = np.random.default_rng(9)
    import scipy.sparse as sp
    for (m, k, n, d) in [(37, 300, 21, 0.05), (5, 7, 3, 0.9), (300, 17, 290, 0.2), (16, 16, 16, 1.0), (1, 1, 1, 1.0)]:
        a = sp.random(m, k, density=d, random_state=rng, format="csr")
        b = sp.random(k, n, density=d, random_state=rng, format="csr")
        a.data = rng.uniform(-1, 1, a.nnz)
        b.data = rng.uniform(-1, 1, b.nnz)
        spgemm_check(ctx, ts.Csr.from_scipy(a), ts.Csr.from_scipy(b), label=f"rect {m}x{k}x{n}")


def test_empty_operands(ctx):
    z = ts.Csr(64, 64, np.zeros(65, np.int32), np.zeros(0, np.int32), np.zeros(0))
    st = spgemm_check(ctx, z, label="empty")
    assert st.numtileC == 0 and st.nnzC == 0
    a = ts.gen_stencil5(8, 8)
    zz = ts.Csr(64, 10, np.zeros(65, np.int32), np.zeros(0, np.int32), np.zeros(0))
    spgemm_check(ctx, a, zz, label="A * empty")


def _coo(n, rows, cols, rng):
    import scipy.sparse as sp
    a = sp.csr_matrix((rng.uniform(-1, 1, len(rows)), (rows, cols)), shape=(n, n))
    a.sum_duplicates()
    a.sort_indices()
    return ts.Csr.from_scipy(a)


@pytest.mark.parametrize("case", ["all_single", "mixed", "empty_b_rows", "one_residue"])
def test_row_kernel_single_entry_tiles(ctx, case):
    """The row-wise numeric kernel's single-entry K-steps (an A tile with one
    nonzero (r, k) is never staged: the wave owning C row r loads row k of
    each B tile itself) on tile-rows made for it, all within the row-wise
    kernel's caps: only single-entry tiles (no staged step at all, several
    per owning wave back to back), single- and two-entry tiles mixed, single
    entries whose B tile-row is empty, and entries in one row residue mod 4
    only (every task and every single-entry step on one wave)."""
    rng = np.random.default_rng(11)
    n = 16 * 48
    i = np.arange(n)
    if case == "all_single":  # one entry per row, columns a permutation
        rows, cols = i, rng.permutation(n)
    elif case == "mixed":  # a random entry per row, and beside it every third row
        c = rng.integers(0, n - 1, n)
        two = i % 3 == 0
        rows, cols = np.concatenate([i, i[two]]), np.concatenate([c, c[two] + 1])
    elif case == "empty_b_rows":  # rows of tile-rows 36..47 empty, columns anywhere
        live = i[i < 16 * 36]
        rows = np.concatenate([live, live])
        cols = np.concatenate([rng.integers(0, n, len(live)), rng.integers(16 * 36, n, len(live))])
    else:  # rows 4q + 3 only, one or two entries each
        live = i[i % 4 == 3]
        c = rng.integers(0, n - 1, len(live))
        two = rng.random(len(live)) < 0.5
        rows, cols = np.concatenate([live, live[two]]), np.concatenate([c, c[two] + 1])
    a = _coo(n, rows, cols, rng)
    st = spgemm_check(ctx, a, label=f"single-entry tiles: {case}")
    assert st.t_step3_rows_ms > 0, "the row-wise numeric kernel took no tile-row"
    spgemm_check(ctx, a, t(a), label=f"single-entry tiles: {case}, AA^T")


def test_step1_kept_bitmaps_runs_cross_waves(ctx):
    """C = I * B with B's tile-rows holding runs of consecutive tile columns
    (kept as step-1 bitmaps) of lengths that put the 64-tile boundaries of
    the bitmap builder inside runs at every lane offset, next to sparse
    tile-rows that are not kept: C's tile list is B's, compared with the
    oracle array by array."""
    rng = np.random.default_rng(17)
    ntc = 400
    n = 16 * ntc
    rows, cols = [], []
    for K in range(n // 16):
        if K % 5 == 4:  # sparse: a few tiles far apart (not kept)
            js = np.unique(rng.integers(0, ntc, 3))
        else:  # a run of consecutive tile columns, some gaps
            L = 1 + (K * 37) % 150
            j0 = int(rng.integers(0, ntc - L + 1))
            js = np.arange(j0, j0 + L)
            js = js[rng.random(L) < 0.9] if L > 4 else js
        for J in js:
            rows.append(16 * K + int(rng.integers(0, 16)))
            cols.append(16 * int(J) + int(rng.integers(0, 16)))
    rows, cols = np.array(rows), np.array(cols)
    b = _coo(n, rows, cols, rng)
    eye = _coo(n, np.arange(n), np.arange(n), rng)
    st = spgemm_check(ctx, eye, b, label="I * B, kept bitmap runs")
    assert st.numtileC == len(set(zip((rows // 16).tolist(), (cols // 16).tolist())))


def test_wide_tilerow_uses_tile_parallel_step2(ctx):
    """A hub row touching 700 tile columns makes C tile-row 0 hold 700 tiles
    (> the 512 the row-wise step-2 kernel keeps in LDS), so that row goes
    through the tile-parallel step-2 kernel while the others do not."""
    n = 16 * 700
    rows, cols = [], []
    for i in range(n):
        cs = {i}
        if i == 0:
            cs |= set(range(0, n, 16))
        if i % 97 == 3:
            cs |= {(i * 7919) % n, (i * 104729) % n}
        rows += [i] * len(cs)
        cols += sorted(cs)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    rng = np.random.default_rng(1)
    a = ts.Csr(n, n, rp, np.array(cols, np.int32), rng.uniform(-1, 1, len(cols)))
    spgemm_check(ctx, a, label="hub row 700 tiles")
    spgemm_check(ctx, a, t(a), label="hub row 700 tiles AA^T")


def test_dense_block_full_tiles(ctx):
    """A fully dense 64x64 block: every C tile holds 256 nonzeros (the
    reference's 'ful' bin) and 64 products per entry."""
    n = 64
    rng = np.random.default_rng(3)
    rp = np.arange(0, n * n + 1, n, dtype=np.int32)
    ci = np.tile(np.arange(n, dtype=np.int32), n)
    a = ts.Csr(n, n, rp, ci, rng.uniform(-1, 1, n * n))
    st = spgemm_check(ctx, a, label="dense64")
    assert st.tiles_dense_acc == 16 and st.nnzC == n * n


@pytest.mark.parametrize("th", [1, 33, 225, 257])
def test_accumulator_threshold_does_not_change_results(ctx, th):
    """Every C tile through the dense accumulator (th=1), the reference's
    lrg/dns split points, or none (257): all bit-identical to the oracle."""
    tune = ts.default_tune()
    tune.dense_acc_min_nnz = th
    a = ts.gen_er(4000, 20.0, seed=8)
    spgemm_check(ctx, a, tune=tune, label=f"er th={th}")
    b = ts.gen_stencil27(16, 16, 16, True)
    spgemm_check(ctx, b, tune=tune, label=f"27p16 th={th}")


def test_repeat_is_deterministic(ctx):
    a = ts.gen_rmat(12, 16, seed=5)
    da = ctx.upload(a)
    ta = ctx.csr2tile(da)
    ctx.build_csc(ta)
    outs = []
    for _ in range(3):
        tc, _ = ctx.spgemm(ta, ta)
        c = ctx.download_tiles(tc)
        outs.append(c["val"].view(np.int64).copy())
        ctx.free(tc)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    ctx.free(ta)
    ctx.free(da)


def test_wide_step1_window_rmat(ctx):
    """R-MAT scale 14 A*A^T, whose rows' candidates span the whole column
    range: the check runs every step-1 window (256, 1024 and 2048 words), all
    arrays identical to the oracle."""
    a = ts.gen_rmat(14, 8, seed=21)
    spgemm_check(ctx, a, t(a), label="rmat14 AAT windows")


def test_wide_b_element_hash_unpacked(ctx):
    """C = A*B with B 3000 x 200000 (12,500 tile columns, random entries):
    C's tile-rows span more than 8192 tile columns, so the element-wise
    numeric kernel keeps its hash keys and positions in separate arrays
    (narrower products pack both into one 32-bit slot); every array of C
    is compared with the oracle through all element-wise modes."""
    rng = np.random.default_rng(77)
    a = ts.gen_er(3000, 6.0, seed=77)
    m, n, nnz = 3000, 200000, 30000
    key = np.unique(rng.integers(0, m, nnz).astype(np.int64) * n + rng.integers(0, n, nnz))
    r, c = (key // n).astype(np.int32), (key % n).astype(np.int32)
    rp = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=m))]).astype(np.int32)
    b = ts.Csr(m, n, rp, c, rng.uniform(-1, 1, len(c)))
    spgemm_check(ctx, a, b, label="ER3000 x wide B (200000 columns)")


def test_bad_step1_window_rejected(ctx):
    a = ts.gen_stencil5(32, 32)
    da = ctx.upload(a)
    ta = ctx.csr2tile(da)
    ctx.build_csc(ta)
    tune = ts.default_tune()
    tune.s1_window_words = 100
    with pytest.raises(ts.TsgError):
        ctx.spgemm(ta, ta, tune)
    ctx.free(ta)
    ctx.free(da)
