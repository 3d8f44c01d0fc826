"""Shared GPU-vs-oracle checks for the parity tests."""
import sys

import numpy as np

import oracle
import tilespgemm as ts


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_tiles(ctx, a: ts.Csr):
    d = ctx.upload(a)
    t = ctx.csr2tile(d)
    ctx.build_csc(t)
    return d, t


def check_tiles_equal(got, ref, packed_cols=None):
    for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "csr_ptr", "mask", "val"):
        assert np.array_equal(got[k], ref[k]), k
    if packed_cols is None:
        assert np.array_equal(got["csr_col"], ref["csr_col"]), "csr_col"
    else:
        assert np.array_equal(packed_cols, ref["csr_col"]), "packed csr_col"


def spgemm_check(ctx, a: ts.Csr, b: ts.Csr | None = None, tune=None, label="", values=True):
    """C = A*B (B = A when None) on the device; compares every array of C
    with the oracle: step-1 tile list, step-2 masks / u8 offsets / tile
    offsets, and C's CSR (tile2csr) -- structure bit-exact, values
    bit-exact (and, as the stated bar, within 1e-12 relative to
    sum |a_ik b_kj|)."""
    log(f"[spgemm_check] {label} A {a.m}x{a.n} nnz {a.nnz}" + ("" if b is None else f" B {b.m}x{b.n} nnz {b.nnz}"))
    da, ta = device_tiles(ctx, a)
    if b is None:
        db, tb, bb = da, ta, a
    else:
        db, tb = device_tiles(ctx, b)
        bb = b
    tc, st = ctx.spgemm(ta, tb, tune)
    C = ctx.download_tiles(tc)
    # oracle
    oa = oracle.csr2tile(a)
    ob = oa if b is None else oracle.csr2tile(bb)
    cptr, ccol = oracle.step1_tiles(oa, ob)
    assert np.array_equal(C["tile_ptr"], cptr), "step-1 tile pointer"
    assert np.array_equal(C["tile_colidx"], ccol), "step-1 tile columns"
    assert np.array_equal(C["tile_rowidx"], np.repeat(np.arange(oa["tilem"], dtype=np.int32), np.diff(cptr)))
    rpc, cc = oracle.spgemm_structure(a, bb)
    mask, ptr16, tnnz = oracle.c_tile_masks(a.m, cptr, ccol, rpc, cc)
    assert np.array_equal(C["mask"], mask), "step-2 masks"
    assert np.array_equal(C["csr_ptr"], ptr16), "step-2 row offsets"
    assert np.array_equal(C["tile_nnz"].astype(np.int64), tnnz), "step-2 tile offsets"
    assert st.numtileC == len(ccol) and st.nnzC == rpc[-1]
    dc = ctx.tile2csr(tc)
    c = ctx.download(dc)
    assert np.array_equal(c.rowptr.astype(np.int64), rpc), "C row pointers"
    assert np.array_equal(c.colidx, cc), "C column indices"
    if values:
        vc = oracle.spgemm_values(a, bb, rpc, cc)
        # stated bar: |dv| <= 1e-12 * sum|a||b| per entry; achieved: bit-exact
        absa = ts.Csr(a.m, a.n, a.rowptr, a.colidx, np.abs(a.val))
        absb = ts.Csr(bb.m, bb.n, bb.rowptr, bb.colidx, np.abs(bb.val))
        bnd = oracle.spgemm_values(absa, absb, rpc, cc)
        assert np.all(np.abs(c.val - vc) <= 1e-12 * bnd), "values beyond 1e-12"
        assert np.array_equal(c.val.view(np.int64), vc.view(np.int64)), "values not bit-exact"
    # the tile-parallel kernels must give the identical C (every array)
    forced = ts.default_tune() if tune is None else type(tune).from_buffer_copy(tune)
    forced.force_tile_parallel = 1
    tc2, st2 = ctx.spgemm(ta, tb, forced)
    C2 = ctx.download_tiles(tc2)
    for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "csr_ptr", "mask", "csr_col", "val"):
        assert np.array_equal(C2[k], C[k]), f"tile-parallel path differs in {k}"
    ctx.free(tc2)
    # every step-1 bitmap window gives the same tile list
    for words in (256, 1024, 2048):
        tw = ts.default_tune() if tune is None else type(tune).from_buffer_copy(tune)
        tw.s1_window_words = words
        tcw, _ = ctx.spgemm(ta, tb, tw)
        Cw = ctx.download_tiles(tcw)
        for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "mask", "val"):
            assert np.array_equal(Cw[k], C[k]), f"step-1 window {words} differs in {k}"
        ctx.free(tcw)
    # element-wise numeric step with B's row index: for every tile-row, and
    # for the tile-rows the row-wise tile kernel cannot hold
    ctx.build_rowindex(tb)
    # first with an A that has no row index (A's entries decoded from its tiles)
    da2, ta2 = device_tiles(ctx, a)
    for mode in (2, 0):
        t3 = ts.default_tune() if tune is None else type(tune).from_buffer_copy(tune)
        t3.element_mode = mode
        tc3, _ = ctx.spgemm(ta2, tb, t3)
        C3 = ctx.download_tiles(tc3)
        for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "csr_ptr", "mask", "csr_col", "val"):
            assert np.array_equal(C3[k], C[k]), f"element mode {mode}, A tiles only, differs in {k}"
        ctx.free(tc3)
    ctx.free(ta2)
    ctx.free(da2)
    # with A's row index too, and rows above 64 products computed one wave
    # each (the heavy-row path: column parts of 4096 columns in LDS, one run
    # each; with items of ~256 products, runs of column parts go to
    # different waves and each run's later parts start from the cursors)
    configs = [(2, 0, 0), (0, 0, 0)]
    if b is not None:
        ctx.build_rowindex(ta)
    configs += [(2, 64, 0), (2, 64, 256)]
    for mode, heavy, item in configs:
        t3 = ts.default_tune() if tune is None else type(tune).from_buffer_copy(tune)
        t3.element_mode = mode
        t3.heavy_min_products = heavy
        t3.heavy_item_products = item
        tc3, _ = ctx.spgemm(ta, tb, t3)
        C3 = ctx.download_tiles(tc3)
        for k in ("tile_ptr", "tile_colidx", "tile_rowidx", "tile_nnz", "csr_ptr", "mask", "csr_col", "val"):
            assert np.array_equal(C3[k], C[k]), f"element mode {mode} / heavy {heavy} / item {item} differs in {k}"
        ctx.free(tc3)
    for obj in (dc, tc):
        ctx.free(obj)
    ctx.free(ta)
    ctx.free(da)
    if b is not None:
        ctx.free(tb)
        ctx.free(db)
    return st
