"""The R-MAT scale-20, batched and served workloads at their full sizes.

Products this large cannot be downloaded whole or checked by the oracle
whole, so each test combines
  * order-free device checksums of C (tsg_tiles_checksum: every tile's
    masks / row offsets / nonzero count and every entry's (row, column,
    value bits)), pinned below against a numpy restatement on a small
    product, which add up over tile-row chunks;
  * size-independent properties (nnz(C) invariant under orderings, the
    chunk plan's products equal nnzCub, A*A^T bit-symmetric);
  * sampled rows -- the 16 highest-degree rows plus random ones --
    bit-exact against the oracle (oracle.spgemm_rowlist, the serial SPA
    value order of src/external/cusparse/spgemm_serialref_spa.h:18-29).
"""
import threading
import time

import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import device_tiles, log

pytestmark = pytest.mark.gpu

U64 = np.uint64


def np_mix(x):
    x = x ^ (x >> U64(30))
    x = x * U64(0xBF58476D1CE4E5B9)
    x = x ^ (x >> U64(27))
    x = x * U64(0x94D049BB133111EB)
    return x ^ (x >> U64(31))


def np_checksum(T, csr, row_base=0, jbase=0):
    """Restatement of tsg_tiles_checksum from downloaded tiles and their CSR."""
    tp = T["tile_ptr"].astype(np.int64)
    tilem = len(tp) - 1
    I = np.repeat(np.arange(tilem, dtype=np.int64), np.diff(tp))
    J = T["tile_colidx"].astype(np.int64) - jbase
    tk = np_mix(((U64(row_base >> 4) + I.astype(U64)) << U64(32)) ^ J.astype(U64))
    nz = np.diff(T["tile_nnz"].astype(np.int64)).astype(U64)
    M = (1 << 64) - 1
    ht = int(np.sum(np_mix(tk + nz), dtype=U64))
    mask = T["mask"].reshape(-1, 16).astype(U64)
    ptr = T["csr_ptr"].reshape(-1, 16).astype(U64)
    for r in range(16):
        ok = 16 * I + r < csr.m
        ht = (ht + int(np.sum(np_mix(tk[ok] ^ (U64(r + 1) << U64(40)) ^ (mask[ok, r] << U64(8)) ^ ptr[ok, r]),
                              dtype=U64))) & M
    rows = (np.repeat(np.arange(csr.m, dtype=np.int64), np.diff(csr.rowptr)) + row_base).astype(U64)
    cols = csr.colidx.astype(np.int64).astype(U64)
    vb = csr.val.view(U64)
    he = np.sum(np_mix(np_mix((rows << U64(32)) ^ cols) ^ vb), dtype=U64)
    hx = np.sum(np_mix(np_mix((cols << U64(32)) ^ rows) ^ vb), dtype=U64)
    return ht, int(he), int(hx)


def transpose_host(a):
    tp, tc, tv = oracle.transpose(a)
    return ts.Csr(a.n, a.m, tp, tc, tv)


def check_rows(got: ts.Csr, ref, label):
    rp, cc, vv = ref
    assert np.array_equal(np.asarray(got.rowptr, np.int64), rp), f"{label}: row lengths"
    assert np.array_equal(got.colidx, cc), f"{label}: columns"
    assert np.array_equal(got.val.view(np.int64), vv.view(np.int64)), f"{label}: values not bit-exact"


def sample_rows(a: ts.Csr, nrand, seed, hubs=16):
    deg = np.diff(a.rowptr)
    top = np.argsort(-deg, kind="stable")[:hubs]
    rnd = np.random.default_rng(seed).choice(a.m, min(nrand, a.m), replace=False)
    return np.unique(np.concatenate([top, rnd])).astype(np.int32), top


# ------------------------------------------------------------------ checksum pin
def test_checksum_restatement_and_chunk_sums(ctx):
    """The device checksums equal their numpy restatement on a small A*A^T,
    the chunks' checksums add up to the one-shot product's, entries ==
    entries_t (bit-symmetric), and a planted change in C's values changes
    the restated checksum (the checksum sees every value bit)."""
    a = ts.gen_rmat(12, 16, seed=3)
    da, ta = device_tiles(ctx, a)
    dat = ctx.transpose(da)
    tb = ctx.csr2tile(dat)
    ctx.build_csc(tb)
    ctx.build_rowindex(tb)
    tc, st = ctx.spgemm(ta, tb)
    cs = ctx.checksum(tc)
    T = ctx.download_tiles(tc)
    dc = ctx.tile2csr(tc)
    c = ctx.download(dc)
    ht, he, hx = np_checksum(T, c)
    assert (cs.numtile, cs.nnz) == (st.numtileC, st.nnzC)
    assert (cs.tiles, cs.entries, cs.entries_t) == (ht, he, hx)
    assert cs.entries == cs.entries_t, "A*A^T not bit-symmetric"
    c2 = ts.Csr(c.m, c.n, c.rowptr, c.colidx, c.val.copy())
    c2.val[c.nnz // 2] = np.nextafter(c2.val[c.nnz // 2], 2.0)
    assert np_checksum(T, c2)[1] != he
    bounds, _ = ctx.chunk_plan(ta, tb, 300000)
    assert len(bounds) > 3
    tot = ts.Csum()
    for r0, r1 in zip(bounds, bounds[1:]):
        cc, _ = ctx.spgemm_tile_rows(ta, tb, r0, r1)
        tot.add(ctx.checksum(cc, row_base=16 * r0))
        rows = np.arange(16 * r0, min(16 * r1, a.m), 7)
        got = ctx.tile_rows_csr(cc, rows - 16 * r0)
        ref = (np.concatenate([[0], np.cumsum(np.diff(c.rowptr)[rows])]),
               np.concatenate([c.colidx[c.rowptr[i]:c.rowptr[i + 1]] for i in rows]),
               np.concatenate([c.val[c.rowptr[i]:c.rowptr[i + 1]] for i in rows]))
        check_rows(got, ref, f"chunk [{r0},{r1}) rows")
        ctx.free(cc)
    assert tot.key() == cs.key() and tot.entries_t == cs.entries_t
    for o in (dc, tc, tb, dat, ta, da):
        ctx.free(o)


# ------------------------------------------------------------------ R-MAT scale 20
def test_rmat20_aat_orderings_fullsize(ctx):
    """R-MAT scale 20 (edge factor 16, (0.57, 0.19, 0.19, 0.05),
    symmetrized, values U(0,1], seed 1), C = A*A^T under the device random,
    RCM and degree orders, in tile-row chunks of <= 1.5e9 products, each
    chunk checked and dropped."""
    t0 = time.perf_counter()
    a = ts.gen_rmat(20, 16, seed=1)
    log(f"[rmat20] generated n={a.m} nnz={a.nnz} in {time.perf_counter() - t0:.1f} s")
    d = ctx.upload(a)
    nnz_by_order = {}
    for algo in ("random", "rcm", "degree"):
        t0 = time.perf_counter()
        if algo == "rcm":
            old, _ = ctx.order_rcm(d)
        elif algo == "degree":
            old = ctx.order_degree(d)
        else:
            old = ctx.order_random(a.m, 1)
        assert np.array_equal(np.sort(old), np.arange(a.m))
        new = np.empty_like(old)
        new[old] = np.arange(a.m, dtype=np.int32)
        dp = ctx.permute(d, old, new)
        ap = ctx.download(dp)
        apt = transpose_host(ap)
        ta = ctx.csr2tile(dp)
        ctx.build_csc(ta)
        ctx.build_rowindex(ta)
        dpt = ctx.transpose(dp)
        tb = ctx.csr2tile(dpt)
        ctx.build_csc(tb)
        ctx.build_rowindex(tb)
        bounds, prods = ctx.chunk_plan(ta, tb, 1_500_000_000)
        assert prods == oracle.nnzcub(ap, apt), "chunk products != nnzCub"
        rows, hubs = sample_rows(ap, 2000, seed=7)
        got_rows = {}
        tot = ts.Csum()
        nnz = tiles = 0
        for r0, r1 in zip(bounds, bounds[1:]):
            cc, _ = ctx.spgemm_tile_rows(ta, tb, r0, r1, timed=False)
            cs = ctx.checksum(cc, row_base=16 * r0)
            assert (cs.numtile, cs.nnz) == (cc.numtile, cc.nnz)
            tot.add(cs)
            nnz += cc.nnz
            tiles += cc.numtile
            sel = rows[(rows >= 16 * r0) & (rows < 16 * r1)]
            if len(sel):
                part = ctx.tile_rows_csr(cc, sel - 16 * r0)
                for k, i in enumerate(sel):
                    s, e = part.rowptr[k], part.rowptr[k + 1]
                    got_rows[int(i)] = (part.colidx[s:e].copy(), part.val[s:e].copy())
            ctx.free(cc)
        assert tot.nnz == nnz and tot.numtile == tiles
        assert tot.entries == tot.entries_t, f"{algo}: A*A^T not bit-symmetric"
        ref = oracle.spgemm_rowlist(ap, apt, rows)
        got = ts.Csr(len(rows), a.m,
                     np.concatenate([[0], np.cumsum([len(got_rows[int(i)][0]) for i in rows])]),
                     np.concatenate([got_rows[int(i)][0] for i in rows]),
                     np.concatenate([got_rows[int(i)][1] for i in rows]))
        check_rows(got, ref, f"rmat20 {algo} sampled rows")
        nnz_by_order[algo] = nnz
        log(f"[rmat20] {algo}: {len(bounds) - 1} chunks, nnzCub {prods}, nnzC {nnz}, tiles {tiles}, "
            f"{len(rows)} rows (16 hubs of degree up to {int(np.diff(ap.rowptr)[hubs].max())}) bit-exact, "
            f"{time.perf_counter() - t0:.1f} s")
        for o in (tb, dpt, ta, dp):
            ctx.free(o)
        ctx.trim()
    assert len(set(nnz_by_order.values())) == 1, nnz_by_order
    ctx.free(d)
    ctx.trim()


# ------------------------------------------------------------------ batched 512
def test_batch512_fullsize(ctx):
    """The batched workload as benchmarked: 512 problems (seed 2, n
    log-uniform in [1000, 65536], Erdos-Renyi or banded, A^2 / A*A^T).
    Every problem's C from the graph-launched batch has the same tile
    count, nonzeros and checksums as a plain call on the same operands
    (two graph launches); 16 problems, the 8 largest among them, match the
    oracle's CSR bit for bit."""
    specs = ts.gen_batch_spec(512, seed=2, nmin=1000, nmax=65536)
    mats = [ts.gen_batch_problem(s) for s in specs]
    ops = [s.op for s in specs]
    log(f"[batch512] generated, total nnz {sum(m.nnz for m in mats)}")
    plain = []
    for i, (a, op) in enumerate(zip(mats, ops)):
        da, ta = device_tiles(ctx, a)
        tb, dat = ta, None
        if op == ts.OP_AAT:
            dat = ctx.transpose(da)
            tb = ctx.csr2tile(dat)
            ctx.build_csc(tb)
        tc, _ = ctx.spgemm(ta, tb, timed=False)
        plain.append(ctx.checksum(tc))
        for o in (tc, ta, da) + ((tb, dat) if dat is not None else ()):
            ctx.free(o)
    log("[batch512] plain calls done")
    dps = [ctx.upload(a) for a in mats]
    b = ts.Batch(ctx, dps, ops)
    try:
        for rep in range(2):
            b.run()
            ctx.sync()
            for i in range(512):
                got = b.result_checksum(i)
                assert got.key() == plain[i].key(), (rep, i, got.key(), plain[i].key())
        info = b.info()
        log(f"[batch512] {info.count} problems, {info.groups} groups, nnzC {info.nnzC}: all checksums equal")
        assert info.nnzC == sum(p.nnz for p in plain)
        ns = np.array([m.m for m in mats])
        pick = list(np.argsort(-ns, kind="stable")[:8])
        pick += list(np.random.default_rng(4).choice([i for i in range(512) if i not in pick], 8, replace=False))
        for i in pick:
            a = mats[i]
            bb = a if ops[i] == ts.OP_A2 else transpose_host(a)
            rpc, cc, vc = oracle.spgemm(a, bb)
            c = b.result_csr(int(i))
            assert np.array_equal(c.rowptr.astype(np.int64), rpc), i
            assert np.array_equal(c.colidx, cc), i
            assert np.array_equal(c.val.view(np.int64), vc.view(np.int64)), i
        log(f"[batch512] oracle: problems {sorted(int(x) for x in pick)} bit-exact (n up to {ns.max()})")
    finally:
        b.close()
        for d in dps:
            ctx.free(d)
        ctx.trim()


# ------------------------------------------------------------------ served pool
def serve_pool(seed=3, n_er=24, n_rmat=2):
    rng = np.random.default_rng(seed)
    ers = []
    for i in range(n_er):
        n = int(round(np.exp(rng.uniform(np.log(2000), np.log(262144)))))
        ers.append(ts.gen_er(n, float(rng.uniform(4, 32)), seed=1000 + i))
    return ers + [ts.gen_rmat(18, 16, seed=2000 + i) for i in range(n_rmat)], rng


def test_serve_bench_pool_poisson(ctx):
    """The served workload's pool (24 Erdos-Renyi matrices, n log-uniform in
    [2k, 256k], mean nnz/row U[4,32]; 2 R-MAT scale-18) under a 64 GiB
    budget.  Every (matrix, op) in the request list runs alone first:
    sampled rows (16 highest-degree + 48 random) bit-exact against the
    oracle, and its checksums recorded.  Then 48 Poisson requests (80%
    Erdos-Renyi / 20% R-MAT, A^2 / A*A^T 50/50) arrive at 8 req/s from a
    submitter thread while the main thread takes each result as it
    completes, checks that its checksums equal the request's alone, and
    releases it."""
    mats, rng = serve_pool()
    n_er, n_req = 24, 48
    kinds = rng.uniform(size=n_req) < 0.8
    pick = [int(rng.integers(0, n_er)) if k else n_er + int(rng.integers(0, 2)) for k in kinds]
    ops = [int(x) for x in rng.integers(0, 2, n_req)]
    srv = ts.Server(0, ts.server_config(budget_bytes=64 << 30))
    try:
        t0 = time.perf_counter()
        hs = [srv.register(a) for a in mats]
        log(f"[serve] registered {len(mats)} matrices in {time.perf_counter() - t0:.1f} s, "
            f"resident {srv.stats().resident_bytes / 2**30:.2f} GiB")
        alone = {}
        for key in sorted(set(zip(pick, ops))):
            i, op = key
            a = mats[i]
            rid = srv.submit(hs[i], op)
            info = srv.wait(rid, 600000)
            assert info.status == 0 and info.fused_with == 1
            cs = srv.result_checksum(rid)
            assert cs.nnz == info.nnzC == srv.size(hs[i], op)["nnzC"]
            rows, _ = sample_rows(a, 48, seed=i)
            got = srv.result_rows(rid, rows)
            srv.release(rid)
            b = a if op == ts.OP_A2 else transpose_host(a)
            check_rows(got, oracle.spgemm_rowlist(a, b, rows), f"matrix {i} op {op} alone")
            if op == ts.OP_AAT:
                assert cs.entries == cs.entries_t, f"matrix {i}: A*A^T not bit-symmetric"
            alone[key] = cs.key()
        log(f"[serve] {len(alone)} distinct products alone: sampled rows bit-exact")
        arrival = np.cumsum(rng.exponential(1000.0 / 8.0, n_req))
        ids = [None] * n_req
        ready = [threading.Event() for _ in range(n_req)]
        err = []

        def submitter():
            try:
                start = time.perf_counter()
                for k in range(n_req):
                    delay = arrival[k] / 1e3 - (time.perf_counter() - start)
                    if delay > 0:
                        time.sleep(delay)
                    ids[k] = srv.submit(hs[pick[k]], ops[k])
                    ready[k].set()
            except Exception as e:  # surfaced by the main thread
                err.append(e)
                for ev in ready:
                    ev.set()

        th = threading.Thread(target=submitter)
        th.start()
        fused = checked = 0
        pending = set(range(n_req))
        deadline = time.perf_counter() + 300
        try:
            # results are taken in completion order and released at once:
            # holding a finished result while waiting for an earlier request
            # could keep that request from being admitted
            while pending:
                assert time.perf_counter() < deadline, f"{len(pending)} requests not done"
                assert not err, err
                for k in sorted(pending):
                    if not ready[k].is_set():
                        continue
                    info = srv.poll(ids[k], 5.0)
                    if info is None:
                        continue
                    assert info.status == 0, (k, info.status)
                    fused += info.fused_with > 1
                    cs = srv.result_checksum(ids[k])
                    assert cs.key() == alone[(pick[k], ops[k])], (k, pick[k], ops[k], info.fused_with)
                    srv.release(ids[k])
                    pending.discard(k)
                    checked += 1
                    if checked % 8 == 0:
                        log(f"[serve] {checked}/{n_req} mixed results equal to alone")
                if pending and not any(ready[k].is_set() for k in pending):
                    time.sleep(0.01)
        finally:
            th.join()
        st = srv.stats()
        log(f"[serve] mix: {fused} fused, peak footprint {st.peak_footprint_bytes / 2**30:.1f} GiB")
        assert st.peak_footprint_bytes <= 64 << 30
        assert st.reserved_bytes == 0
    finally:
        srv.close()
