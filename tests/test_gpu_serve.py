"""Serving on the GPU (tsg_server): continuous batching of A^2 / AA^T
requests on resident matrices.  Every result must be bit-identical when
the request runs alone and when it runs inside a Poisson mix (fused with
others, overlapping on several streams), and equal to the plain
tsg_spgemm product; admission must respect the budget."""
import numpy as np
import pytest

import tilespgemm as ts
from tsg_check import device_tiles, log

pytestmark = pytest.mark.gpu


def plain_csr(ctx, a, op):
    da, ta = device_tiles(ctx, a)
    tb, dat = ta, None
    if op == ts.OP_AAT:
        dat = ctx.transpose(da)
        tb = ctx.csr2tile(dat)
        ctx.build_csc(tb)
    tc, _ = ctx.spgemm(ta, tb)
    dc = ctx.tile2csr(tc)
    c = ctx.download(dc)
    for o in (dc, tc, ta, da) + ((tb, dat) if dat is not None else ()):
        ctx.free(o)
    return c


def same_csr(x, y):
    return (x.m == y.m and x.n == y.n and np.array_equal(x.rowptr, y.rowptr) and np.array_equal(x.colidx, y.colidx)
            and np.array_equal(x.val.view(np.int64), y.val.view(np.int64)))


def pool():
    mats = [ts.gen_er(n, d, seed=i + 1) for i, (n, d) in enumerate([(2000, 4.0), (3500, 20.0), (7000, 9.0),
                                                                      (1200, 31.0), (15000, 6.0)])]
    mats.append(ts.gen_banded(5000, 17, seed=9))
    mats.append(ts.gen_rmat(12, 16, seed=4))
    rect = ts.gen_er(3000, 8.0, seed=12)
    mats.append(ts.Csr(1000, rect.n, rect.rowptr[:1001], rect.colidx[:rect.rowptr[1000]],
                       rect.val[:rect.rowptr[1000]]))
    return mats


def test_serve_alone_vs_mix(ctx):
    mats = pool()
    srv = ts.Server(0, ts.server_config(budget_bytes=8 << 30, slots=4, max_batch=16))
    try:
        hs = [srv.register(a) for a in mats]
        ops_of = {h: ([ts.OP_A2, ts.OP_AAT] if mats[h].m == mats[h].n else [ts.OP_AAT]) for h in hs}
        alone = {}
        for h in hs:
            for op in ops_of[h]:
                sz = srv.size(h, op)
                rid = srv.submit(h, op)
                info = srv.wait(rid, 60000)
                assert info.fused_with == 1
                c = srv.result_csr(rid)
                srv.release(rid)
                assert c.nnz == sz["nnzC"]
                assert same_csr(c, plain_csr(ctx, mats[h], op)), (h, op)
                alone[(h, op)] = c
        # Poisson mix, fast enough that requests queue, fuse and overlap
        rng = np.random.default_rng(3)
        keys = list(alone)
        pick = rng.integers(0, len(keys), 60)
        handles = [keys[k][0] for k in pick]
        ops = [keys[k][1] for k in pick]
        arrival = np.cumsum(rng.exponential(0.3, 60)).tolist()
        ids, infos = srv.replay(handles, ops, arrival, keep=True)
        fused = sum(1 for i in infos if i.fused_with > 1)
        st = srv.stats()
        log(f"[serve] mix: {fused} of 60 requests fused; fused launches {st.fused_launches}, "
            f"single {st.single_launches}; peak reserved {st.peak_reserved_bytes}")
        assert fused > 0
        for rid, h, op, info in zip(ids, handles, ops, infos):
            assert info.status == 0 and info.t_done_ms >= info.t_admit_ms >= info.t_arrival_ms
            c = srv.result_csr(rid)
            assert same_csr(c, alone[(h, op)]), (rid, h, op, info.fused_with)
            srv.release(rid)
        assert srv.stats().reserved_bytes == 0
    finally:
        srv.close()


def test_serve_chunked_request(ctx):
    """A product cut into many tile-row chunks (small chunk cap) equals the
    one-shot product, alone and fused."""
    a = ts.gen_rmat(12, 16, seed=8)
    srv = ts.Server(0, ts.server_config(budget_bytes=4 << 30, slots=2, chunk_products=200000))
    try:
        h = srv.register(a)
        sz = srv.size(h, ts.OP_AAT)
        log(f"[serve] chunks {sz['chunks']}")
        assert sz["chunks"] > 3
        rid = srv.submit(h, ts.OP_AAT)
        srv.wait(rid, 60000)
        c = srv.result_csr(rid)
        srv.release(rid)
        assert same_csr(c, plain_csr(ctx, a, ts.OP_AAT))
    finally:
        srv.close()


def test_serve_budget(ctx):
    a = ts.gen_er(4000, 16.0, seed=5)
    b = ts.gen_banded(3000, 9, seed=6)
    srv = ts.Server(0, ts.server_config(budget_bytes=256 << 20, slots=3))
    try:
        ha = srv.register(a)
        hb = srv.register(b)
        need = srv.size(ha, ts.OP_A2)["bytes"]
        st = srv.stats()
        log(f"[serve] resident {st.resident_bytes}, A^2 needs {need}")
        # at most floor((budget - resident) / need) such requests may be in flight together
        ids, infos = srv.replay([ha] * 12 + [hb] * 4, [ts.OP_A2] * 16, [0.0] * 16, keep=False)
        assert all(i.status == 0 for i in infos)
        assert srv.stats().peak_reserved_bytes + st.resident_bytes <= 256 << 20
        # with the workers' cached blocks counted too
        assert srv.stats().peak_footprint_bytes <= 256 << 20
        big = ts.gen_er(60000, 30.0, seed=7)
        with pytest.raises(ts.TsgError):
            hbig = srv.register(big)
            srv.submit(hbig, ts.OP_A2)
    finally:
        srv.close()
