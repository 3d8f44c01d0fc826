"""The headline workload at full size: C = A^2 for the 27-point stencil on a
128^3 grid (n = 2,097,152, 56.6M nonzeros), periodic and Dirichlet.

The full oracle does not finish in seconds here, so parity is checked by
size-independent properties: closed-form counts (tiles of A, step-1 C tiles,
structurally empty tiles, nnz(C) = 125 n), every row's structure against
the stencil geometry, and bit-exact rows against the oracle on a sample of
rows spread over the matrix (first/last tile-rows, wrap-around planes,
random interior rows)."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import log

pytestmark = pytest.mark.gpu

N = 128


def sample_rows(n, seed):
    rng = np.random.default_rng(seed)
    fixed = list(range(0, 48)) + list(range(n - 48, n)) + list(range(N * N * 64, N * N * 64 + 32))
    return sorted(set(fixed) | set(rng.choice(n, 3000, replace=False).tolist()))


@pytest.mark.parametrize("periodic", [True, False], ids=["27p", "27d"])
def test_27pt_128(ctx, periodic):
    log(f"[fullsize] 27{'p' if periodic else 'd'} 128^3: generate")
    a = ts.gen_stencil27(N, N, N, periodic=periodic, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    n = a.m
    da = ctx.upload(a)
    ta = ctx.csr2tile(da)
    ctx.build_csc(ta)
    log("[fullsize] spgemm")
    tc, st = ctx.spgemm(ta, ta)
    log(f"[fullsize] stats {st.as_dict()}")
    if periodic:
        assert ta.numtile == 3_538_944
        assert st.numtileC == 16_384_000 and st.tiles_empty == 6_553_600
        assert st.nnzC == 125 * n == 262_144_000
    else:
        assert ta.numtile == 3_210_328
        assert st.numtileC == 13_666_504 and st.tiles_empty == 4_823_472
        assert st.nnzC == 254_840_104
    ctx.free(ta)
    log("[fullsize] tile2csr + download")
    dc = ctx.tile2csr(tc)
    ctx.free(tc)
    c = ctx.download(dc)
    ctx.free(dc)
    ctx.free(da)
    # structure of every row: the 5x5x5 neighbourhood (clipped when Dirichlet)
    if periodic:
        assert np.all(np.diff(c.rowptr) == 125)
    rows = sample_rows(n, 3)
    log(f"[fullsize] oracle on {len(rows)} rows")
    for r in rows:
        rp, cc, vv = oracle.spgemm_rows(a, a, r, r + 1)
        s, e = c.rowptr[r], c.rowptr[r + 1]
        assert e - s == rp[1], r
        assert np.array_equal(c.colidx[s:e], cc), r
        assert np.array_equal(c.val[s:e].view(np.int64), vv.view(np.int64)), r
        x, y, z = r % N, (r // N) % N, r // (N * N)
        exp = set()
        for dz in range(-2, 3):
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    xx, yy, zz = x + dx, y + dy, z + dz
                    if periodic:
                        xx, yy, zz = xx % N, yy % N, zz % N
                    elif not (0 <= xx < N and 0 <= yy < N and 0 <= zz < N):
                        continue
                    exp.add(xx + N * (yy + N * zz))
        assert list(cc) == sorted(exp), r
