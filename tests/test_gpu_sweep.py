"""Seeded randomized sweep: Erdos-Renyi, banded, R-MAT and rectangular
products, A^2 and A*A^T, each through every kernel path of spgemm_check
(row-wise, tile-parallel, both step-1 windows, element-wise with and without
A's row index, heavy rows with single and split work items), every array of
C compared with the oracle, values bit-exact."""
import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import spgemm_check

pytestmark = pytest.mark.gpu


def _t(c):
    tp, tc, tv = oracle.transpose(c)
    return ts.Csr(c.n, c.m, tp, tc, tv)


def _case(rng, kind):
    seed = int(rng.integers(1, 1 << 30))
    if kind == 0:
        return ts.gen_er(int(rng.integers(100, 30000)), float(rng.uniform(1, 24)), seed=seed)
    if kind == 1:
        return ts.gen_banded(int(rng.integers(50, 20000)), int(rng.integers(1, 60)), seed=seed)
    if kind == 2:
        return ts.gen_rmat(int(rng.integers(8, 14)), int(rng.integers(4, 20)), seed=seed)
    # rectangular m x n with random entries (C = A A^T)
    m, n = int(rng.integers(100, 8000)), int(rng.integers(100, 40000))
    nnz = int(m * rng.uniform(1, 20))
    key = np.unique(rng.integers(0, m, nnz).astype(np.int64) * n + rng.integers(0, n, nnz))
    r, c = (key // n).astype(np.int32), (key % n).astype(np.int32)
    rp = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=m))]).astype(np.int32)
    return ts.Csr(m, n, rp, c, rng.uniform(-1, 1, len(c)))


@pytest.mark.parametrize("block", range(4))
def test_random_sweep(ctx, block):
    rng = np.random.default_rng(1000 + block)
    for it in range(6):
        kind = (it + block) % 4
        a = _case(rng, kind)
        b = _t(a) if (kind == 3 or rng.uniform() < 0.5) else None
        spgemm_check(ctx, a, b, label=f"sweep{block}.{it} kind {kind} {'AAT' if b is not None else 'A2'}")
