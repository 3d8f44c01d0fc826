"""The only known answers the reference holds for this path: cant.mtx A^2
counts (tests/golden/cant_a2_results.json, from the reference's
data/results_tile.csv:1).  cant.mtx itself is not shipped; set TSG_CANT_MTX
to its path to run these checks."""
import json
import os

import pytest

import tilespgemm as ts

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "cant_a2_results.json")))
MTX = os.environ.get("TSG_CANT_MTX", "")

needs_file = pytest.mark.skipif(not os.path.isfile(MTX), reason="cant.mtx not supplied (TSG_CANT_MTX)")


def test_fixture_is_consistent():
    assert GOLD["m"] == GOLD["n"] == 62451
    assert abs(GOLD["nnzCub"] / GOLD["nnzC"] - GOLD["compression_rate"]) < 1e-5


def test_read_mtx_returns_matrix_and_flag():
    """The reader the cant.mtx checks use, on a committed fixture."""
    a, sym = ts.read_mtx(os.path.join(HERE, "golden", "stencil5_8x8_sym.mtx"))
    assert sym and (a.m, a.n, a.nnz) == (64, 64, 288)


@needs_file
def test_cant_read_counts():
    a, _ = ts.read_mtx(MTX)
    assert (a.m, a.n, a.nnz) == (GOLD["m"], GOLD["n"], GOLD["nnzA"])


@needs_file
@pytest.mark.gpu
def test_cant_a2_counts(ctx):
    a, _ = ts.read_mtx(MTX)
    a = ts.set_values(a, ts.VAL_MOD10)
    da = ctx.upload(a)
    assert ctx.nnzcub(da, da) == GOLD["nnzCub"]
    ta = ctx.csr2tile(da)
    ctx.build_csc(ta)
    tc, st = ctx.spgemm(ta, ta)
    assert st.nnzC == GOLD["nnzC"]
    for o in (tc, ta, da):
        ctx.free(o)
