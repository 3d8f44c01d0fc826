"""CPU-side checks of the product library: it loads, exports every function
declared in include/*.h, and its host-side logic (generators, Matrix Market
reader, value modes) is right.  No device calls."""
import os
import re

import numpy as np
import pytest

import tilespgemm as ts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"typedef\s+struct\s*\w*\s*\{.*?\}\s*\w+\s*;", "", src, flags=re.S)
        src = re.sub(r"enum\s*\{.*?\}\s*;", "", src, flags=re.S)
        src = re.sub(r"typedef[^;]*\(\s*\*\s*\w+\s*\)[^;]*;", "", src)  # function-pointer types
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(ts.lib, n)]
    assert not missing, missing


def test_build_info():
    assert "gfx950" in ts.build_info()


def _check_sorted_unique(a):
    for i in range(a.m):
        row = a.colidx[a.rowptr[i]:a.rowptr[i + 1]]
        assert np.all(np.diff(row) > 0)
    assert a.colidx.min(initial=0) >= 0 and a.colidx.max(initial=0) < a.n


def test_gen_stencil5():
    a = ts.gen_stencil5(64, 64)
    assert (a.m, a.nnz) == (4096, 20224)
    _check_sorted_unique(a)
    d = a.to_scipy().diagonal()
    assert np.all(d == 4.0)
    assert np.all(np.isin(a.val, [4.0, -1.0]))


@pytest.mark.parametrize("N,periodic,nnz", [(6, True, 27 * 216), (6, False, (3 * 6 - 2) ** 3), (32, True, 27 * 32 ** 3)])
def test_gen_stencil27(N, periodic, nnz):
    a = ts.gen_stencil27(N, N, N, periodic=periodic, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    assert a.nnz == nnz
    _check_sorted_unique(a)
    assert a.val.min() >= -1.0 and a.val.max() < 1.0
    b = ts.gen_stencil27(N, N, N, periodic=periodic, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    assert np.array_equal(a.val, b.val)
    c = ts.gen_stencil27(N, N, N, periodic=periodic, valmode=ts.VAL_UNIFORM_PM1, seed=2)
    assert not np.array_equal(a.val, c.val)
    assert abs(a.val.mean()) < 0.1


def test_gen_er_banded_rmat():
    a = ts.gen_er(5000, 12.0, seed=3)
    _check_sorted_unique(a)
    assert 11.0 < a.nnz / a.m < 13.0
    b = ts.gen_banded(1000, 7)
    _check_sorted_unique(b)
    assert b.nnz == sum(min(999, i + 7) - max(0, i - 7) + 1 for i in range(1000))
    r = ts.gen_rmat(10, 16, seed=1)
    _check_sorted_unique(r)
    s = r.to_scipy()
    s.data[:] = 1.0
    assert (s != s.T).nnz == 0  # symmetrised pattern (values are per-entry draws)
    assert r.val.min() > 0.0 and r.val.max() <= 1.0
    deg = np.diff(r.rowptr)
    assert deg.max() > 20 * max(1.0, deg.mean()) / 4  # skewed degrees


def test_value_modes():
    a = ts.gen_stencil5(8, 8)
    ts.set_values(a, ts.VAL_MOD10)
    assert np.array_equal(a.val, np.arange(a.nnz) % 10)
    ts.set_values(a, ts.VAL_UNIFORM_01, seed=5)
    assert a.val.min() > 0 and a.val.max() <= 1


def test_read_mtx_symmetric_pattern(tmp_path):
    p = tmp_path / "a.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% c\n4 4 4\n1 1\n2 1\n4 3\n3 2\n")
    a, sym = ts.read_mtx(str(p))
    assert sym
    d = a.to_scipy().toarray()
    ref = np.zeros((4, 4))
    for r, c in [(0, 0), (1, 0), (0, 1), (3, 2), (2, 3), (2, 1), (1, 2)]:
        ref[r, c] = 1.0
    assert np.array_equal(d, ref)
    _check_sorted_unique(a)


def test_read_mtx_general_real(tmp_path):
    p = tmp_path / "b.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 5 3\n3 5 2.5\n1 2 -1\n1 1 4\n")
    a, sym = ts.read_mtx(str(p))
    assert not sym and (a.m, a.n, a.nnz) == (3, 5, 3)
    assert list(a.colidx) == [0, 1, 4] and list(a.val) == [4.0, -1.0, 2.5]
