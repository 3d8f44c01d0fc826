"""The reference driver flow (src/main.cu:15-638) run in-process through
tsg_driver_main, and the reorder half of the boundary through the C ABI.

Every driver run writes C (after restore, in the original order) with -o;
it is compared with the oracle's product of the ORIGINAL operands as the
reference driver prepares them: mmio_allinone, values i % 10
(src/main.cu:191-197), rows sorted, B = A, A^T or the -b matrix.  With
i % 10 values every product and sum is an exact integer, so C is
bit-exact whatever order a permutation gives the inner sums."""
import os

import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import log

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
SYM = os.path.join(G, "stencil5_8x8_sym.mtx")
SQ = os.path.join(G, "er_96_general.mtx")
RECT = os.path.join(G, "er_80x96_general.mtx")


def write_mtx(path, csr, symmetric=False):
    rows = np.repeat(np.arange(csr.m), np.diff(csr.rowptr))
    keep = rows >= csr.colidx if symmetric else np.ones(csr.nnz, bool)
    r, c, v = rows[keep], csr.colidx[keep], csr.val[keep]
    order = np.lexsort((r, c))  # column-major, as SuiteSparse stores files
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {'symmetric' if symmetric else 'general'}\n")
        f.write(f"{csr.m} {csr.n} {len(order)}\n")
        for k in order:
            f.write(f"{int(r[k]) + 1} {int(c[k]) + 1} {float(v[k])!r}\n")


def load_like_driver(spec, keep=False):
    """The operand as tsg_driver_main prepares it."""
    if spec.startswith("gen:"):
        p = spec.split(":")
        if p[1] == "5pt":
            a = ts.gen_stencil5(int(p[2]), int(p[3]))
        elif p[1] in ("27p", "27d"):
            n = int(p[2])
            a = ts.gen_stencil27(n, n, n, periodic=p[1] == "27p", valmode=ts.VAL_UNIFORM_PM1, seed=1)
        elif p[1] == "er":
            a = ts.gen_er(int(p[2]), float(p[3]), seed=int(p[4]))
        elif p[1] == "rmat":
            a = ts.gen_rmat(int(p[2]), int(p[3]), seed=int(p[4]))
        else:
            raise ValueError(spec)
        if not keep:
            ts.set_values(a, ts.VAL_MOD10)
        return a
    st, raw, _ = ts.mmio_allinone(spec)
    assert st == 0
    if not keep:
        raw.val[:] = np.arange(raw.nnz) % 10
    return ts.canonicalize(raw)


def expected_c(a_spec, aat, b_spec=None, keep=False):
    a = load_like_driver(a_spec, keep)
    if b_spec:
        b = load_like_driver(b_spec, keep)
    elif aat:
        tp, tc, tv = oracle.transpose(a)
        b = ts.Csr(a.n, a.m, tp, tc, tv)
    else:
        b = a
    return a, b, oracle.spgemm(a, b)


def run_driver(tmp_path, a_spec, aat, b_spec=None, reorder=None, keep=False):
    out = str(tmp_path / "c.bin")
    args = ["-d", 0, "-aat", 1 if aat else 0, a_spec]
    if b_spec:
        args += ["-b", b_spec]
    if reorder:
        args += ["-reorder", 1, "-reorder_algo", reorder]
    if keep:
        args += ["-keepvalues"]
    args += ["-o", out]
    log(f"[driver] {' '.join(str(x) for x in args)}")
    rc = ts.driver(args)
    assert rc == 0, f"driver exit {rc}"
    return ts.read_csr_bin(out)


def assert_c(c, ref, exact=True, a=None, b=None):
    rpc, cc, vc = ref
    assert np.array_equal(c.rowptr.astype(np.int64), rpc)
    assert np.array_equal(c.colidx, cc)
    if exact:
        assert np.array_equal(c.val, vc)
    else:
        aa, bb = a.to_scipy().copy(), b.to_scipy().copy()
        aa.data, bb.data = np.abs(aa.data), np.abs(bb.data)
        bound = (aa @ bb).tocsr()
        rows = np.repeat(np.arange(c.m), np.diff(c.rowptr))
        scale = np.asarray(bound[rows, c.colidx]).ravel()
        assert np.all(np.abs(c.val - vc) <= 1e-12 * scale)


ORDERS = [None, "rcm", "degree", "random", "original", "tasa"]


@pytest.mark.parametrize("reorder", ORDERS)
def test_driver_aa_symmetric_mtx(tmp_path, reorder):
    c = run_driver(tmp_path, SYM, False, reorder=reorder)
    assert_c(c, expected_c(SYM, False)[2])


@pytest.mark.parametrize("reorder", ORDERS[:5])
def test_driver_aat_general_mtx(tmp_path, reorder):
    c = run_driver(tmp_path, SQ, True, reorder=reorder)
    assert_c(c, expected_c(SQ, True)[2])


def test_driver_aat_rectangular(tmp_path):
    c = run_driver(tmp_path, RECT, True)
    assert (c.m, c.n) == (80, 80)
    assert_c(c, expected_c(RECT, True)[2])


@pytest.mark.parametrize("reorder", ORDERS)
def test_driver_ab(tmp_path, reorder):
    bpath = str(tmp_path / "b.mtx")
    write_mtx(bpath, ts.gen_er(96, 5.0, seed=7))
    c = run_driver(tmp_path, SQ, False, b_spec=bpath, reorder=reorder)
    assert_c(c, expected_c(SQ, False, bpath)[2])


def test_driver_ab_rectangular(tmp_path):
    bpath = str(tmp_path / "b.mtx")
    write_mtx(bpath, ts.gen_er(96, 3.0, seed=8))
    c = run_driver(tmp_path, RECT, False, b_spec=bpath)
    assert (c.m, c.n) == (80, 96)
    assert_c(c, expected_c(RECT, False, bpath)[2])


@pytest.mark.parametrize("spec,aat,reorder", [
    ("gen:5pt:40:30", False, None),
    ("gen:27p:12", False, "rcm"),
    ("gen:27d:10", False, "degree"),
    ("gen:er:3000:6:4", True, "random"),
    ("gen:rmat:11:8:5", True, "degree"),
    ("gen:rmat:10:16:2", False, "rcm"),
])
def test_driver_generated(tmp_path, spec, aat, reorder):
    c = run_driver(tmp_path, spec, aat, reorder=reorder)
    assert_c(c, expected_c(spec, aat)[2])


def test_driver_keepvalues(tmp_path):
    """File values kept: no permutation -> bit-exact; under RCM the inner
    sums change order, so values agree within 1e-12 of sum |a_ik b_kj|."""
    bpath = str(tmp_path / "sym.mtx")
    write_mtx(bpath, ts.gen_er(500, 6.0, seed=3), symmetric=False)
    a, b, ref = expected_c(bpath, True, keep=True)
    assert_c(run_driver(tmp_path, bpath, True, keep=True), ref)
    assert_c(run_driver(tmp_path, bpath, True, reorder="rcm", keep=True), ref, exact=False, a=a, b=b)


# ------------------------------------------------------------------ reorder boundary
def test_external_plan_degree_on_symmetric_mtx_matches_file_degrees(tmp_path):
    """build_external_reorder_plan(AA, "degree") on a symmetric file, with
    the CSR mmio_allinone builds: ids by (-deg, id), deg counted from the
    file's lines as order_utils.py:30-45 does (stored entry + mirror)."""
    rpath = str(tmp_path / "r.mtx")
    write_mtx(rpath, ts.gen_rmat(9, 8, seed=4), symmetric=True)
    lines = open(rpath).read().splitlines()
    m, n, nz = lines[1].split()
    off = next(ln for ln in lines[2:] if ln.split()[0] != ln.split()[1])
    lines[1] = f"{m} {n} {int(nz) + 1}"
    with open(rpath, "w") as f:  # one duplicate line: the file's degrees count it twice
        f.write("\n".join(lines + [off]) + "\n")
    for path in (SYM, rpath):
        text = open(path).read()
        lines = text.splitlines()
        body = [ln for ln in lines[1:] if ln.strip() and not ln.startswith("%")]
        n = int(body[0].split()[0])
        deg = [0] * n
        for ln in body[1:]:
            r, c = (int(x) - 1 for x in ln.split()[:2])
            deg[r] += 1
            if r != c:
                deg[c] += 1
        ref = sorted(range(n), key=lambda i: (-deg[i], i))
        st, csr, sym = ts.mmio_allinone(path)
        assert st == 0 and sym
        hm = ts.HostSMatrix(csr, symmetric=True)
        st, info, _ = ts.reorder_plan(ts.PROBLEM_AA, "degree", hm)
        assert st == 0
        for ax in ("row_perm", "inner_perm", "col_perm"):
            assert list(info[ax][0]) == ref
        assert info["shared_symmetric_csr"] == 1 and info["mode"] == "degree"


@pytest.mark.parametrize("kind", [ts.PROBLEM_AA, ts.PROBLEM_AAT, ts.PROBLEM_AB])
@pytest.mark.parametrize("algo", ["rcm", "degree", "random", "original"])
def test_plan_axis_sharing_and_estimates(kind, algo):
    """AA/AAT: one order from A on all three axes; AB: A's order for rows
    and the inner axis, B's own order for columns (src/external_reorder.h:
    226-304).  The tile estimates are exact device counts."""
    a = ts.set_values(ts.gen_er(700, 5.0, seed=11), ts.VAL_MOD10)
    b = ts.set_values(ts.gen_banded(700, 3), ts.VAL_MOD10)
    ha, hb = ts.HostSMatrix(a), ts.HostSMatrix(b)
    other = None if kind == ts.PROBLEM_AA else (ha if kind == ts.PROBLEM_AAT else hb)
    st, info, _ = ts.reorder_plan(kind, algo, ha, other)
    assert st == 0
    row, inner, col = (info[k][0] for k in ("row_perm", "inner_perm", "col_perm"))
    assert np.array_equal(np.sort(row), np.arange(700))
    assert np.array_equal(row, inner)
    st_b, info_b, _ = ts.reorder_plan(ts.PROBLEM_AA, algo, hb)
    if kind == ts.PROBLEM_AB and algo != "original":
        assert np.array_equal(col, info_b["row_perm"][0])
        if algo != "random":  # a seeded shuffle depends only on n and the seed
            assert not np.array_equal(col, row)
    else:
        assert np.array_equal(col, row)
    assert info["input_tiles_A_before"] > 0 and info["estimated_c_tiles_before"] > 0
    assert info["estimated_c_tiles_after"] > 0


def test_apply_restore_ownership_and_roundtrip():
    """apply mallocs the output CSR; restore frees the matrix's arrays and
    installs new ones holding the original matrix (src/reorder.h:2506-2551)."""
    a = ts.set_values(ts.gen_er(900, 7.0, seed=12), ts.VAL_UNIFORM_PM1)
    ha = ts.HostSMatrix(a)
    st, info, _ = ts.reorder_plan(ts.PROBLEM_AA, "rcm", ha)
    assert st == 0
    perm = ts.AxisPermutation()
    old, new = info["row_perm"]
    perm.n = a.m
    perm.old_id = old.ctypes.data_as(ts.C.POINTER(ts.C.c_int))
    perm.new_id = new.ctypes.data_as(ts.C.POINTER(ts.C.c_int))
    out = ts.SMatrix()
    assert ts.lib.apply_bipermutation_csr(ha.ref, ts.C.byref(perm), ts.C.byref(perm), ts.C.byref(out)) == 0
    assert out.isSymmetric == 0 and out.nnz == a.nnz
    got = ts.Csr(out.m, out.n, np.ctypeslib.as_array(out.rowpointer, shape=(a.m + 1,)),
                 np.ctypeslib.as_array(out.columnindex, shape=(a.nnz,)),
                 np.ctypeslib.as_array(out.value, shape=(a.nnz,)))
    ref = oracle.permute_csr(a, old, new)
    assert np.array_equal(got.rowptr, ref[0]) and np.array_equal(got.colidx, ref[1])
    assert np.array_equal(got.val, ref[2])
    before = [ts.C.cast(p, ts.C.c_void_p).value for p in (out.rowpointer, out.columnindex, out.value)]
    assert ts.lib.restore_bipermutation_csr(ts.C.byref(out), ts.C.byref(perm), ts.C.byref(perm)) == 0
    after = [ts.C.cast(p, ts.C.c_void_p).value for p in (out.rowpointer, out.columnindex, out.value)]
    # new arrays are allocated while the old ones are still live, so every address changes
    assert all(x != y for x, y in zip(before, after))
    back = ts.Csr(out.m, out.n, np.ctypeslib.as_array(out.rowpointer, shape=(a.m + 1,)).copy(),
                  np.ctypeslib.as_array(out.columnindex, shape=(a.nnz,)).copy(),
                  np.ctypeslib.as_array(out.value, shape=(a.nnz,)).copy())
    assert np.array_equal(back.rowptr, a.rowptr) and np.array_equal(back.colidx, a.colidx)
    assert np.array_equal(back.val, a.val)
    h = ts.HCsr(out.m, out.n, out.nnz, out.rowpointer, out.columnindex, out.value)
    ts.lib.tsg_hcsr_free(ts.C.byref(h))
