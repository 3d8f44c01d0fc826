"""CPU checks of the reference-facing host layer: the Matrix Market loader
with the reference's semantics (src/mmio_highlevel.h:593-759), the order
file contract (src/external_reorder.h:53-98), the argument checks and axis
sharing of the reorder plans (src/external_reorder.h:173-319,
src/reorder.h:2407-2551) and the order_gen adapter contract
(reorder_algorithm/README.md).  Nothing here needs a device."""
import os
import stat

import numpy as np
import pytest
import scipy.sparse as sp

import tilespgemm as ts

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


# ------------------------------------------------------------------ Matrix Market
def parse_entries(text):
    """(header words, m, n, [(r, c, value or None)]) straight from the text."""
    lines = text.splitlines()
    head = lines[0].lower().split()
    body = [ln for ln in lines[1:] if ln.strip() and not ln.startswith("%")]
    m, n, nz = (int(x) for x in body[0].split()[:3])
    ent = []
    for ln in body[1:1 + nz]:
        p = ln.split()
        ent.append((int(p[0]) - 1, int(p[1]) - 1, float(p[2]) if len(p) > 2 else None))
    return head, m, n, ent


def reference_rows(text):
    """Rows the reference loader builds: entries in file order, an
    off-diagonal entry of a symmetric/hermitian file also appended to the
    mirrored row at the same moment; pattern -> 1.0."""
    head, m, n, ent = parse_entries(text)
    sym = head[4] in ("symmetric", "hermitian")
    rows = [[] for _ in range(m)]
    for r, c, v in ent:
        v = 1.0 if v is None else v
        rows[r].append((c, v))
        if sym and r != c:
            rows[c].append((r, v))
    return m, n, rows, sym


def rows_of(csr):
    return [list(zip(csr.colidx[csr.rowptr[i]:csr.rowptr[i + 1]].tolist(),
                     csr.val[csr.rowptr[i]:csr.rowptr[i + 1]].tolist())) for i in range(csr.m)]


CASES = {
    "general_unsorted_dups": "%%MatrixMarket matrix coordinate real general\n% c\n3 4 6\n"
                             "2 3 1.5\n1 4 -2\n2 1 3\n1 1 7e-1\n2 3 0.25\n3 2 -1e-3\n",
    "symmetric_lower": "%%MatrixMarket matrix coordinate real symmetric\n4 4 6\n"
                       "1 1 2\n3 1 -1\n4 1 5\n2 2 3\n4 2 -2\n4 4 9\n",
    "symmetric_upper_mixed": "%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n1 3 1\n2 1 2\n3 3 3\n1 2 4\n",
    "pattern_symmetric": "%%MatrixMarket matrix coordinate pattern symmetric\n% c\n4 4 4\n1 1\n2 1\n4 3\n3 2\n",
    "integer_general": "%%MatrixMarket Matrix Coordinate Integer General\n2 3 3\n2 3 7\n1 2 -4\n2 1 1\n",
    "complex_hermitian": "%%MatrixMarket matrix coordinate complex hermitian\n2 2 2\n1 1 1.5 0\n2 1 -2 3\n",
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_mmio_allinone_reference_semantics(tmp_path, name):
    text = CASES[name]
    p = tmp_path / f"{name}.mtx"
    p.write_text(text)
    st, a, sym = ts.mmio_allinone(str(p))
    assert st == 0
    m, n, rows, rsym = reference_rows(text)
    assert (a.m, a.n, sym) == (m, n, rsym)
    assert a.nnz == sum(len(r) for r in rows)
    assert rows_of(a) == rows


def test_mmio_skew_symmetric_is_not_expanded(tmp_path):
    p = tmp_path / "k.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1\n3 2 -4\n")
    st, a, sym = ts.mmio_allinone(str(p))
    assert st == 0 and not sym and a.nnz == 2
    assert rows_of(a) == [[], [(0, 1.0)], [(1, -4.0)]]


@pytest.mark.parametrize("text,code", [
    (None, -1),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -2),
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix coordinate real weird\n1 1 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n3 3\n1 1 1\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1\n4 1 2\n", -5),
    ("%%MatrixMarket matrix coordinate real general\n3 3 3\n1 1 1\n2 2 2\n", -5),
    ("%%MatrixMarket matrix coordinate real symmetric\n3 4 1\n1 4 1\n", -5),
])
def test_mmio_status_codes(tmp_path, text, code):
    p = tmp_path / "bad.mtx"
    if text is not None:
        p.write_text(text)
    st, a, _ = ts.mmio_allinone(str(p))
    assert st == code and a is None


@pytest.mark.parametrize("name", sorted(CASES))
def test_read_mtx_is_canonical_form(tmp_path, name):
    """tsg_read_mtx = the reference CSR with rows sorted and duplicates
    summed; checked against SciPy's COO -> CSR conversion of the same
    expanded entry list."""
    text = CASES[name]
    p = tmp_path / f"{name}.mtx"
    p.write_text(text)
    a, sym = ts.read_mtx(str(p))
    m, n, rows, _ = reference_rows(text)
    r = [i for i in range(m) for _ in rows[i]]
    c = [cv[0] for i in range(m) for cv in rows[i]]
    v = [cv[1] for i in range(m) for cv in rows[i]]
    ref = sp.coo_matrix((v, (r, c)), shape=(m, n)).tocsr()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(a.rowptr, ref.indptr) and np.array_equal(a.colidx, ref.indices)
    assert np.array_equal(a.val, ref.data)


@pytest.mark.parametrize("fn", ["stencil5_8x8_sym.mtx", "er_96_general.mtx", "er_80x96_general.mtx"])
def test_committed_mtx_fixtures(fn):
    """The committed fixtures through both readers.  Column-major files give
    the reference loader sorted rows already, so the canonical form is the
    reference CSR itself; the row-shuffled file needs the sort."""
    path = os.path.join(GOLDEN, fn)
    text = open(path).read()
    st, raw, sym = ts.mmio_allinone(path)
    assert st == 0
    m, n, rows, rsym = reference_rows(text)
    assert rows_of(raw) == rows and sym == rsym
    can, sym2 = ts.read_mtx(path)
    assert sym2 == sym
    for i in range(can.m):
        assert np.all(np.diff(can.colidx[can.rowptr[i]:can.rowptr[i + 1]]) > 0)
    if fn != "er_80x96_general.mtx":
        assert np.array_equal(raw.rowptr, can.rowptr) and np.array_equal(raw.colidx, can.colidx)
        assert np.array_equal(raw.val, can.val)
    if fn == "stencil5_8x8_sym.mtx":
        ref = ts.gen_stencil5(8, 8)
        assert np.array_equal(can.rowptr, ref.rowptr) and np.array_equal(can.colidx, ref.colidx)
        assert np.array_equal(can.val, ref.val)


def test_canonicalize_matches_scipy():
    rng = np.random.default_rng(5)
    m, n = 50, 40
    rows = [sorted(rng.integers(0, n, size=rng.integers(0, 12)).tolist(), key=lambda _: rng.random())
            for _ in range(m)]
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.array([c for r in rows for c in r], np.int32)
    va = rng.uniform(-1, 1, size=len(ci))
    out = ts.canonicalize(ts.Csr(m, n, rp, ci, va))
    ref = sp.csr_matrix((va, ci, rp), shape=(m, n))
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(out.rowptr, ref.indptr) and np.array_equal(out.colidx, ref.indices)
    assert np.allclose(out.val, ref.data, rtol=0, atol=1e-15)


# ------------------------------------------------------------------ order files
@pytest.mark.parametrize("body,n,code", [
    ("2\n0\n1\n", 3, 0),
    ("2 0\n  1\n", 3, 0),
    ("", 0, 0),
    ("2\n0\n", 3, -11),
    ("2\nx\n1\n", 3, -11),
    ("2\n3\n1\n", 3, -12),
    ("2\n-1\n1\n", 3, -12),
    ("2\n2\n1\n", 3, -13),
    ("2\n0\n1\n1\n", 3, -14),
])
def test_order_file_contract(tmp_path, body, n, code):
    p = tmp_path / "x.order"
    p.write_text(body)
    st, old, new = ts.read_order_file(str(p), n)
    assert st == code
    if code == 0:
        ids = [int(t) for t in body.split()]
        assert list(old) == ids
        assert all(new[o] == i for i, o in enumerate(ids))


def test_order_file_missing(tmp_path):
    assert ts.read_order_file(str(tmp_path / "none.order"), 4)[0] == -10


# ------------------------------------------------------------------ plan argument checks
def smat(m, n, seed=0):
    rng = np.random.default_rng(seed)
    a = sp.random(m, n, density=0.3, random_state=rng, format="csr")
    a.sort_indices()
    return ts.HostSMatrix(ts.Csr.from_scipy(a))


def test_external_plan_status_codes(tmp_path, monkeypatch):
    monkeypatch.setenv("TASA_REORDER_ALGORITHM_ROOT", str(tmp_path))
    sq4, sq4b, r34, r43, sq5, r45 = smat(4, 4), smat(4, 4, 1), smat(3, 4), smat(4, 3), smat(5, 5), smat(4, 5)
    E = ts.PROBLEM_AA, ts.PROBLEM_AAT, ts.PROBLEM_AB
    plan = ts.reorder_plan
    assert plan(E[0], "rcm", None)[0] == -1
    assert plan(E[1], "rcm", sq4, None)[0] == -2
    assert plan(E[2], "rcm", sq4, sq5)[0] == -3
    assert plan(E[0], "rcm", r34)[0] == -4
    assert plan(E[1], "rcm", r34, r43)[0] == -5
    assert plan(E[2], "rcm", r34, sq4)[0] == -6
    assert plan(E[2], "rcm", sq4, r45)[0] == -7
    assert plan(7, "rcm", sq4, sq4b)[0] == -9
    # no adapter named "nosuch" under the root: the generator cannot run
    assert plan(E[0], "nosuch", sq4, None, path_a="m.mtx")[0] == -20
    assert plan(E[0], None, sq4)[0] == -20
    # build_reorder_plan's own codes (src/reorder.h:2407-2427)
    assert ts.reorder_plan(E[0], None, None, builder="tasa")[0] == -1
    assert ts.reorder_plan(E[1], None, sq4, None, builder="tasa")[0] == -3
    assert ts.reorder_plan(E[2], None, sq4, None, builder="tasa")[0] == -3
    assert ts.reorder_plan(5, None, sq4, sq4b, builder="tasa")[0] == -2
    # build_original_reorder_plan's
    assert ts.reorder_plan(E[1], None, sq4, None, builder="original")[0] == -2


def test_apply_and_restore_argument_checks():
    a = smat(4, 4)
    p = ts.AxisPermutation()
    ts.lib.axis_permutation_init(ts.C.byref(p))
    try:
        assert ts.lib.axis_permutation_alloc(ts.C.byref(p), 3) == 0
        out = ts.SMatrix()
        assert ts.lib.apply_bipermutation_csr(a.ref, ts.C.byref(p), ts.C.byref(p), ts.C.byref(out)) == -1
        assert ts.lib.apply_bipermutation_csr(a.ref, None, None, ts.C.byref(out)) == -1
        assert ts.lib.restore_bipermutation_csr(a.ref, ts.C.byref(p), ts.C.byref(p)) == -1
        assert ts.lib.restore_bipermutation_csr(a.ref, None, None) == 0
    finally:
        ts.lib.axis_permutation_destroy(ts.C.byref(p))


# ------------------------------------------------------------------ order_gen adapters
ADAPTER = r"""#!/bin/sh
# order_gen <matrix.mtx> <out.order>: n from the size line; the order is
# chosen by a "% order: <kind>" comment in the matrix file
in="$1"; out="$2"
kind=$(sed -n 's/^% order: //p' "$in" | head -n 1)
n=$(grep -v '^%' "$in" | head -n 1 | cut -d' ' -f1)
case "$kind" in
  reverse) i=$((n-1)); : > "$out"; while [ $i -ge 0 ]; do echo $i >> "$out"; i=$((i-1)); done ;;
  rotate)  : > "$out"; i=1; while [ $i -lt $n ]; do echo $i >> "$out"; i=$((i+1)); done; echo 0 >> "$out" ;;
  dup)     : > "$out"; i=0; while [ $i -lt $n ]; do echo 0 >> "$out"; i=$((i+1)); done ;;
  missing) exit 42 ;;
  *) exit 3 ;;
esac
"""

needs_no_gpu = pytest.mark.skipif(os.path.exists("/dev/kfd"),
                                  reason="spawns a child process; run on a host without a GPU")


def write_matrix(path, kind, csr):
    rows = np.repeat(np.arange(csr.m), np.diff(csr.rowptr))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"% order: {kind}\n{csr.m} {csr.n} {csr.nnz}\n")
        for r, c, v in zip(rows, csr.colidx, csr.val):
            f.write(f"{int(r) + 1} {int(c) + 1} {float(v)!r}\n")


@needs_no_gpu
def test_adapter_orders_and_axis_sharing(tmp_path, monkeypatch):
    root = tmp_path / "algos"
    (root / "myorder").mkdir(parents=True)
    exe = root / "myorder" / "order_gen"
    exe.write_text(ADAPTER)
    exe.chmod(exe.stat().st_mode | stat.S_IXUSR)
    cache = tmp_path / "orders"
    monkeypatch.setenv("TASA_REORDER_ALGORITHM_ROOT", str(root))
    monkeypatch.setenv("TASA_REORDER_ORDER_DIR", str(cache))
    n = 6
    a, b = smat(n, n, 3), smat(n, n, 4)
    pa, pb, pd, pm = (str(tmp_path / f) for f in ("a.mtx", "b.mtx", "d.mtx", "m.mtx"))
    write_matrix(pa, "reverse", a.csr)
    write_matrix(pb, "rotate", b.csr)
    write_matrix(pd, "dup", a.csr)
    write_matrix(pm, "missing", a.csr)
    rev = list(range(n - 1, -1, -1))
    rot = list(range(1, n)) + [0]
    for kind in (ts.PROBLEM_AA, ts.PROBLEM_AAT):
        st, info, ms = ts.reorder_plan(kind, "myorder", a, a if kind else None, path_a=pa, path_b=pb)
        assert st == 0 and info["enabled"] == 1 and info["mode"] == "myorder"
        assert info["risk_reason"] == "external_order" and ms > 0
        for ax in ("row_perm", "inner_perm", "col_perm"):
            old, new = info[ax]
            assert list(old) == rev and all(new[o] == i for i, o in enumerate(old))
        assert info["shared_symmetric_csr"] == (kind == ts.PROBLEM_AA)
        assert info["estimated_c_tiles_after"] == -1  # no device: no tile estimates
    st, info, _ = ts.reorder_plan(ts.PROBLEM_AB, "myorder", a, b, path_a=pa, path_b=pb)
    assert st == 0 and info["shared_symmetric_csr"] == 0
    assert list(info["row_perm"][0]) == rev and list(info["inner_perm"][0]) == rev
    assert list(info["col_perm"][0]) == rot
    # AB without a B path orders B's axis from A's file
    st, info, _ = ts.reorder_plan(ts.PROBLEM_AB, "myorder", a, b, path_a=pa)
    assert st == 0 and list(info["col_perm"][0]) == rev
    # the adapter's order file is removed after a successful read
    assert not [f for f in os.listdir(cache) if f.endswith(".order")]
    assert ts.reorder_plan(ts.PROBLEM_AA, "myorder", a, path_a=pd)[0] == -13
    assert ts.reorder_plan(ts.PROBLEM_AA, "myorder", a, path_a=pm)[0] == -20
    assert ts.reorder_plan(ts.PROBLEM_AA, "myorder", a, path_a=str(tmp_path / "x.mtx"))[0] == -20


# ------------------------------------------------------------------ driver refusals (before any device work)
@pytest.mark.parametrize("args,rc", [
    (["-d", "0", "-aat", "1", os.path.join(GOLDEN, "stencil5_8x8_sym.mtx")], 0),     # AA^T of a symmetric file
    (["-d", "0", "-aat", "0", os.path.join(GOLDEN, "er_80x96_general.mtx")], 0),     # A^2 of a non-square A
    (["-d", "0", "-aat", "1", os.path.join(GOLDEN, "er_96_general.mtx"),
      "-b", os.path.join(GOLDEN, "er_96_general.mtx")], 0),                           # A*B with -aat 1
    (["-d", "0", "-aat", "0", os.path.join(GOLDEN, "er_80x96_general.mtx"),
      "-b", os.path.join(GOLDEN, "stencil5_8x8_sym.mtx")], 0),                        # colA != rowB
    (["-d", "0", "-aat", "0", "/nonexistent/a.mtx"], 1),                              # cannot load
    (["-d", "0", "-aat"], 0),                                                         # usage
    (["-d", "0", "-aat", "0", os.path.join(GOLDEN, "er_96_general.mtx"), "-reorder"], 0),  # missing value
])
def test_driver_refusals(tmp_path, args, rc):
    out = tmp_path / "c.bin"
    assert ts.driver(args + ["-o", str(out)] if len(args) > 4 and args[-1] != "-reorder" else args) == rc
    assert not out.exists()


@needs_no_gpu
def test_driver_unknown_reorder_adapter(tmp_path, monkeypatch):
    monkeypatch.setenv("TASA_REORDER_ALGORITHM_ROOT", str(tmp_path))
    out = tmp_path / "c.bin"
    args = ["-d", "0", "-aat", "0", os.path.join(GOLDEN, "er_96_general.mtx"), "-reorder", "1",
            "-reorder_algo", "nosuch", "-o", str(out)]
    assert ts.driver(args) == 0 and not out.exists()
