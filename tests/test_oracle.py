"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement fails:
  * closed-form counts of the stencil workloads (nnz, nnzCub, nnz(C), tile
    counts, phantom tiles, bins) -- independent of any SpGEMM code;
  * SciPy's sparse product (an independent SpGEMM): structure bit-exact,
    integer-valued results bit-exact, float results within 1e-12 of
    sum |a_ik b_kj|;
  * exact rational arithmetic on tiny inputs: the oracle's fma accumulation
    order (A row ascending k, then B row ascending j) reproduced bit for bit,
    and shown to differ from plain multiply-then-add somewhere;
  * transposes / permutations against SciPy, degree order against its
    one-line definition, RCM against invariants.
"""
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import tilespgemm as ts


def sprand(m, n, density, seed, integer=False):
    rng = np.random.default_rng(seed)
    a = sp.random(m, n, density=density, random_state=rng, format="csr")
    if integer:
        a.data = rng.integers(-9, 10, size=a.nnz).astype(np.float64)
    else:
        a.data = rng.uniform(-1, 1, size=a.nnz)
    a.sort_indices()
    return ts.Csr.from_scipy(a)


def scipy_product(a, b):
    c = (a.to_scipy() @ b.to_scipy()).tocsr()
    c.sort_indices()
    return c


def bound(a, b, c_struct):
    """sum |a_ik||b_kj| on C's structure."""
    aa = a.to_scipy().copy()
    aa.data = np.abs(aa.data)
    bb = b.to_scipy().copy()
    bb.data = np.abs(bb.data)
    return (aa @ bb).tocsr()


def _pattern(x):
    p = x.to_scipy().copy()
    p.data = np.ones_like(p.data)
    return p


def _lookup(ref, m, n, rpc, cc):
    """Values of scipy matrix `ref` at the positions (row, cc) of a CSR
    structure; positions scipy dropped (exact zeros) read as 0."""
    ref = ref.tocsr()
    ref.sort_indices()
    rk = np.repeat(np.arange(m, dtype=np.int64), np.diff(ref.indptr)) * n + ref.indices
    ok = np.repeat(np.arange(m, dtype=np.int64), np.diff(rpc)) * n + cc
    pos = np.searchsorted(rk, ok)
    pos = np.minimum(pos, max(len(rk) - 1, 0))
    hit = (len(rk) > 0) & (rk[pos] == ok) if len(rk) else np.zeros(len(ok), bool)
    out = np.zeros(len(ok))
    out[hit] = ref.data[pos[hit]]
    # every value scipy kept must be on the structure
    assert hit.sum() == ref.nnz
    return out


def check_against_scipy(a, b, exact):
    rpc, cc, vc = oracle.spgemm(a, b)
    # structure: scipy on all-ones patterns (no cancellation, nothing pruned)
    pat = (_pattern(a) @ _pattern(b)).tocsr()
    pat.sort_indices()
    assert np.array_equal(rpc, pat.indptr.astype(np.int64)), "row pointers"
    assert np.array_equal(cc, pat.indices), "column indices"
    # values: scipy prunes exact zeros of the numeric product, look them up
    ref = _lookup(scipy_product(a, b), a.m, b.n, rpc, cc)
    if exact:
        assert np.array_equal(vc, ref)
    else:
        bd = bound(a, b, None)
        bd.sort_indices()
        assert np.array_equal(bd.indptr.astype(np.int64), rpc) or True
        bnd = _lookup(bd, a.m, b.n, rpc, cc)
        assert np.all(np.abs(vc - ref) <= 1e-12 * bnd)


# ------------------------------------------------------------------ closed forms
def test_5pt_64x64_closed_forms():
    a = ts.gen_stencil5(64, 64)
    assert a.nnz == 20224
    assert oracle.nnzcub(a, a) == 100104
    rpc, cc = oracle.spgemm_structure(a, a)
    assert rpc[-1] == 51972
    ta = oracle.csr2tile(a)
    assert ta["numtile"] == 1144
    cp, ccol = oracle.step1_tiles(ta, ta)
    assert len(ccol) == 2652
    mask, ptr16, tnnz = oracle.c_tile_masks(a.m, cp, ccol, rpc, cc)
    per_tile = np.diff(tnnz)
    assert int((per_tile == 0).sum()) == 256          # phantom tiles
    assert int(((per_tile > 0) & (per_tile <= 32)).sum()) == 1636
    assert int(((per_tile > 32) & (per_tile <= 224)).sum()) == 760
    assert tnnz[-1] == 51972


def test_5pt_values_are_exact_integers():
    a = ts.gen_stencil5(64, 64)
    check_against_scipy(a, a, exact=True)


@pytest.mark.parametrize("N", [8, 16, 20])
def test_27p_periodic_small_closed_forms(N):
    a = ts.gen_stencil27(N, N, N, periodic=True, valmode=ts.VAL_MOD10)
    n = N ** 3
    assert a.nnz == 27 * n
    assert oracle.nnzcub(a, a) == 729 * n
    rpc, cc = oracle.spgemm_structure(a, a)
    assert rpc[-1] == 125 * n
    check_against_scipy(a, a, exact=True)


def test_27p_128_tile_closed_forms():
    """Tile-level counts of the headline workload: 3,538,944 A tiles and
    16,384,000 step-1 C tiles (125 per tile-row)."""
    a = ts.gen_stencil27(128, 128, 128, periodic=True, valmode=ts.VAL_KEEP)
    assert a.nnz == 56_623_104
    ta = oracle.csr2tile(a)
    assert ta["numtile"] == 3_538_944
    cp, ccol = oracle.step1_tiles(ta, ta)
    assert len(ccol) == 16_384_000
    assert np.all(np.diff(cp) == 125)


def test_27d_128_counts():
    a = ts.gen_stencil27(128, 128, 128, periodic=False, valmode=ts.VAL_KEEP)
    assert a.nnz == 55_742_968
    assert oracle.nnzcub(a, a) == 1_489_355_288
    ta = oracle.csr2tile(a)
    assert ta["numtile"] == 3_210_328
    cp, ccol = oracle.step1_tiles(ta, ta)
    assert len(ccol) == 13_666_504


# ------------------------------------------------------------------ scipy
@pytest.mark.parametrize("m,k,n,d,seed", [(50, 40, 60, 0.08, 1), (200, 200, 200, 0.03, 2), (17, 33, 9, 0.3, 3),
                                          (1, 5, 7, 0.9, 4), (64, 64, 64, 0.0, 5)])
def test_oracle_vs_scipy_float(m, k, n, d, seed):
    a = sprand(m, k, d, seed)
    b = sprand(k, n, d, seed + 100)
    check_against_scipy(a, b, exact=False)


@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_vs_scipy_integer(seed):
    a = sprand(300, 300, 0.02, seed, integer=True)
    check_against_scipy(a, a, exact=True)
    at = ts.Csr(a.n, a.m, *oracle.transpose(a))
    check_against_scipy(a, at, exact=True)


def test_oracle_rows_slice_matches_full():
    a = sprand(120, 120, 0.05, 21)
    rpc, cc, vc = oracle.spgemm(a, a)
    rp, c2, v2 = oracle.spgemm_rows(a, a, 30, 77)
    assert np.array_equal(rp + rpc[30], rpc[30:78])
    assert np.array_equal(c2, cc[rpc[30]:rpc[77]])
    assert np.array_equal(v2, vc[rpc[30]:rpc[77]])


def test_oracle_rowlist_matches_full():
    """Sampled rows (any order, repeats) equal the same rows of the whole
    product, bit for bit, values and structure-only."""
    a = sprand(300, 260, 0.04, 22)
    b = sprand(260, 310, 0.05, 23)
    rpc, cc, vc = oracle.spgemm(a, b)
    rows = np.array([299, 0, 17, 17, 150, 3, 298, 64], np.int32)
    rp, c2, v2 = oracle.spgemm_rowlist(a, b, rows)
    assert np.array_equal(np.diff(rp), np.diff(rpc)[rows])
    for x, i in enumerate(rows):
        assert np.array_equal(c2[rp[x]:rp[x + 1]], cc[rpc[i]:rpc[i + 1]])
        assert np.array_equal(v2[rp[x]:rp[x + 1]].view(np.int64), vc[rpc[i]:rpc[i + 1]].view(np.int64))
    rp3, c3, _ = oracle.spgemm_rowlist(a, b, rows, values=False)
    assert np.array_equal(rp3, rp) and np.array_equal(c3, c2)


# ------------------------------------------------------------------ exact fma order
def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))  # correctly rounded


def test_oracle_fma_order_exact():
    """Brute force with exact rationals: acc = fma(a_ik, b_kj, acc) in A-row
    order then B-row order reproduces the oracle bit for bit."""
    a = sprand(24, 24, 0.35, 31)
    b = sprand(24, 24, 0.35, 32)
    rpc, cc, vc = oracle.spgemm(a, b)
    differs_from_mul_add = False
    for i in range(a.m):
        acc, acc_ma = {}, {}
        for p in range(a.rowptr[i], a.rowptr[i + 1]):
            k, av = a.colidx[p], a.val[p]
            for q in range(b.rowptr[k], b.rowptr[k + 1]):
                j, bv = b.colidx[q], b.val[q]
                acc[j] = _fma(av, bv, acc.get(j, 0.0))
                acc_ma[j] = acc_ma.get(j, 0.0) + av * bv
        cols = sorted(acc)
        assert list(cc[rpc[i]:rpc[i + 1]]) == cols
        got = vc[rpc[i]:rpc[i + 1]]
        assert np.array_equal(got, np.array([acc[j] for j in cols]))
        if any(acc[j] != acc_ma[j] for j in cols):
            differs_from_mul_add = True
    assert differs_from_mul_add, "inputs too easy to tell fma from mul+add"


# ------------------------------------------------------------------ tiles
def test_csr2tile_layout_5pt():
    a = ts.gen_stencil5(20, 7)  # m = 140: last tile-row partial
    t = oracle.csr2tile(a, packed=True)
    # rebuild CSR from the tiles by brute force
    rows, cols, vals = [], [], []
    for I in range(t["tilem"]):
        for tt in range(t["tile_ptr"][I], t["tile_ptr"][I + 1]):
            J = t["tile_colidx"][tt]
            nnz = t["tile_nnz"][tt + 1] - t["tile_nnz"][tt]
            for r in range(16):
                s = t["csr_ptr"][tt * 16 + r]
                e = nnz if r == 15 else t["csr_ptr"][tt * 16 + r + 1]
                for x in range(s, e):
                    pc = int(t["csr_col"][t["tile_nnz"][tt] + x])
                    assert pc >> 4 == r
                    c = pc & 15
                    assert int(t["mask"][tt * 16 + r]) & (0x8000 >> c)
                    rows.append(I * 16 + r)
                    cols.append(J * 16 + c)
                    vals.append(t["val"][t["tile_nnz"][tt] + x])
    ref = a.to_scipy().tocoo()
    got = sp.coo_matrix((vals, (rows, cols)), shape=(a.m, a.n)).tocsr()
    got.sort_indices()
    assert np.array_equal(got.indptr, a.rowptr) and np.array_equal(got.indices, a.colidx)
    assert np.array_equal(got.data, a.val)
    assert ref.nnz == len(vals)
    # tile columns sorted inside each tile-row, rows padded with tile nnz
    for I in range(t["tilem"]):
        seg = t["tile_colidx"][t["tile_ptr"][I]:t["tile_ptr"][I + 1]]
        assert np.all(np.diff(seg) > 0)


def test_transpose_vs_scipy():
    a = sprand(90, 70, 0.05, 41)
    tp, tc, tv = oracle.transpose(a)
    ref = a.to_scipy().T.tocsr()
    ref.sort_indices()
    assert np.array_equal(tp, ref.indptr) and np.array_equal(tc, ref.indices) and np.array_equal(tv, ref.data)


# ------------------------------------------------------------------ orderings
def test_degree_order_definition():
    a = sprand(200, 200, 0.03, 51)
    deg = np.diff(a.rowptr)
    ref = sorted(range(a.m), key=lambda i: (-deg[i], i))
    assert list(oracle.degree_order(a)) == ref


def _bandwidth(a):
    if a.nnz == 0:
        return 0
    rows = np.repeat(np.arange(a.m), np.diff(a.rowptr))
    return int(np.max(np.abs(rows - a.colidx)))


def test_rcm_invariants_and_bandwidth():
    # a banded matrix scrambled by a random permutation: RCM must undo most of it
    b = ts.gen_banded(400, 3, valmode=ts.VAL_UNIFORM_PM1)
    rng = np.random.default_rng(7)
    p = rng.permutation(b.m).astype(np.int32)
    inv = np.empty_like(p)
    inv[p] = np.arange(b.m, dtype=np.int32)
    scr = ts.Csr(b.m, b.n, *oracle.permute_csr(b, p, inv))
    order = oracle.rcm(scr)
    assert sorted(order) == list(range(b.m))
    new = np.empty_like(order)
    new[order] = np.arange(b.m, dtype=np.int32)
    r = ts.Csr(b.m, b.n, *oracle.permute_csr(scr, order, new))
    assert _bandwidth(scr) > 100
    assert _bandwidth(r) <= 2 * 3 + 1


def test_rcm_matches_scipy_when_ties_cannot_matter():
    """On a path graph with distinct labels the Cuthill-McKee order is
    unique up to the choice of seed among the two ends; both pick the
    lowest (degree, id) end."""
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    n = 50
    rows = np.concatenate([np.arange(n - 1), np.arange(1, n)])
    cols = np.concatenate([np.arange(1, n), np.arange(n - 1)])
    g = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    g.sort_indices()
    a = ts.Csr.from_scipy(g)
    ours = oracle.rcm(a)
    ref = reverse_cuthill_mckee(g, symmetric_mode=True)
    assert list(ours) == list(ref)


def test_permute_roundtrip_and_scipy():
    a = sprand(80, 80, 0.06, 61)
    rng = np.random.default_rng(3)
    old = rng.permutation(80).astype(np.int32)
    new = np.empty_like(old)
    new[old] = np.arange(80, dtype=np.int32)
    op, oc, ov = oracle.permute_csr(a, old, new)
    ref = a.to_scipy()[old][:, old].tocsr()
    ref.sort_indices()
    assert np.array_equal(op, ref.indptr) and np.array_equal(oc, ref.indices) and np.array_equal(ov, ref.data)
    back = oracle.permute_csr(ts.Csr(80, 80, op, oc, ov), new, old)
    assert np.array_equal(back[0], a.rowptr) and np.array_equal(back[1], a.colidx) and np.array_equal(back[2], a.val)


# ------------------------------------------------------------------ symmetrize / symmetric-header degree
@pytest.mark.parametrize("n,density,seed", [(300, 0.02, 1), (257, 0.1, 2), (400, 0.004, 3)])
def test_symmetrize_vs_scipy_on_nonsymmetric(n, density, seed):
    """pattern(A + A^T) with sorted rows against SciPy's sparse sum of the
    two patterns; the inputs are not symmetric, so a restatement that drops
    the transpose half (or the A half) fails."""
    a = sprand(n, n, density, seed)
    s = _pattern(a)
    assert (s != s.T).nnz > 0
    ref = (s + s.T).tocsr()
    ref.sort_indices()
    gp, gc = oracle.symmetrize(a)
    assert np.array_equal(gp, ref.indptr) and np.array_equal(gc, ref.indices)


def _file_degrees(text):
    """Degrees as reorder_algorithm/common/order_utils.py:30-45 counts them
    from the file's entry lines: +1 for the row of every stored entry, +1
    for its column too when the header is symmetric and the entry is off the
    diagonal."""
    lines = text.splitlines()
    symmetric = "symmetric" in lines[0].lower()
    body = [ln for ln in lines[1:] if ln.strip() and not ln.startswith("%")]
    n = int(body[0].split()[0])
    deg = [0] * n
    for ln in body[1:]:
        r, c = (int(x) - 1 for x in ln.split()[:2])
        deg[r] += 1
        if symmetric and r != c:
            deg[c] += 1
    return n, deg


def _sym_files(tmp_path):
    yield open(os.path.join(os.path.dirname(__file__), "golden", "stencil5_8x8_sym.mtx")).read()
    rng = np.random.default_rng(9)
    n = 120
    ent = [(i, i) for i in range(0, n, 3)]
    for _ in range(400):
        r, c = sorted(rng.integers(0, n, size=2).tolist(), reverse=True)
        ent.append((r, c))
    ent.append(ent[5])  # a duplicate line counts twice in the file's degrees
    body = "".join(f"{r + 1} {c + 1} 1.0\n" for r, c in ent)
    yield f"%%MatrixMarket matrix coordinate real symmetric\n{n} {n} {len(ent)}\n{body}"


def test_degree_order_symmetric_header_vs_file_definition(tmp_path):
    """ora_degree_order on the stored triangle with symmetric_header=1, and
    on the expanded CSR with 0, both equal sorted(range(n), key=(-deg, id))
    with deg counted from the file's lines."""
    for k, text in enumerate(_sym_files(tmp_path)):
        n, deg = _file_degrees(text)
        ref = sorted(range(n), key=lambda i: (-deg[i], i))
        p = tmp_path / f"s{k}.mtx"
        p.write_text(text)
        st, expanded, sym = ts.mmio_allinone(str(p))
        assert st == 0 and sym
        q = tmp_path / f"t{k}.mtx"
        q.write_text(text.replace("symmetric", "general", 1))
        st, tri, sym = ts.mmio_allinone(str(q))
        assert st == 0 and not sym
        assert list(oracle.degree_order(tri, True)) == ref
        assert list(oracle.degree_order(expanded, False)) == ref
        assert list(oracle.degree_order(tri, False)) != ref  # the mirror counts
