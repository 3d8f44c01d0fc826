"""Python binding of the CPU oracle (oracle.c).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker / CPU baseline; the product library never uses it.

Every function takes and returns numpy arrays; see oracle.c for the
reference routine each one restates (file:line).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()
_p = C.c_void_p
_i = C.c_int
_ll = C.c_longlong


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


for name, res, args in [
    ("ora_spa_structure", _ll, [_i, _p, _p, _p, _p, _i, _p, _p, _i]),
    ("ora_spa_values", _i, [_i, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p]),
    ("ora_spgemm_rows", _ll, [_i, _i, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p]),
    ("ora_spgemm_rowlist", _ll, [_i, _p, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p]),
    ("ora_nnzcub", C.c_ulonglong, [_i, _p, _p]),
    ("ora_csr2tile", _ll, [_i, _i, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p]),
    ("ora_step1_tiles", _ll, [_i, _p, _p, _i, _p, _p, _p, _p]),
    ("ora_c_tile_masks", _i, [_i, _i, _p, _p, _ll, _p, _p, _p, _p, _p]),
    ("ora_transpose", _i, [_i, _i, _p, _p, _p, _p, _p, _p]),
    ("ora_degree_order", _i, [_i, _p, _p, _i, _p]),
    ("ora_symmetrize", _i, [_i, _p, _p, _p, _p, _i]),
    ("ora_rcm", _i, [_i, _p, _p, _p]),
    ("ora_permute_csr", _i, [_i, _p, _p, _p, _p, _p, _p, _p, _p]),
    ("ora_num_threads", _i, []),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def spgemm_structure(a, b):
    """Row pointers (int64) and sorted columns of C = A*B, by the reference
    serial SPA (count pass + fill pass)."""
    rpa, ca, rpb, cb = _i32(a.rowptr), _i32(a.colidx), _i32(b.rowptr), _i32(b.colidx)
    rpc = np.zeros(a.m + 1, np.int64)
    nnz = lib.ora_spa_structure(a.m, _ptr(rpa), _ptr(ca), _ptr(rpb), _ptr(cb), b.n, _ptr(rpc), None, 1)
    if nnz < 0:
        raise MemoryError("oracle count pass")
    cc = np.zeros(max(nnz, 1), np.int32)
    lib.ora_spa_structure(a.m, _ptr(rpa), _ptr(ca), _ptr(rpb), _ptr(cb), b.n, _ptr(rpc), _ptr(cc), 0)
    return rpc, cc[:nnz]


def spgemm_values(a, b, rpc, cc):
    """Values of C on the given structure, compute_dense_row order with fma."""
    out = np.zeros(max(len(cc), 1), np.float64)
    st = lib.ora_spa_values(a.m, _ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), _ptr(_f64(a.val)),
                            _ptr(_i32(b.rowptr)), _ptr(_i32(b.colidx)), _ptr(_f64(b.val)), b.n,
                            _ptr(np.ascontiguousarray(rpc, np.int64)), _ptr(_i32(cc)), _ptr(out))
    if st:
        raise MemoryError("oracle values")
    return out[:len(cc)]


def spgemm(a, b):
    rpc, cc = spgemm_structure(a, b)
    return rpc, cc, spgemm_values(a, b, rpc, cc)


def spgemm_rows(a, b, r0, r1, values=True):
    """Structure (+values) of rows [r0, r1) of C = A*B."""
    rp = np.zeros(r1 - r0 + 1, np.int64)
    args = (_ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), _ptr(_f64(a.val)) if values else None,
            _ptr(_i32(b.rowptr)), _ptr(_i32(b.colidx)), _ptr(_f64(b.val)), b.n)
    nnz = lib.ora_spgemm_rows(r0, r1, *args, _ptr(rp), None, None)
    cc = np.zeros(max(nnz, 1), np.int32)
    vv = np.zeros(max(nnz, 1), np.float64)
    lib.ora_spgemm_rows(r0, r1, *args, _ptr(rp), _ptr(cc), _ptr(vv) if values else None)
    return rp, cc[:nnz], vv[:nnz]


def spgemm_rowlist(a, b, rows, values=True):
    """Structure (+values) of the rows `rows` of C = A*B (any order):
    returns (rowptr int64[len(rows)+1], columns, values)."""
    r = np.ascontiguousarray(rows, dtype=np.int32)
    cnt = np.zeros(len(r), np.int64)
    args = (_ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), _ptr(_f64(a.val)) if values else None,
            _ptr(_i32(b.rowptr)), _ptr(_i32(b.colidx)), _ptr(_f64(b.val)), b.n)
    tot = lib.ora_spgemm_rowlist(len(r), _ptr(r), *args, _ptr(cnt), None, None, None)
    assert tot >= 0, "oracle allocation failed"
    rp = np.zeros(len(r) + 1, np.int64)
    np.cumsum(cnt, out=rp[1:])
    cc = np.zeros(max(tot, 1), np.int32)
    vv = np.zeros(max(tot, 1), np.float64)
    lib.ora_spgemm_rowlist(len(r), _ptr(r), *args, None, _ptr(rp), _ptr(cc), _ptr(vv) if values else None)
    return rp, cc[:tot], vv[:tot]


def num_threads():
    return int(lib.ora_num_threads())


def spa_structure_rows_timed(a, b, r0, r1):
    """Both passes of the reference SPA (count, then fill) over rows
    [r0, r1) of A, as src/external/cusparse/main.cu:191-222 times them;
    returns (seconds, nnz of the slice)."""
    import time
    rpa, ca, rpb, cb = _i32(a.rowptr), _i32(a.colidx), _i32(b.rowptr), _i32(b.colidx)
    m = r1 - r0
    rp_view = rpa[r0:]
    rpc = np.zeros(m + 1, np.int64)
    t0 = time.perf_counter()
    nnz = lib.ora_spa_structure(m, _ptr(rp_view), _ptr(ca), _ptr(rpb), _ptr(cb), b.n, _ptr(rpc), None, 1)
    cc = np.empty(max(nnz, 1), np.int32)
    lib.ora_spa_structure(m, _ptr(rp_view), _ptr(ca), _ptr(rpb), _ptr(cb), b.n, _ptr(rpc), _ptr(cc), 0)
    return time.perf_counter() - t0, int(nnz)


def nnzcub(a, b):
    return int(lib.ora_nnzcub(a.nnz, _ptr(_i32(a.colidx)), _ptr(_i32(b.rowptr))))


def csr2tile(a, packed=False):
    tilem, tilen = (a.m + 15) // 16, (a.n + 15) // 16
    rp, ci, va = _i32(a.rowptr), _i32(a.colidx), _f64(a.val)
    tp = np.zeros(tilem + 1, np.int32)
    nt = lib.ora_csr2tile(a.m, a.n, _ptr(rp), _ptr(ci), _ptr(va), int(packed), _ptr(tp), None, None, None, None,
                          None, None, None)
    out = dict(tilem=tilem, tilen=tilen, numtile=int(nt), tile_ptr=tp,
               tile_colidx=np.zeros(nt, np.int32), tile_rowidx=np.zeros(nt, np.int32),
               tile_nnz=np.zeros(nt + 1, np.int32), csr_ptr=np.zeros(nt * 16, np.uint8),
               csr_col=np.zeros(a.nnz, np.uint8), val=np.zeros(a.nnz, np.float64),
               mask=np.zeros(nt * 16, np.uint16))
    lib.ora_csr2tile(a.m, a.n, _ptr(rp), _ptr(ci), _ptr(va), int(packed), _ptr(tp), _ptr(out["tile_colidx"]),
                     _ptr(out["tile_rowidx"]), _ptr(out["tile_nnz"]), _ptr(out["csr_ptr"]), _ptr(out["csr_col"]),
                     _ptr(out["val"]), _ptr(out["mask"]))
    return out


def step1_tiles(ta, tb):
    """C tile pointer and sorted tile columns of the tile-level boolean
    product (phantom tiles included)."""
    cp = np.zeros(ta["tilem"] + 1, np.int32)
    nt = lib.ora_step1_tiles(ta["tilem"], _ptr(ta["tile_ptr"]), _ptr(ta["tile_colidx"]), tb["tilen"],
                             _ptr(tb["tile_ptr"]), _ptr(tb["tile_colidx"]), _ptr(cp), None)
    cc = np.zeros(max(nt, 1), np.int32)
    lib.ora_step1_tiles(ta["tilem"], _ptr(ta["tile_ptr"]), _ptr(ta["tile_colidx"]), tb["tilen"],
                        _ptr(tb["tile_ptr"]), _ptr(tb["tile_colidx"]), _ptr(cp), _ptr(cc))
    return cp, cc[:nt]


def c_tile_masks(m, cptr, ccol, rpc, colc):
    nt = len(ccol)
    mask = np.zeros(nt * 16, np.uint16)
    ptr16 = np.zeros(nt * 16, np.uint8)
    tnnz = np.zeros(nt + 1, np.int64)
    st = lib.ora_c_tile_masks(m, (m + 15) // 16, _ptr(_i32(cptr)), _ptr(_i32(ccol)), nt,
                              _ptr(np.ascontiguousarray(rpc, np.int64)), _ptr(_i32(colc)), _ptr(mask), _ptr(ptr16),
                              _ptr(tnnz))
    if st:
        raise RuntimeError(f"ora_c_tile_masks failed {st}")
    return mask, ptr16, tnnz


def transpose(a):
    tp = np.zeros(a.n + 1, np.int32)
    tc = np.zeros(max(a.nnz, 1), np.int32)
    tv = np.zeros(max(a.nnz, 1), np.float64)
    lib.ora_transpose(a.m, a.n, _ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), _ptr(_f64(a.val)), _ptr(tp), _ptr(tc),
                      _ptr(tv))
    return tp, tc[:a.nnz], tv[:a.nnz]


def degree_order(a, symmetric_header=False):
    out = np.zeros(a.m, np.int32)
    lib.ora_degree_order(a.m, _ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), int(symmetric_header), _ptr(out))
    return out


def symmetrize(a):
    cap = 2 * a.nnz + 1
    gp = np.zeros(a.m + 1, np.int32)
    gc = np.zeros(cap, np.int32)
    n = lib.ora_symmetrize(a.m, _ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), _ptr(gp), _ptr(gc), cap)
    if n < 0:
        raise RuntimeError("ora_symmetrize failed")
    return gp, gc[:n]


def rcm(a):
    gp, gc = symmetrize(a)
    out = np.zeros(a.m, np.int32)
    lib.ora_rcm(a.m, _ptr(gp), _ptr(gc), _ptr(out))
    return out


def permute_csr(a, row_old, col_new):
    op = np.zeros(a.m + 1, np.int32)
    oc = np.zeros(max(a.nnz, 1), np.int32)
    ov = np.zeros(max(a.nnz, 1), np.float64)
    lib.ora_permute_csr(a.m, _ptr(_i32(a.rowptr)), _ptr(_i32(a.colidx)), _ptr(_f64(a.val)), _ptr(_i32(row_old)),
                        _ptr(_i32(col_new)), _ptr(op), _ptr(oc), _ptr(ov))
    return op, oc[:a.nnz], ov[:a.nnz]
