"""The drop-in entry points of include/tilespgemm.h, called exactly as the
reference driver calls them (src/main.cu:391-593): SMatrix in, SMatrix out,
host arrays in the reference layout.  Checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
import tilespgemm as ts
from tsg_check import log

pytestmark = pytest.mark.gpu

libc = C.CDLL(None)
libc.free.argtypes = [C.c_void_p]


def smatrix(csr, keep):
    s = ts.SMatrix()
    s.m, s.n, s.nnz = csr.m, csr.n, csr.nnz
    keep += [csr.rowptr, csr.colidx, csr.val]
    s.rowpointer = csr.rowptr.ctypes.data_as(C.POINTER(C.c_int))
    s.columnindex = csr.colidx.ctypes.data_as(C.POINTER(C.c_int))
    s.value = csr.val.ctypes.data_as(C.POINTER(C.c_double))
    return s


def arr(p, dtype, n):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def col_major_expected(ref):
    order = np.lexsort((ref["tile_rowidx"], ref["tile_colidx"]))
    sizes = np.diff(ref["tile_nnz"])
    nnz_csc = np.concatenate([[0], np.cumsum(sizes[order])]).astype(np.int32)
    val = np.concatenate([ref["val"][ref["tile_nnz"][t]:ref["tile_nnz"][t + 1]] for t in order]) if len(order) else np.zeros(0)
    col = np.concatenate([ref["csr_col"][ref["tile_nnz"][t]:ref["tile_nnz"][t + 1]] for t in order]) if len(order) else np.zeros(0, np.uint8)
    return dict(order=order, tile_nnz=nnz_csc, val=val, csr_col=col,
                csr_ptr=ref["csr_ptr"].reshape(-1, 16)[order].ravel(), mask=ref["mask"].reshape(-1, 16)[order].ravel(),
                csc_ptr=np.concatenate([[0], np.cumsum(np.bincount(ref["tile_colidx"], minlength=ref["tilen"]))]).astype(np.int32),
                csc_rowidx=ref["tile_rowidx"][order])


@pytest.mark.parametrize("name,aat", [("5pt", 0), ("er", 1), ("27p", 0)])
def test_reference_driver_flow(name, aat):
    log(f"[compat] {name} aat={aat}")
    keep = []
    if name == "5pt":
        a = ts.gen_stencil5(64, 64)
    elif name == "er":
        a = ts.gen_er(2000, 10.0, seed=3)
    else:
        a = ts.gen_stencil27(24, 24, 24, True)
    a.val = (np.arange(a.nnz) % 10).astype(np.float64)      # src/main.cu:191-192
    A = smatrix(a, keep)
    if aat:
        tp = np.zeros(a.n + 1, np.int32)
        tc = np.zeros(a.nnz, np.int32)
        tv = np.zeros(a.nnz, np.float64)
        ts.lib.matrix_transposition(a.m, a.n, a.nnz, A.rowpointer, A.columnindex, A.value,
                                    tc.ctypes.data_as(C.POINTER(C.c_int)), tp.ctypes.data_as(C.POINTER(C.c_int)),
                                    tv.ctypes.data_as(C.POINTER(C.c_double)))
        rtp, rtc, rtv = oracle.transpose(a)
        assert np.array_equal(tp, rtp) and np.array_equal(tc, rtc) and np.array_equal(tv, rtv)
        b = ts.Csr(a.n, a.m, tp, tc, tv)
    else:
        b = a
    B = smatrix(b, keep)
    nnzcub = ts.lib.tsg_nnzcub(C.byref(A), C.byref(B))
    assert nnzcub == oracle.nnzcub(a, b)

    ts.lib.csr2tile_row_major(C.byref(A))
    ref = oracle.csr2tile(a, packed=True)
    nt = A.numtile
    assert nt == ref["numtile"] and A.tilem == ref["tilem"] and A.tilen == ref["tilen"]
    assert np.array_equal(arr(A.tile_ptr, np.int32, A.tilem + 1), ref["tile_ptr"])
    assert np.array_equal(arr(A.tile_columnidx, np.int32, nt), ref["tile_colidx"])
    assert np.array_equal(arr(A.tile_nnz, np.int32, nt + 1), ref["tile_nnz"])
    assert np.array_equal(arr(A.tile_csr_Ptr, np.uint8, 16 * nt), ref["csr_ptr"])
    assert np.array_equal(arr(A.tile_csr_Col, np.uint8, a.nnz), ref["csr_col"])
    assert np.array_equal(arr(A.tile_csr_Value, np.float64, a.nnz), ref["val"])
    assert np.array_equal(arr(A.mask, np.uint16, 16 * nt), ref["mask"])

    ts.lib.csr2tile_col_major(C.byref(B))
    rb = oracle.csr2tile(b, packed=False)
    exp = col_major_expected(rb)
    nb = B.numtile
    assert nb == rb["numtile"]
    assert np.array_equal(arr(B.tile_ptr, np.int32, B.tilem + 1), rb["tile_ptr"])
    assert np.array_equal(arr(B.tile_columnidx, np.int32, nb), rb["tile_colidx"])
    assert np.array_equal(arr(B.csc_tile_ptr, np.int32, B.tilen + 1), exp["csc_ptr"])
    assert np.array_equal(arr(B.csc_tile_rowidx, np.int32, nb), exp["csc_rowidx"])
    assert np.array_equal(arr(B.tile_nnz, np.int32, nb + 1), exp["tile_nnz"])
    assert np.array_equal(arr(B.tile_csr_Ptr, np.uint8, 16 * nb), exp["csr_ptr"])
    assert np.array_equal(arr(B.tile_csr_Col, np.uint8, b.nnz), exp["csr_col"])
    assert np.array_equal(arr(B.tile_csr_Value, np.float64, b.nnz), exp["val"])
    assert np.array_equal(arr(B.mask, np.uint16, 16 * nb), exp["mask"])

    Cm = ts.SMatrix()
    out = [C.c_ulonglong(), C.c_double(), C.c_double(), C.c_double(), C.c_double(), C.c_double(), C.c_double(),
           C.c_double()]
    ts.lib.tilespgemm(C.byref(A), C.byref(B), C.byref(Cm), None, None, 0, 0.0, 0.0, nnzcub, C.byref(out[0]),
                      C.byref(out[1]), C.byref(out[2]), C.byref(out[3]), b"test", C.byref(out[4]), C.byref(out[5]),
                      C.byref(out[6]), C.byref(out[7]))
    rpc, cc, vc = oracle.spgemm(a, b)
    assert out[0].value == rpc[-1] == Cm.nnz
    assert out[2].value > 0 and out[3].value > 0
    cptr, ccol = oracle.step1_tiles(ref, rb)
    assert Cm.numtile == len(ccol)
    assert np.array_equal(arr(Cm.tile_ptr, np.int32, Cm.tilem + 1), cptr)
    assert np.array_equal(arr(Cm.tile_columnidx, np.int32, Cm.numtile), ccol)
    ts.lib.tile2csr(C.byref(Cm))
    assert np.array_equal(arr(Cm.rowpointer, np.int64, a.m + 1), rpc)
    assert np.array_equal(arr(Cm.columnindex, np.int32, Cm.nnz), cc)
    assert np.array_equal(arr(Cm.value, np.float64, Cm.nnz), vc)   # integer values: exact
    ts.lib.matrix_destroy(C.byref(A))
    ts.lib.matrix_destroy(C.byref(B))
    ts.lib.matrix_destroy(C.byref(Cm))
    for p in (Cm.rowpointer, Cm.columnindex, Cm.value):
        libc.free(C.cast(p, C.c_void_p))
