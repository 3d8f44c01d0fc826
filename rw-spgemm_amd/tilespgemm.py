"""Python binding of the MI355X TileSpGEMM library (librwspgemm.so).

Mirrors the C ABI of include/tsg.h and include/tilespgemm.h with ctypes.
The shared library is built in-tree (``make -C rw-spgemm_amd``); importing
this module without it raises, there is no fallback path.

Typical use::

    import tilespgemm as ts
    ctx = ts.Context(0)
    a = ts.gen_stencil27(128, 128, 128, periodic=True, valmode=ts.VAL_UNIFORM_PM1, seed=1)
    da = ctx.upload(a)
    ta = ctx.csr2tile(da); ctx.build_csc(ta)
    tc, stats = ctx.spgemm(ta, ta)          # C = A^2, B aliases A
    c = ctx.download(ctx.tile2csr(tc))
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librwspgemm.so")

TSG_OK = 0
VAL_KEEP, VAL_UNIFORM_PM1, VAL_UNIFORM_01, VAL_MOD10 = 0, 1, 2, 3


class TsgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tsg status {status}: {msg}")
        self.status = status


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run make -C rw-spgemm_amd)")
    return C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)


lib = _load()

_p = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_ull = C.c_ulonglong
_d = C.c_double


class HCsr(C.Structure):
    _fields_ = [("m", _i), ("n", _i), ("nnz", _i), ("rowptr", C.POINTER(_i)),
                ("colidx", C.POINTER(_i)), ("val", C.POINTER(_d))]


class DCsr(C.Structure):
    _fields_ = [("m", _i), ("n", _i), ("nnz", _i), ("rowptr", _p), ("colidx", _p), ("val", _p)]


class DTiles(C.Structure):
    _fields_ = [("m", _i), ("n", _i), ("tilem", _i), ("tilen", _i), ("numtile", _i), ("nnz", _i),
                ("tile_ptr", _p), ("tile_colidx", _p), ("tile_rowidx", _p), ("tile_nnz", _p),
                ("csr_ptr", _p), ("csr_col", _p), ("val", _p), ("mask", _p),
                ("csc_ptr", _p), ("csc_rowidx", _p), ("csc_tileid", _p), ("rowx_ptr", _p), ("rowx_col", _p),
                ("rowx_val", _p)]


class HTiles(C.Structure):
    _fields_ = [("m", _i), ("n", _i), ("tilem", _i), ("tilen", _i), ("numtile", _i), ("nnz", _i),
                ("tile_ptr", C.POINTER(_i)), ("tile_colidx", C.POINTER(_i)),
                ("tile_rowidx", C.POINTER(_i)), ("tile_nnz", C.POINTER(_i)),
                ("csr_ptr", C.POINTER(C.c_ubyte)), ("csr_col", C.POINTER(C.c_ubyte)),
                ("val", C.POINTER(_d)), ("mask", C.POINTER(C.c_ushort)),
                ("csc_ptr", C.POINTER(_i)), ("csc_rowidx", C.POINTER(_i)),
                ("csc_tileid", C.POINTER(_i))]


class Tune(C.Structure):
    _fields_ = [("dense_acc_min_nnz", _i), ("s1_window_words", _i), ("force_tile_parallel", _i),
                ("batch_group_nnz_cap", _i), ("element_mode", _i), ("no_heavy_rows", _i),
                ("heavy_min_products", _i), ("heavy_item_products", _i)]


class Stats(C.Structure):
    _fields_ = [("t_step1_ms", _d), ("t_step2_ms", _d), ("t_step3_ms", _d), ("t_alloc_ms", _d),
                ("t_total_ms", _d), ("numtileC", _ll), ("nnzC", _ll), ("tiles_sparse_acc", _ll),
                ("tiles_dense_acc", _ll), ("tiles_empty", _ll), ("t_step3_rows_ms", _d), ("heavy_rows", _ll),
                ("heavy_bytes", _ll)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SMatrix(C.Structure):
    """Same layout as the reference SMatrix (src/common.h:113-135)."""
    _fields_ = [("m", _i), ("n", _i), ("nnz", _i), ("isSymmetric", _i), ("value", C.POINTER(_d)),
                ("columnindex", C.POINTER(_i)), ("rowpointer", C.POINTER(_i)), ("tilem", _i),
                ("tilen", _i), ("tile_ptr", C.POINTER(_i)), ("tile_columnidx", C.POINTER(_i)),
                ("tile_rowidx", C.POINTER(_i)), ("tile_nnz", C.POINTER(_i)), ("numtile", _i),
                ("tile_csr_Value", C.POINTER(_d)), ("tile_csr_Col", C.POINTER(C.c_ubyte)),
                ("tile_csr_Ptr", C.POINTER(C.c_ubyte)), ("mask", C.POINTER(C.c_ushort)),
                ("csc_tile_ptr", C.POINTER(_i)), ("csc_tile_rowidx", C.POINTER(_i))]


class BatchInfo(C.Structure):
    _fields_ = [("count", _i), ("groups", _i), ("numtileC", _ll), ("nnzC", _ll), ("graph_nodes", _i),
                ("build_ms", _d), ("capture_ms", _d), ("tune", Tune)]


class BatchSpec(C.Structure):
    _fields_ = [("n", _i), ("family", _i), ("op", _i), ("halfbw", _i), ("mean_row", _d), ("seed", _ull)]


OP_A2, OP_AAT = 0, 1


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", _ll), ("slots", _i), ("max_batch", _i), ("small_bytes", _ll),
                ("chunk_products", _ll), ("starve_ms", _d), ("tune", Tune)]


class RequestInfo(C.Structure):
    _fields_ = [("status", _i), ("handle", _i), ("op", _i), ("chunks", _i), ("fused_with", _i),
                ("numtileC", _ll), ("nnzC", _ll), ("bytes", _ll), ("t_arrival_ms", _d), ("t_admit_ms", _d),
                ("t_done_ms", _d), ("t_gpu_ms", _d)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Csum(C.Structure):
    """tsg_csum: order-free checksums of a C tile matrix (include/tsg.h)."""
    _fields_ = [("numtile", _ll), ("nnz", _ll), ("tiles", _ull), ("entries", _ull), ("entries_t", _ull)]

    def key(self):
        return (self.numtile, self.nnz, self.tiles, self.entries)

    def add(self, o: "Csum"):
        """Chunk sums add up (mod 2^64) to the whole product's."""
        m = (1 << 64) - 1
        self.numtile += o.numtile
        self.nnz += o.nnz
        self.tiles = (self.tiles + o.tiles) & m
        self.entries = (self.entries + o.entries) & m
        self.entries_t = (self.entries_t + o.entries_t) & m
        return self


class ServerStats(C.Structure):
    _fields_ = [("resident_bytes", _ll), ("reserved_bytes", _ll), ("peak_reserved_bytes", _ll), ("queued", _i),
                ("fused_launches", _ll), ("single_launches", _ll), ("cached_bytes", _ll),
                ("peak_footprint_bytes", _ll)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("tsg_last_error", C.c_char_p, [])
_sig("tsg_build_info", C.c_char_p, [])
_sig("tsg_ctx_create", _i, [_i, C.POINTER(_p)])
_sig("tsg_ctx_destroy", None, [_p])
_sig("tsg_ctx_sync", _i, [_p])
_sig("tsg_ctx_stream", _p, [_p])
_sig("tsg_ctx_bytes_in_use", _ll, [_p])
_sig("tsg_ctx_bytes_cached", _ll, [_p])
_sig("tsg_ctx_trim", None, [_p])
_sig("tsg_hcsr_free", None, [C.POINTER(HCsr)])
_sig("tsg_hcsr_set_values", _i, [C.POINTER(HCsr), _i, _ull])
_sig("tsg_gen_stencil5", _i, [_i, _i, C.POINTER(HCsr)])
_sig("tsg_gen_stencil27", _i, [_i, _i, _i, _i, _i, _ull, C.POINTER(HCsr)])
_sig("tsg_gen_er", _i, [_i, _d, _i, _ull, C.POINTER(HCsr)])
_sig("tsg_gen_banded", _i, [_i, _i, _i, _ull, C.POINTER(HCsr)])
_sig("tsg_gen_rmat", _i, [_i, _i, _d, _d, _d, _i, _i, _ull, C.POINTER(HCsr)])
_sig("tsg_read_mtx", _i, [C.c_char_p, C.POINTER(HCsr), C.POINTER(_i)])
_sig("tsg_dcsr_upload", _i, [_p, C.POINTER(HCsr), C.POINTER(DCsr)])
_sig("tsg_dcsr_download", _i, [_p, C.POINTER(DCsr), C.POINTER(HCsr)])
_sig("tsg_dcsr_free", None, [_p, C.POINTER(DCsr)])
_sig("tsg_dcsr_transpose", _i, [_p, C.POINTER(DCsr), C.POINTER(DCsr)])
_sig("tsg_dcsr_nnzcub", _i, [_p, C.POINTER(DCsr), C.POINTER(DCsr), C.POINTER(_ull)])
_sig("tsg_csr2tile", _i, [_p, C.POINTER(DCsr), C.POINTER(DTiles)])
_sig("tsg_tiles_build_csc", _i, [_p, C.POINTER(DTiles)])
_sig("tsg_tiles_build_rowindex", _i, [_p, C.POINTER(DTiles)])
_sig("tsg_tile2csr", _i, [_p, C.POINTER(DTiles), C.POINTER(DCsr)])
_sig("tsg_tiles_free", None, [_p, C.POINTER(DTiles)])
_sig("tsg_tiles_download", _i, [_p, C.POINTER(DTiles), C.POINTER(HTiles)])
_sig("tsg_htiles_free", None, [C.POINTER(HTiles)])
_sig("tsg_tune_default", None, [C.POINTER(Tune)])
_sig("tsg_spgemm", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles), C.POINTER(Tune), C.POINTER(DTiles),
                        C.POINTER(Stats)])
_sig("tsg_tiles_packed_cols", _i, [_p, C.POINTER(DTiles), _p])
_sig("tsg_tiles_csc_order", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles)])
_sig("tsg_tiles_from_csc_order", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles)])
_sig("tsg_dmalloc", _i, [_p, C.POINTER(_p), C.c_size_t])
_sig("tsg_dfree", None, [_p, _p])
_sig("tsg_h2d", _i, [_p, _p, _p, C.c_size_t])
_sig("tsg_d2h", _i, [_p, _p, _p, C.c_size_t])
_sig("tsg_copy_kernel", _i, [_p, _p, _p, C.c_size_t])
_sig("tsg_spgemm_count_tiles", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles), C.POINTER(_ll)])
_sig("tsg_spgemm_count_tiles_noalloc", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles), _p])
NOALLOC_MAX_TILE_ROWS = 4192256
_sig("tsg_perm_validate", _i, [_p, _i, _p])
_sig("tsg_perm_invert", _i, [_p, _i, _p, _p])
_sig("tsg_order_degree", _i, [_p, C.POINTER(DCsr), _i, _p])
_sig("tsg_order_rcm", _i, [_p, C.POINTER(DCsr), _p, C.POINTER(_i)])
_sig("tsg_order_random", _i, [_p, _i, _ull, _p])
_sig("tsg_dcsr_symmetrize", _i, [_p, C.POINTER(DCsr), C.POINTER(DCsr)])
_sig("tsg_dcsr_permute", _i, [_p, C.POINTER(DCsr), _p, _p, C.POINTER(DCsr)])
# drop-in reference entry points (include/tilespgemm.h)
_sig("csr2tile_row_major", None, [C.POINTER(SMatrix)])
_sig("csr2tile_col_major", None, [C.POINTER(SMatrix)])
_sig("tile2csr", None, [C.POINTER(SMatrix)])
_sig("matrix_destroy", None, [C.POINTER(SMatrix)])
_sig("matrix_transposition", None, [_i, _i, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d),
                                    C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)])
_sig("tilespgemm", None, [C.POINTER(SMatrix), C.POINTER(SMatrix), C.POINTER(SMatrix), _p, _p, _i, _d, _d,
                          _ull, C.POINTER(_ull), C.POINTER(_d), C.POINTER(_d), C.POINTER(_d), C.c_char_p,
                          C.POINTER(_d), C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)])
_sig("tsg_nnzcub", _ull, [C.POINTER(SMatrix), C.POINTER(SMatrix)])
_sig("tsg_spgemm_chunk_plan", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles), _ll, _i, C.POINTER(_i), C.POINTER(_i),
                                   C.POINTER(_ull)])
_sig("tsg_spgemm_tile_rows", _i, [_p, C.POINTER(DTiles), C.POINTER(DTiles), _i, _i, C.POINTER(Tune),
                                  C.POINTER(DTiles), C.POINTER(Stats)])
_sig("tsg_batch_create", _i, [_p, C.POINTER(DCsr), C.POINTER(_i), _i, C.POINTER(Tune), C.POINTER(_p)])
_sig("tsg_batch_run", _i, [_p])
_sig("tsg_batch_autotune", _i, [_p, _i, C.POINTER(Tune), C.POINTER(_d)])
_sig("tsg_batch_get_info", _i, [_p, C.POINTER(BatchInfo)])
_sig("tsg_batch_result", _i, [_p, _i, C.POINTER(HTiles), C.POINTER(HCsr)])
_sig("tsg_batch_destroy", None, [_p])
_sig("tsg_batch_result_checksum", _i, [_p, _i, C.POINTER(Csum)])
_sig("tsg_tiles_checksum", _i, [_p, C.POINTER(DTiles), _ll, _i, C.POINTER(Csum)])
_sig("tsg_tiles_rows", _i, [_p, C.POINTER(DTiles), _i, C.POINTER(_i), _i, C.POINTER(HCsr)])
_sig("tsg_server_result_checksum", _i, [_p, _ull, C.POINTER(Csum)])
_sig("tsg_server_result_rows", _i, [_p, _ull, C.POINTER(_i), _i, C.POINTER(HCsr)])
_sig("tsg_server_config_default", None, [C.POINTER(ServerConfig)])
_sig("tsg_server_create", _i, [_i, C.POINTER(ServerConfig), C.POINTER(_p)])
_sig("tsg_server_destroy", None, [_p])
_sig("tsg_server_register", _i, [_p, C.POINTER(HCsr), C.POINTER(_i)])
_sig("tsg_server_size", _i, [_p, _i, _i, C.POINTER(_ll), C.POINTER(_ll), C.POINTER(_ll), C.POINTER(_i)])
_sig("tsg_server_submit", _i, [_p, _i, _i, C.POINTER(_ull)])
_sig("tsg_server_wait", _i, [_p, _ull, _d, C.POINTER(RequestInfo)])
_sig("tsg_server_result_csr", _i, [_p, _ull, C.POINTER(HCsr)])
_sig("tsg_server_release", _i, [_p, _ull])
_sig("tsg_server_stats", _i, [_p, C.POINTER(ServerStats)])
_sig("tsg_server_replay", _i, [_p, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d), _i, C.POINTER(_ull),
                               C.POINTER(RequestInfo)])
REPLAY_PROGRESS = C.CFUNCTYPE(None, _i, _i, _p)
_sig("tsg_server_replay_progress", _i, [_p, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d), _i, C.POINTER(_ull),
                                        C.POINTER(RequestInfo), REPLAY_PROGRESS, _p])
_sig("tsg_gen_batch_spec", _i, [_i, _ull, _i, _i, C.POINTER(BatchSpec)])
_sig("tsg_gen_batch_problem", _i, [C.POINTER(BatchSpec), C.POINTER(HCsr)])
_sig("tsg_compat_set_device", _i, [_i])
_sig("tsg_hcsr_canonicalize", _i, [C.POINTER(HCsr)])
_sig("mmio_allinone", _i, [C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(C.POINTER(_i)),
                           C.POINTER(C.POINTER(_i)), C.POINTER(C.POINTER(_d)), C.c_char_p])
_sig("tsg_driver_main", _i, [_i, C.POINTER(C.c_char_p)])


class AxisPermutation(C.Structure):
    """Same layout as src/reorder.h:29-34."""
    _fields_ = [("n", _i), ("old_id", C.POINTER(_i)), ("new_id", C.POINTER(_i))]


class ReorderInfo(C.Structure):
    """Same layout as src/reorder.h:36-63."""
    _fields_ = [("enabled", _i), ("n", _i), ("row_perm", AxisPermutation), ("inner_perm", AxisPermutation),
                ("col_perm", AxisPermutation), ("input_tiles_A_before", _ll), ("input_tiles_A_after", _ll),
                ("input_tiles_B_before", _ll), ("input_tiles_B_after", _ll), ("estimated_c_tiles_before", _ll),
                ("estimated_c_tiles_after", _ll), ("reorder_time_ms", _d), ("signature_time_ms", _d),
                ("tile_profile_time_ms", _d), ("affinity_time_ms", _d), ("sort_time_ms", _d),
                ("local_refine_time_ms", _d), ("proxy_time_ms", _d), ("permute_time_ms", _d),
                ("risk_input_tile_ratio", _d), ("risk_c_tile_ratio", _d), ("local_refine_window", _i),
                ("proxy_sample_rows", _i), ("guard_degraded_to_identity", _i), ("structural_conservative", _i),
                ("shared_symmetric_csr", _i), ("mode", C.c_char * 64), ("risk_reason", C.c_char * 128)]


class ReorderProblem(C.Structure):
    _fields_ = [("kind", _i), ("matrixA", C.POINTER(SMatrix)), ("matrixB", C.POINTER(SMatrix))]


PROBLEM_AA, PROBLEM_AAT, PROBLEM_AB = 0, 1, 2

_sig("axis_permutation_init", None, [C.POINTER(AxisPermutation)])
_sig("axis_permutation_destroy", None, [C.POINTER(AxisPermutation)])
_sig("axis_permutation_alloc", _i, [C.POINTER(AxisPermutation), _i])
_sig("reorder_info_init", None, [C.POINTER(ReorderInfo)])
_sig("reorder_info_destroy", None, [C.POINTER(ReorderInfo)])
_sig("read_external_order_file", _i, [C.c_char_p, _i, C.POINTER(AxisPermutation)])
_sig("build_external_reorder_plan", _i, [_i, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(SMatrix),
                                         C.POINTER(SMatrix), C.POINTER(ReorderInfo), C.POINTER(_d)])
_sig("build_original_reorder_plan", _i, [_i, C.POINTER(SMatrix), C.POINTER(SMatrix), C.POINTER(ReorderInfo)])
_sig("build_reorder_plan", _i, [C.POINTER(ReorderProblem), C.POINTER(ReorderInfo)])
_sig("apply_bipermutation_csr", _i, [C.POINTER(SMatrix), C.POINTER(AxisPermutation), C.POINTER(AxisPermutation),
                                     C.POINTER(SMatrix)])
_sig("restore_bipermutation_csr", _i, [C.POINTER(SMatrix), C.POINTER(AxisPermutation), C.POINTER(AxisPermutation)])


def _check(status):
    if status != TSG_OK:
        raise TsgError(status, lib.tsg_last_error().decode(errors="replace"))


def build_info():
    return lib.tsg_build_info().decode()


# ---------------------------------------------------------------- host CSR
class Csr:
    """Host CSR held in numpy arrays (int32 rowptr/colidx, float64 val)."""

    def __init__(self, m, n, rowptr, colidx, val):
        self.m, self.n = int(m), int(n)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        assert self.rowptr.shape == (self.m + 1,)

    @property
    def nnz(self):
        return int(self.colidx.shape[0])

    def as_c(self):
        h = HCsr(self.m, self.n, self.nnz, self.rowptr.ctypes.data_as(C.POINTER(_i)),
                 self.colidx.ctypes.data_as(C.POINTER(_i)), self.val.ctypes.data_as(C.POINTER(_d)))
        return h

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.colidx, self.rowptr), shape=(self.m, self.n))

    @staticmethod
    def from_scipy(a):
        a = a.tocsr()
        a.sort_indices()
        return Csr(a.shape[0], a.shape[1], a.indptr, a.indices, a.data)


def _take_hcsr(h):
    """Copy a library-allocated HCsr into numpy and free it."""
    try:
        m, n, nnz = h.m, h.n, h.nnz
        rp = np.ctypeslib.as_array(h.rowptr, shape=(m + 1,)).copy()
        if nnz:
            ci = np.ctypeslib.as_array(h.colidx, shape=(nnz,)).copy()
            va = np.ctypeslib.as_array(h.val, shape=(nnz,)).copy()
        else:
            ci = np.zeros(0, np.int32)
            va = np.zeros(0, np.float64)
        return Csr(m, n, rp, ci, va)
    finally:
        lib.tsg_hcsr_free(C.byref(h))


def _take_htiles(h) -> dict:
    """Copy a library-allocated HTiles into numpy arrays and free it."""
    try:
        nt = h.numtile
        out = dict(m=h.m, n=h.n, tilem=h.tilem, tilen=h.tilen, numtile=nt, nnz=h.nnz)
        out["tile_ptr"] = _arr(h.tile_ptr, np.int32, h.tilem + 1)
        out["tile_colidx"] = _arr(h.tile_colidx, np.int32, nt)
        out["tile_rowidx"] = _arr(h.tile_rowidx, np.int32, nt)
        out["tile_nnz"] = _arr(h.tile_nnz, np.int32, nt + 1)
        out["csr_ptr"] = _arr(h.csr_ptr, np.uint8, nt * 16)
        out["csr_col"] = _arr(h.csr_col, np.uint8, h.nnz)
        out["val"] = _arr(h.val, np.float64, h.nnz)
        out["mask"] = _arr(h.mask, np.uint16, nt * 16)
        out["csc_ptr"] = _arr(h.csc_ptr, np.int32, h.tilen + 1) if h.csc_ptr else None
        out["csc_rowidx"] = _arr(h.csc_rowidx, np.int32, nt) if h.csc_rowidx else None
        out["csc_tileid"] = _arr(h.csc_tileid, np.int32, nt) if h.csc_tileid else None
        return out
    finally:
        lib.tsg_htiles_free(C.byref(h))


def gen_batch_spec(count, seed=2, nmin=1000, nmax=65536):
    """The batched workload's problem list (tsg_gen_batch_spec)."""
    arr = (BatchSpec * max(1, count))()
    _check(lib.tsg_gen_batch_spec(count, seed, nmin, nmax, arr))
    return [arr[i] for i in range(count)]


def gen_batch_problem(spec: BatchSpec) -> Csr:
    h = HCsr()
    _check(lib.tsg_gen_batch_problem(C.byref(spec), C.byref(h)))
    return _take_hcsr(h)


class Batch:
    """Many SpGEMMs per step (tsg_batch): packed block-diagonally, planned
    once, one HIP graph launch per run."""

    def __init__(self, ctx: "Context", problems, ops, tune: Tune | None = None):
        self.ctx = ctx
        n = len(problems)
        arr = (DCsr * n)(*problems)
        o = (_i * n)(*[int(x) for x in ops])
        p = _p()
        _check(lib.tsg_batch_create(ctx.h, arr, o, n, C.byref(tune) if tune else None, C.byref(p)))
        self.h = p

    def run(self):
        _check(lib.tsg_batch_run(self.h))

    def autotune(self, reps=3):
        t = Tune()
        ms = _d()
        _check(lib.tsg_batch_autotune(self.h, reps, C.byref(t), C.byref(ms)))
        return t, ms.value

    def info(self) -> BatchInfo:
        bi = BatchInfo()
        _check(lib.tsg_batch_get_info(self.h, C.byref(bi)))
        return bi

    def result_tiles(self, i) -> dict:
        h = HTiles()
        _check(lib.tsg_batch_result(self.h, i, C.byref(h), None))
        return _take_htiles(h)

    def result_csr(self, i) -> "Csr":
        h = HCsr()
        _check(lib.tsg_batch_result(self.h, i, None, C.byref(h)))
        return _take_hcsr(h)

    def result_checksum(self, i) -> Csum:
        cs = Csum()
        _check(lib.tsg_batch_result_checksum(self.h, i, C.byref(cs)))
        return cs

    def close(self):
        if self.h:
            lib.tsg_batch_destroy(self.h)
            self.h = None


def server_config(**kw) -> ServerConfig:
    c = ServerConfig()
    lib.tsg_server_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Server:
    """Continuous-batching SpGEMM server (tsg_server) on one device."""

    def __init__(self, device=0, config: ServerConfig | None = None):
        p = _p()
        _check(lib.tsg_server_create(device, C.byref(config) if config else None, C.byref(p)))
        self.h = p

    def register(self, csr: Csr) -> int:
        h = csr.as_c()
        out = _i()
        _check(lib.tsg_server_register(self.h, C.byref(h), C.byref(out)))
        return out.value

    def size(self, handle, op):
        t, n, b, c = _ll(), _ll(), _ll(), _i()
        _check(lib.tsg_server_size(self.h, handle, op, C.byref(t), C.byref(n), C.byref(b), C.byref(c)))
        return {"numtileC": t.value, "nnzC": n.value, "bytes": b.value, "chunks": c.value}

    def submit(self, handle, op) -> int:
        rid = _ull()
        _check(lib.tsg_server_submit(self.h, handle, op, C.byref(rid)))
        return rid.value

    def wait(self, rid, timeout_ms=0.0) -> RequestInfo:
        info = RequestInfo()
        _check(lib.tsg_server_wait(self.h, rid, timeout_ms, C.byref(info)))
        return info

    def poll(self, rid, timeout_ms) -> RequestInfo | None:
        """The request's info once done, or None if it is not done within
        timeout_ms (> 0)."""
        info = RequestInfo()
        info.handle = -1  # filled only when the request is done
        st = lib.tsg_server_wait(self.h, rid, float(timeout_ms), C.byref(info))
        if info.handle < 0:
            if st == -6:  # TSG_ERR_STATE: not done yet
                return None
            _check(st)
        return info

    def result_csr(self, rid) -> Csr:
        h = HCsr()
        _check(lib.tsg_server_result_csr(self.h, rid, C.byref(h)))
        return _take_hcsr(h)

    def result_checksum(self, rid) -> Csum:
        cs = Csum()
        _check(lib.tsg_server_result_checksum(self.h, rid, C.byref(cs)))
        return cs

    def result_rows(self, rid, rows) -> Csr:
        """Rows `rows` of a finished request's C (a CSR with len(rows) rows)."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        h = HCsr()
        _check(lib.tsg_server_result_rows(self.h, rid, r.ctypes.data_as(C.POINTER(_i)), len(r), C.byref(h)))
        return _take_hcsr(h)

    def release(self, rid):
        _check(lib.tsg_server_release(self.h, rid))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib.tsg_server_stats(self.h, C.byref(st)))
        return st

    def replay(self, handles, ops, arrival_ms, keep=False, progress=None):
        """Open-loop load (tsg_server_replay); progress(submitted, completed)
        is called as requests are submitted and completed."""
        n = len(handles)
        hs = (_i * n)(*handles)
        os_ = (_i * n)(*ops)
        ar = (_d * n)(*arrival_ms)
        ids = (_ull * n)()
        infos = (RequestInfo * n)()
        if progress is None:
            _check(lib.tsg_server_replay(self.h, n, hs, os_, ar, 1 if keep else 0, ids, infos))
        else:
            cb = REPLAY_PROGRESS(lambda sub, done, _u: progress(sub, done))
            _check(lib.tsg_server_replay_progress(self.h, n, hs, os_, ar, 1 if keep else 0, ids, infos, cb, None))
        return [ids[i] for i in range(n)], [infos[i] for i in range(n)]

    def close(self):
        if self.h:
            lib.tsg_server_destroy(self.h)
            self.h = None


def gen_stencil5(nx, ny):
    h = HCsr()
    _check(lib.tsg_gen_stencil5(nx, ny, C.byref(h)))
    return _take_hcsr(h)


def gen_stencil27(nx, ny, nz, periodic=True, valmode=VAL_UNIFORM_PM1, seed=1):
    h = HCsr()
    _check(lib.tsg_gen_stencil27(nx, ny, nz, int(periodic), valmode, seed, C.byref(h)))
    return _take_hcsr(h)


def gen_er(n, mean_row, valmode=VAL_UNIFORM_PM1, seed=1):
    h = HCsr()
    _check(lib.tsg_gen_er(n, float(mean_row), valmode, seed, C.byref(h)))
    return _take_hcsr(h)


def gen_banded(n, halfbw, valmode=VAL_UNIFORM_PM1, seed=1):
    h = HCsr()
    _check(lib.tsg_gen_banded(n, halfbw, valmode, seed, C.byref(h)))
    return _take_hcsr(h)


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, symmetrize=True, valmode=VAL_UNIFORM_01, seed=1):
    h = HCsr()
    _check(lib.tsg_gen_rmat(scale, edge_factor, a, b, c, int(symmetrize), valmode, seed, C.byref(h)))
    return _take_hcsr(h)


def set_values(csr, mode, seed=1):
    h = csr.as_c()
    _check(lib.tsg_hcsr_set_values(C.byref(h), mode, seed))
    return csr


def read_mtx(path):
    """Canonical CSR of a Matrix Market file (rows sorted, duplicates
    summed) and the header's symmetric flag."""
    h = HCsr()
    sym = _i(0)
    _check(lib.tsg_read_mtx(path.encode(), C.byref(h), C.byref(sym)))
    return _take_hcsr(h), bool(sym.value)


def mmio_allinone(path):
    """The reference loader (src/mmio_highlevel.h:593-759): returns
    (status, csr, symmetric); rows in file order, duplicates kept."""
    m, n, nnz, sym = _i(0), _i(0), _i(0), _i(0)
    rp, ci, va = C.POINTER(_i)(), C.POINTER(_i)(), C.POINTER(_d)()
    st = lib.mmio_allinone(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym), C.byref(rp), C.byref(ci),
                           C.byref(va), path.encode())
    if st != 0:
        return st, None, False
    h = HCsr(m.value, n.value, nnz.value, rp, ci, va)
    return 0, _take_hcsr(h), bool(sym.value)


def canonicalize(csr: Csr) -> Csr:
    """Rows sorted by column, duplicates summed (tsg_hcsr_canonicalize)."""
    rp, ci, va = csr.rowptr.copy(), csr.colidx.copy(), csr.val.copy()
    h = HCsr(csr.m, csr.n, csr.nnz, rp.ctypes.data_as(C.POINTER(_i)), ci.ctypes.data_as(C.POINTER(_i)),
             va.ctypes.data_as(C.POINTER(_d)))
    _check(lib.tsg_hcsr_canonicalize(C.byref(h)))
    return Csr(csr.m, csr.n, rp, ci[:h.nnz], va[:h.nnz])


class HostSMatrix:
    """An SMatrix whose CSR fields point at numpy arrays kept alive here."""

    def __init__(self, csr: Csr, symmetric=False):
        self.csr = csr
        self.s = SMatrix()
        self.s.m, self.s.n, self.s.nnz, self.s.isSymmetric = csr.m, csr.n, csr.nnz, int(symmetric)
        self.s.rowpointer = csr.rowptr.ctypes.data_as(C.POINTER(_i))
        self.s.columnindex = csr.colidx.ctypes.data_as(C.POINTER(_i))
        self.s.value = csr.val.ctypes.data_as(C.POINTER(_d))

    @property
    def ref(self):
        return C.byref(self.s)


def perm_arrays(p: AxisPermutation):
    """(old_id, new_id) numpy copies of an AxisPermutation."""
    if p.n <= 0 or not p.old_id:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    return (np.ctypeslib.as_array(p.old_id, shape=(p.n,)).copy(),
            np.ctypeslib.as_array(p.new_id, shape=(p.n,)).copy())


def read_order_file(path, n):
    """read_external_order_file: (status, old_id, new_id)."""
    p = AxisPermutation()
    lib.axis_permutation_init(C.byref(p))
    try:
        st = lib.read_external_order_file(path.encode() if path is not None else None, n, C.byref(p))
        old, new = perm_arrays(p) if st == 0 else (None, None)
        return st, old, new
    finally:
        lib.axis_permutation_destroy(C.byref(p))


def reorder_plan(kind, algorithm, a: HostSMatrix | None, b: HostSMatrix | None = None, path_a=None, path_b=None,
                 builder="external"):
    """build_external_reorder_plan / build_original_reorder_plan /
    build_reorder_plan: (status, dict of the ReorderInfo fields with numpy
    permutations, ordergen_ms)."""
    info = ReorderInfo()
    lib.reorder_info_init(C.byref(info))
    ms = _d(0.0)
    pa = a.ref if a is not None else None
    pb = b.ref if b is not None else None
    try:
        if builder == "external":
            st = lib.build_external_reorder_plan(kind, algorithm.encode() if algorithm is not None else None,
                                                 path_a.encode() if path_a else None,
                                                 path_b.encode() if path_b else None, pa, pb, C.byref(info),
                                                 C.byref(ms))
        elif builder == "original":
            st = lib.build_original_reorder_plan(kind, pa, pb, C.byref(info))
        else:
            prob = ReorderProblem(kind, C.pointer(a.s) if a is not None else None,
                                  C.pointer(b.s) if b is not None else None)
            st = lib.build_reorder_plan(C.byref(prob), C.byref(info))
        out = {k: getattr(info, k) for k, _ in ReorderInfo._fields_ if not k.endswith("_perm")}
        out["mode"] = info.mode.decode()
        out["risk_reason"] = info.risk_reason.decode()
        for ax in ("row_perm", "inner_perm", "col_perm"):
            out[ax] = perm_arrays(getattr(info, ax))
        return st, out, ms.value
    finally:
        lib.reorder_info_destroy(C.byref(info))


def driver(args):
    """tsg_driver_main(["tsg_driver", *args]) in this process; returns its
    exit code."""
    argv = [b"tsg_driver"] + [str(x).encode() for x in args]
    arr = (C.c_char_p * (len(argv) + 1))(*argv, None)
    return lib.tsg_driver_main(len(argv), arr)


def read_csr_bin(path) -> Csr:
    """C written by the driver's -o option."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"TSGCSR1\0", raw[:8]
    m, n, nnz = np.frombuffer(raw, np.int64, 3, 8)
    o = 32
    rp = np.frombuffer(raw, np.int32, m + 1, o)
    o += 4 * (m + 1)
    ci = np.frombuffer(raw, np.int32, nnz, o)
    o += 4 * nnz
    va = np.frombuffer(raw, np.float64, nnz, o)
    return Csr(m, n, rp, ci, va)


# ---------------------------------------------------------------- device
def _arr(ptr, dtype, n):
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    ctype = np.ctypeslib.as_ctypes_type(dtype)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()


class Context:
    """One device, one HIP stream, a caching allocator (tsg_ctx)."""

    def __init__(self, device=0):
        p = _p()
        _check(lib.tsg_ctx_create(device, C.byref(p)))
        self.h = p

    def close(self):
        if self.h:
            lib.tsg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(lib.tsg_ctx_sync(self.h))

    def trim(self):
        """Return the allocator's cached free blocks to the device."""
        lib.tsg_ctx_trim(self.h)

    def bytes_in_use(self):
        return lib.tsg_ctx_bytes_in_use(self.h)

    # CSR
    def upload(self, csr: Csr) -> DCsr:
        h = csr.as_c()
        d = DCsr()
        _check(lib.tsg_dcsr_upload(self.h, C.byref(h), C.byref(d)))
        return d

    def download(self, d: DCsr) -> Csr:
        h = HCsr()
        _check(lib.tsg_dcsr_download(self.h, C.byref(d), C.byref(h)))
        return _take_hcsr(h)

    def free(self, obj):
        if isinstance(obj, DCsr):
            lib.tsg_dcsr_free(self.h, C.byref(obj))
        elif isinstance(obj, DTiles):
            lib.tsg_tiles_free(self.h, C.byref(obj))

    def transpose(self, d: DCsr) -> DCsr:
        t = DCsr()
        _check(lib.tsg_dcsr_transpose(self.h, C.byref(d), C.byref(t)))
        return t

    def nnzcub(self, a: DCsr, b: DCsr) -> int:
        v = _ull(0)
        _check(lib.tsg_dcsr_nnzcub(self.h, C.byref(a), C.byref(b), C.byref(v)))
        return int(v.value)

    # tiles
    def csr2tile(self, d: DCsr) -> DTiles:
        t = DTiles()
        _check(lib.tsg_csr2tile(self.h, C.byref(d), C.byref(t)))
        return t

    def build_rowindex(self, t: DTiles):
        _check(lib.tsg_tiles_build_rowindex(self.h, C.byref(t)))
        return t

    def build_csc(self, t: DTiles):
        _check(lib.tsg_tiles_build_csc(self.h, C.byref(t)))
        return t

    def tile2csr(self, t: DTiles) -> DCsr:
        d = DCsr()
        _check(lib.tsg_tile2csr(self.h, C.byref(t), C.byref(d)))
        return d

    def csc_order(self, t: DTiles) -> DTiles:
        o = DTiles()
        _check(lib.tsg_tiles_csc_order(self.h, C.byref(t), C.byref(o)))
        return o

    def from_csc_order(self, t: DTiles) -> DTiles:
        o = DTiles()
        _check(lib.tsg_tiles_from_csc_order(self.h, C.byref(t), C.byref(o)))
        return o

    def packed_cols(self, t: DTiles) -> np.ndarray:
        p = _p()
        _check(lib.tsg_dmalloc(self.h, C.byref(p), max(1, t.nnz)))
        try:
            _check(lib.tsg_tiles_packed_cols(self.h, C.byref(t), p))
            out = np.zeros(t.nnz, np.uint8)
            if t.nnz:
                _check(lib.tsg_d2h(self.h, out.ctypes.data, p, t.nnz))
            return out
        finally:
            lib.tsg_dfree(self.h, p)

    def download_tiles(self, t: DTiles) -> dict:
        h = HTiles()
        _check(lib.tsg_tiles_download(self.h, C.byref(t), C.byref(h)))
        return _take_htiles(h)

    # raw device int arrays
    def dev_ints(self, host: np.ndarray | None = None, n: int | None = None):
        n = len(host) if host is not None else n
        p = _p()
        _check(lib.tsg_dmalloc(self.h, C.byref(p), max(1, n) * 4))
        if host is not None and n:
            h = np.ascontiguousarray(host, dtype=np.int32)
            _check(lib.tsg_h2d(self.h, p, h.ctypes.data, n * 4))
        return p

    def host_ints(self, p, n):
        out = np.zeros(n, np.int32)
        if n:
            _check(lib.tsg_d2h(self.h, out.ctypes.data, p, n * 4))
        return out

    def dfree(self, p):
        lib.tsg_dfree(self.h, p)

    def copy_kernel_run(self, nbytes: int):
        """One launch of the plain copy kernel over nbytes (byte-counter
        reference)."""
        s, d = _p(), _p()
        _check(lib.tsg_dmalloc(self.h, C.byref(s), nbytes))
        _check(lib.tsg_dmalloc(self.h, C.byref(d), nbytes))
        try:
            _check(lib.tsg_copy_kernel(self.h, d, s, nbytes))
            self.sync()
        finally:
            lib.tsg_dfree(self.h, s)
            lib.tsg_dfree(self.h, d)

    def count_tiles(self, a: DTiles, b: DTiles) -> int:
        v = _ll(0)
        _check(lib.tsg_spgemm_count_tiles(self.h, C.byref(a), C.byref(b), C.byref(v)))
        return int(v.value)

    def count_tiles_noalloc(self, a: DTiles, b: DTiles, d_out) -> None:
        """Enqueues tsg_spgemm_count_tiles_noalloc: the count lands in the
        caller's device int64 d_out (no allocation, no synchronization)."""
        _check(lib.tsg_spgemm_count_tiles_noalloc(self.h, C.byref(a), C.byref(b), d_out))

    def order_degree(self, d: DCsr, symmetric_header=False) -> np.ndarray:
        p = self.dev_ints(n=d.m)
        try:
            _check(lib.tsg_order_degree(self.h, C.byref(d), int(symmetric_header), p))
            return self.host_ints(p, d.m)
        finally:
            self.dfree(p)

    def order_rcm(self, d: DCsr):
        p = self.dev_ints(n=d.m)
        lv = _i(0)
        try:
            _check(lib.tsg_order_rcm(self.h, C.byref(d), p, C.byref(lv)))
            return self.host_ints(p, d.m), int(lv.value)
        finally:
            self.dfree(p)

    def order_random(self, n, seed) -> np.ndarray:
        p = self.dev_ints(n=n)
        try:
            _check(lib.tsg_order_random(self.h, n, seed, p))
            return self.host_ints(p, n)
        finally:
            self.dfree(p)

    def symmetrize(self, d: DCsr) -> DCsr:
        g = DCsr()
        _check(lib.tsg_dcsr_symmetrize(self.h, C.byref(d), C.byref(g)))
        return g

    def permute(self, d: DCsr, row_old: np.ndarray, col_new: np.ndarray) -> DCsr:
        pr, pc = self.dev_ints(row_old), self.dev_ints(col_new)
        try:
            o = DCsr()
            _check(lib.tsg_dcsr_permute(self.h, C.byref(d), pr, pc, C.byref(o)))
            return o
        finally:
            self.dfree(pr)
            self.dfree(pc)

    def chunk_plan(self, a: DTiles, b: DTiles, max_products: int, max_chunks: int = 4096):
        """Tile-row bounds of chunks with at most max_products products each;
        returns (bounds list, total products)."""
        bounds = (_i * (max_chunks + 1))()
        n = _i()
        tot = _ull()
        _check(lib.tsg_spgemm_chunk_plan(self.h, C.byref(a), C.byref(b), max_products, max_chunks, bounds,
                                         C.byref(n), C.byref(tot)))
        return [bounds[i] for i in range(n.value + 1)], tot.value

    def checksum(self, c: DTiles, row_base=0, jbase=0) -> Csum:
        """Order-free checksums of C (tsg_tiles_checksum); a tile-row chunk's
        row_base is 16 * its first tile-row."""
        cs = Csum()
        _check(lib.tsg_tiles_checksum(self.h, C.byref(c), row_base, jbase, C.byref(cs)))
        return cs

    def tile_rows_csr(self, c: DTiles, rows, jbase=0) -> Csr:
        """Rows `rows` (local to c) of a C tile matrix as a CSR of len(rows) rows."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        h = HCsr()
        _check(lib.tsg_tiles_rows(self.h, C.byref(c), jbase, r.ctypes.data_as(C.POINTER(_i)), len(r), C.byref(h)))
        return _take_hcsr(h)

    def spgemm_tile_rows(self, a: DTiles, b: DTiles, r0: int, r1: int, tune: Tune | None = None, timed=True):
        c = DTiles()
        st = Stats()
        _check(lib.tsg_spgemm_tile_rows(self.h, C.byref(a), C.byref(b), r0, r1, C.byref(tune) if tune else None,
                                        C.byref(c), C.byref(st) if timed else None))
        return c, st

    def spgemm(self, a: DTiles, b: DTiles, tune: Tune | None = None, timed=True):
        c = DTiles()
        st = Stats()
        _check(lib.tsg_spgemm(self.h, C.byref(a), C.byref(b), C.byref(tune) if tune else None, C.byref(c),
                              C.byref(st) if timed else None))
        return c, st


def default_tune() -> Tune:
    t = Tune()
    lib.tsg_tune_default(C.byref(t))
    return t
