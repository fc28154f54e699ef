"""ctypes binding of libmfamd.so, the MI355X Mille-feuille solver library.

The library itself (C-ABI in include/mf.h) is the product: HIP kernels for
gfx950 and C++ host code.  This module only moves numpy arrays across that
boundary for the tests and the benchmark.  It raises if the library is
missing or if a call returns an error -- there is no fallback path.
"""
import atexit
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfamd.so")

MF_CG, MF_BICGSTAB = 0, 1
MF_FP64, MF_MIXED = 0, 1
MF_REORDER_AUTO, MF_REORDER_OFF, MF_REORDER_ON = 0, 1, 2
MF_SYNC_AUTO, MF_SYNC_GLOBAL = 0, 1
MF_VECTORS_AUTO, MF_VECTORS_HBM, MF_VECTORS_ONCHIP = 0, 1, 2
CODE_FP64, CODE_FP8, CODE_FP16, CODE_FP32 = 0, 2, 3, 4
ALLOW_ALL_LOW = (1 << 2) | (1 << 3) | (1 << 4)

_i = C.POINTER(C.c_int)
_u8 = C.POINTER(C.c_ubyte)
_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)


class MFError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class RefTiles(C.Structure):
    _fields_ = [
        ("B", C.c_int), ("rowA", C.c_int), ("colA", C.c_int), ("tilem", C.c_int),
        ("tilen", C.c_int), ("tilenum", C.c_int), ("csrsize", C.c_int), ("csrptrlen", C.c_int),
        ("tile_ptr", _i), ("tile_columnidx", _i), ("tile_nnz", _i), ("Format", C.POINTER(C.c_char)),
        ("blknnz", _i), ("blknnznnz", _u8), ("csr_offset", _i), ("csrptr_offset", _i),
        ("Blockcsr_Val", _d), ("Blockcsr_Val_Low", _f), ("Blockcsr_Ptr", _u8),
        ("Tile_csr_Col", _u8), ("csr_compressedIdx", _u8),
    ]


class Info(C.Structure):
    _fields_ = [("rowA", C.c_int), ("colA", C.c_int), ("nnz", C.c_int), ("B", C.c_int),
                ("tilem", C.c_int), ("tilen", C.c_int), ("tilenum", C.c_int),
                ("tiles_per_code", C.c_longlong * 5), ("nnz_per_code", C.c_longlong * 5),
                ("bytes_alg_fp64", C.c_double), ("bytes_alg_mixed", C.c_double),
                ("bytes_layout_mixed", C.c_double), ("index_width", C.c_int),
                ("pattern_slabs_fp64", C.c_longlong), ("pattern_slabs_mixed", C.c_longlong)]


class Opts(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("tol", C.c_double), ("maxit", C.c_int),
                ("tau", C.c_double), ("allow_mask", C.c_int), ("rr_factor", C.c_double),
                ("ir_max_outer", C.c_int), ("ir_inner_tol", C.c_double), ("grid", C.c_int),
                ("block", C.c_int), ("spin_limit", C.c_longlong), ("index_width", C.c_int), ("profile", C.c_int),
                ("reorder", C.c_int), ("sync", C.c_int), ("vectors", C.c_int)]


class Report(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("outer", C.c_int),
                ("replacements", C.c_int), ("relres_true", C.c_double), ("ms_solve", C.c_double),
                ("ms_setup", C.c_double), ("grid", C.c_int), ("block", C.c_int),
                ("rows_per_thread", C.c_int), ("phase_us", C.c_double * 5), ("reordered", C.c_int),
                ("pad_factor", C.c_double), ("sync_deps", C.c_int), ("streamed", C.c_int),
                ("fallback", C.c_int), ("split_rows", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["phase_us"] = list(self.phase_us)
        return d


class Mem(C.Structure):
    _fields_ = [("ref_memory_MB", C.c_double), ("ref_memory_CP_MB", C.c_double),
                ("csr_fp64_MB", C.c_double), ("layout_fp64_MB", C.c_double),
                ("layout_mixed_MB", C.c_double), ("index_width", C.c_int),
                ("nnz_per_code", C.c_longlong * 5)]


class Compact(C.Structure):
    _fields_ = [("cnt_non_new", C.c_int), ("nonzero_row_new", _i), ("blockrowid_new", _u8),
                ("blockcsr_ptr_new", _u8), ("block_signal", _i), ("tile_rowidx", _i)]


class Partition(C.Structure):
    _fields_ = [("index", C.c_int), ("each_block_nnz", C.c_int), ("block_per_warp", C.c_int),
                ("vector_each_warp_32", C.c_int), ("vector_total_32", C.c_int),
                ("balance_tile_ptr", _i), ("index_each_block", _i), ("row_each_block", _i),
                ("non_each_block_offset", _i), ("balance_tile_ptr_shared_end", _i)]


class TileMatrix(C.Structure):
    """The reference's Tile_matrix record (include/mf_tile_matrix.h, field
    order of PC/format.h:14-86), for Tile_create/Tile_destroy/blockspmv_cpu."""
    _p = C.c_void_p
    _fields_ = [
        ("tilem", C.c_int), ("tilen", C.c_int), ("tilenum", C.c_int), ("tile_ptr", _i), ("tile_columnidx", _i),
        ("tile_nnz", _i), ("Format", C.POINTER(C.c_char)), ("blknnz", _i), ("blknnznnz", _u8),
        ("dnsrowptr", _p), ("dnscolptr", _p), ("tilewidth", C.POINTER(C.c_char)), ("csr_offset", _i),
        ("csrptr_offset", _i), ("coo_offset", _p), ("ell_offset", _p), ("hyb_offset", _p), ("hyb_coocount", _p),
        ("dns_offset", _p), ("dnsrow_offset", _p), ("dnscol_offset", _p), ("new_coocount", _p),
        ("Blockcsr_Val", _d), ("Blockcsr_Val_Low", _f), ("Tile_csr_Col", _u8), ("Blockcsr_Val_mix", _p),
        ("Blockcsr_Ptr", _u8), ("csr_compressedIdx", _u8), ("csrsize", C.c_int), ("csrptrlen", C.c_int),
        ("Blockcoo_Val", _p), ("Blockcoo_Val_Low", _p), ("coo_compressed_Idx", _p), ("coosize", C.c_int),
        ("Blockell_Val", _p), ("Blockell_Val_Low", _p), ("ell_compressedIdx", _p), ("ellsize", C.c_int),
        ("Blockhyb_Val", _p), ("Blockhyb_Val_Low", _p), ("hybIdx", _p), ("hybsize", C.c_int),
        ("hybellsize", C.c_int), ("hybcoosize", C.c_int), ("Blockdense_Val", _p), ("Blockdense_Val_Low", _p),
        ("dnssize", C.c_int), ("Blockdenserow_Val", _p), ("Blockdenserow_Val_Low", _p), ("denserowid", _p),
        ("dnsrowsize", C.c_int), ("Blockdensecol_Val", _p), ("Blockdensecol_Val_Low", _p), ("densecolid", _p),
        ("dnscolsize", C.c_int), ("coototal", C.c_int), ("deferredcoo_val", _p), ("deferredcoo_val_Low", _p),
        ("deferredcoo_colidx", _p), ("deferredcoo_ptr", _p), ("rowblkblock", C.c_int),
        ("flag_bal_tile_rowidx", _p), ("tile_bal_rowidx_colstart", _p), ("tile_bal_rowidx_colstop", _p),
        ("flag_tilerow_start", _p), ("flag_tilerow_stop", _p), ("tile_bal_rowidx_colstart_v2", _p),
        ("tile_bal_rowidx_colstop_v2", _p), ("map", _p),
    ]


# symbols declared in include/mf.h (checked by the CPU tests)
EXPORTS = (
    "mmio_info", "mmio_data", "mmio_data_Low", "blockspmv_cpu", "encode", "decode", "exclusive_scan", "exclusive_scan_char",
    "binary_read_csr", "binary_write_csr", "read_Dmatrix", "read_Dmatrix_32", "read_Dmatrix_convert", "mf_write_mtx",
    "Tile_create", "Tile_destroy", "cg_solve_inc", "bicgstab", "mf_status_str", "mf_last_error",
    "mf_init", "mf_finalize", "mf_matrix_create", "mf_matrix_destroy", "mf_matrix_export", "mf_ref_tiles_free",
    "mf_matrix_encode_cols", "mf_matrix_classes", "mf_matrix_get_info", "mf_opts_default",
    "mf_solve", "mf_solve_device", "mf_solve_device_noalloc", "mf_vendor_solve", "mf_vendor_solve_device", "mf_matrix_layout", "mf_matrix_set_index_width", "mf_spmv", "mf_spmv_bench",
    "mf_solve_batched", "mf_mc_classify", "mf_memory_report", "mf_compact_rows", "mf_partition_build", "mf_free",
    "mf_gen_laplace2d", "mf_gen_poisson3d", "mf_gen_convdiff3d", "mf_gen_random_spd",
    "mf_gen_stencil_lognormal", "mf_gen_uniform", "mf_gen_normal",
    "mf_batch_create", "mf_batch_solve", "mf_batch_info", "mf_batch_destroy", "mf_batch_tune", "mf_abi_sizes",
)


def abi_sizes():
    """Sizes of the header's structs as compiled into the library, beside
    this binding's mirrors of them: {name: (library, binding)}."""
    names = ("mf_ref_tiles", "mf_matrix_info", "mf_opts", "mf_report", "mf_batch_info_t", "mf_memory",
             "mf_compact", "mf_partition", "Tile_matrix")
    mirrors = (RefTiles, Info, Opts, Report, BatchInfo, Mem, Compact, Partition, TileMatrix)
    out = (C.c_size_t * len(names))()
    k = lib().mf_abi_sizes(out, len(names))
    assert k == len(names), k
    return {n: (int(out[i]), C.sizeof(m)) for i, (n, m) in enumerate(zip(names, mirrors))}

_L = None


def lib():
    """Loads libmfamd.so (raises when it was not built)."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.mf_status_str.restype = C.c_char_p
    L.mf_last_error.restype = C.c_char_p
    L.mf_matrix_create.argtypes = [C.c_int, C.c_int, C.c_int, _i, _i, _d, _f, C.c_int, C.POINTER(C.c_void_p)]
    L.mf_matrix_destroy.argtypes = [C.c_void_p]
    L.mf_matrix_export.argtypes = [C.c_void_p, C.POINTER(RefTiles)]
    L.mf_ref_tiles_free.argtypes = [C.POINTER(RefTiles)]
    L.mf_matrix_encode_cols.argtypes = [C.c_void_p, _u8]
    L.mf_matrix_classes.argtypes = [C.c_void_p, C.c_double, C.c_int, _u8]
    L.mf_matrix_get_info.argtypes = [C.c_void_p, C.POINTER(Info)]
    L.mf_matrix_layout.argtypes = [C.c_void_p, C.c_double, C.c_int]
    L.mf_matrix_set_index_width.argtypes = [C.c_void_p, C.c_int]
    L.mf_spmv.argtypes = [C.c_void_p, C.c_int, _d, _d]
    L.mf_spmv_bench.argtypes = [C.c_void_p, C.c_int, C.c_int, _d]
    L.mf_opts_default.argtypes = [C.POINTER(Opts)]
    L.mf_solve.argtypes = [C.c_void_p, _d, _d, C.POINTER(Opts), C.POINTER(Report)]
    L.mf_solve_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Report)]
    L.mf_solve_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Report)]
    L.mf_mc_classify.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.mf_mc_classify.restype = C.c_double
    L.mf_compact_rows.argtypes = [C.POINTER(RefTiles), C.POINTER(Compact)]
    L.mf_partition_build.argtypes = [C.POINTER(RefTiles), C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(Partition)]
    L.mf_free.argtypes = [C.c_void_p]
    L.mmio_info.argtypes = [_i, _i, _i, _i, C.c_char_p]
    L.mmio_data.argtypes = [_i, _i, _d, C.c_char_p]
    L.encode.argtypes = [_u8, _u8, C.c_int, C.c_int]
    L.decode.argtypes = [_u8, _u8, C.c_int]
    L.exclusive_scan.argtypes = [_i, C.c_int]
    L.exclusive_scan_char.argtypes = [_u8, C.c_int]
    L.mf_gen_laplace2d.argtypes = [C.c_int, C.c_int, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]
    L.mf_gen_poisson3d.argtypes = [C.c_int, C.c_int, C.c_int, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]
    L.mf_gen_convdiff3d.argtypes = [C.c_int, C.c_double, C.c_double, C.c_ulonglong, _i, C.POINTER(_i),
                                    C.POINTER(_i), C.POINTER(_d)]
    L.mf_gen_random_spd.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_ulonglong, _i,
                                    C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]
    L.mf_gen_stencil_lognormal.argtypes = [C.c_int, C.c_int, C.c_double, C.c_ulonglong, _i, C.POINTER(_i),
                                           C.POINTER(_i), C.POINTER(_d)]
    L.mf_gen_uniform.argtypes = [C.c_ulonglong, C.c_int, C.c_double, C.c_double, _d]
    L.mf_gen_normal.argtypes = [C.c_ulonglong, C.c_int, _d]
    L.mf_solve_batched.argtypes = [C.c_int, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d), C.c_int,
                                   C.POINTER(_d), C.POINTER(_d), C.POINTER(Opts), _i, _d, C.POINTER(Report)]
    L.mf_batch_create.argtypes = [C.c_int, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d), C.c_int,
                                  C.POINTER(Opts), C.POINTER(C.c_void_p)]
    L.mf_batch_solve.argtypes = [C.c_void_p, C.POINTER(_d), C.POINTER(_d), _i, _d, C.POINTER(Report)]
    L.mf_batch_info.argtypes = [C.c_void_p, C.POINTER(BatchInfo)]
    L.mf_batch_destroy.argtypes = [C.c_void_p]
    L.mf_batch_tune.argtypes = [C.c_void_p, C.POINTER(_d), C.POINTER(Report)]
    L.mf_finalize.argtypes = []
    L.mf_finalize.restype = None
    _L = L
    # release device state at interpreter exit, before the runtime (and a
    # profiler attached to it) shuts down; objects collected later find
    # their handles released and do nothing
    atexit.register(L.mf_finalize)
    return L


def _check(rc):
    if rc != 0:
        L = lib()
        raise MFError(rc, f"{L.mf_status_str(rc).decode()}: {L.mf_last_error().decode()}")


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _copy(ptr, n, dtype):
    if n <= 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def ref_tiles_to_dict(t):
    n1, m1 = t.tilenum + 1, t.tilem + 1
    csz = t.csrsize // 2 + (t.csrsize % 2)
    fmt = np.ctypeslib.as_array(C.cast(t.Format, _u8), shape=(max(t.tilenum, 1),))[: t.tilenum]
    return {
        "tilem": t.tilem, "tilen": t.tilen, "tilenum": t.tilenum, "csrsize": t.csrsize,
        "csrptrlen": t.csrptrlen,
        "tile_ptr": _copy(t.tile_ptr, m1, np.int32),
        "tile_columnidx": _copy(t.tile_columnidx, t.tilenum, np.int32),
        "tile_nnz": _copy(t.tile_nnz, n1, np.int32),
        "Format": fmt.astype(np.uint8),
        "blknnz": _copy(t.blknnz, n1, np.int32),
        "blknnznnz": _copy(t.blknnznnz, n1, np.uint8),
        "csr_offset": _copy(t.csr_offset, n1, np.int32),
        "csrptr_offset": _copy(t.csrptr_offset, n1, np.int32),
        "Blockcsr_Val": _copy(t.Blockcsr_Val, t.csrsize, np.float64),
        "Blockcsr_Val_Low": _copy(t.Blockcsr_Val_Low, t.csrsize, np.float32),
        "Blockcsr_Ptr": _copy(t.Blockcsr_Ptr, t.csrptrlen, np.uint8),
        "Tile_csr_Col": _copy(t.Tile_csr_Col, t.csrsize, np.uint8),
        "csr_compressedIdx": _copy(t.csr_compressedIdx, csz, np.uint8),
    }


class BatchInfo(C.Structure):
    _fields_ = [("nsys", C.c_int), ("nrhs", C.c_int), ("B", C.c_int), ("G", C.c_int),
                ("rows", C.c_longlong), ("tilenum", C.c_longlong), ("bytes_alg", C.c_double),
                ("ms_setup", C.c_double), ("nnz_per_code", C.c_longlong * 5), ("mixed", C.c_int),
                ("tau", C.c_double), ("tuned", C.c_char * 512), ("index_width", C.c_int),
                ("on_chip", C.c_int), ("fell_back", C.c_int)]


class Batch:
    """Many independent SPD systems (CSR triples) with nrhs right-hand sides
    each, solved by one persistent CG launch captured in a hipGraph."""

    def __init__(self, systems, nrhs=1, **opts):
        L = lib()
        self.nsys = len(systems)
        self.nrhs = nrhs
        self._rp = [np.ascontiguousarray(s[0], np.int32) for s in systems]
        self._ci = [np.ascontiguousarray(s[1], np.int32) for s in systems]
        self._va = [np.ascontiguousarray(s[2], np.float64) for s in systems]
        self.n = np.array([len(r) - 1 for r in self._rp], np.int32)
        rp = (_i * self.nsys)(*[_p(a, C.c_int) for a in self._rp])
        ci = (_i * self.nsys)(*[_p(a, C.c_int) for a in self._ci])
        va = (_d * self.nsys)(*[_p(a, C.c_double) for a in self._va])
        o = default_opts(**opts)
        h = C.c_void_p()
        _check(L.mf_batch_create(self.nsys, _p(self.n, C.c_int), rp, ci, va, nrhs, C.byref(o), C.byref(h)))
        self.h = h

    def info(self):
        i = BatchInfo()
        _check(lib().mf_batch_info(self.h, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in BatchInfo._fields_ if k not in ("nnz_per_code", "tuned")}
        d["nnz_per_code"] = list(i.nnz_per_code)
        d["tuned"] = i.tuned.decode()
        return d

    def _ptrs(self, b):
        bs = [np.ascontiguousarray(np.reshape(bb, (self.nrhs, -1)), np.float64) for bb in b]
        return bs, (_d * self.nsys)(*[_p(a, C.c_double) for a in bs])

    def tune(self, b):
        """Autotunes on right-hand sides b (trial solves); returns the report
        of the last trial."""
        bs, bp = self._ptrs(b)
        rep = Report()
        _check(lib().mf_batch_tune(self.h, bp, C.byref(rep)))
        return rep.as_dict()

    def solve(self, b):
        """b: list of arrays of shape (nrhs, n_s) (or (n_s,) for nrhs = 1).
        Returns (x list of the same shapes, per-system iterations, per-system
        true relative residual, report)."""
        bs = [np.ascontiguousarray(np.reshape(bb, (self.nrhs, -1)), np.float64) for bb in b]
        xs = [np.zeros_like(bb) for bb in bs]
        bp = (_d * self.nsys)(*[_p(a, C.c_double) for a in bs])
        xp = (_d * self.nsys)(*[_p(a, C.c_double) for a in xs])
        it = np.zeros(self.nsys, np.int32)
        rel = np.zeros(self.nsys, np.float64)
        rep = Report()
        _check(lib().mf_batch_solve(self.h, bp, xp, _p(it, C.c_int), _p(rel, C.c_double), C.byref(rep)))
        if self.nrhs == 1:
            xs = [x[0] for x in xs]
        return xs, it, rel, rep.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().mf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_opts(**kw):
    o = Opts()
    lib().mf_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Matrix:
    """A device-resident matrix; its reference tile layout is built on the GPU."""

    def __init__(self, rowptr, col, val, rowA=None, colA=None, B=16, val_low=None):
        L = lib()
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.val = np.ascontiguousarray(val, np.float64)
        self.rowA = len(self.rowptr) - 1 if rowA is None else int(rowA)
        self.colA = self.rowA if colA is None else int(colA)
        low = None if val_low is None else np.ascontiguousarray(val_low, np.float32)
        h = C.c_void_p()
        _check(L.mf_matrix_create(self.rowA, self.colA, len(self.col), _p(self.rowptr, C.c_int),
                                  _p(self.col, C.c_int), _p(self.val, C.c_double),
                                  _p(low, C.c_float) if low is not None else None, B, C.byref(h)))
        self.h = h
        self.B = B

    def close(self):
        if getattr(self, "h", None):
            lib().mf_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self):
        t = RefTiles()
        _check(lib().mf_matrix_export(self.h, C.byref(t)))
        d = ref_tiles_to_dict(t)
        self._last_export = t
        return d

    def export_struct(self):
        t = RefTiles()
        _check(lib().mf_matrix_export(self.h, C.byref(t)))
        return t

    def encode_cols(self):
        info = self.info()
        out = np.zeros((info["nnz"] + 1) // 2 + 1, np.uint8)
        _check(lib().mf_matrix_encode_cols(self.h, _p(out, C.c_ubyte)))
        return out[: (info["nnz"] + 1) // 2]

    def classes(self, tau=0.0, allow_mask=ALLOW_ALL_LOW):
        info = self.info()
        out = np.zeros(max(info["tilenum"], 1), np.uint8)
        _check(lib().mf_matrix_classes(self.h, tau, allow_mask, _p(out, C.c_ubyte)))
        return out[: info["tilenum"]]

    def info(self):
        i = Info()
        _check(lib().mf_matrix_get_info(self.h, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in Info._fields_}
        d["tiles_per_code"] = list(i.tiles_per_code)
        d["nnz_per_code"] = list(i.nnz_per_code)
        return d

    def layout(self, tau=0.0, allow_mask=ALLOW_ALL_LOW):
        _check(lib().mf_matrix_layout(self.h, tau, allow_mask))

    def set_index_width(self, width):
        """Slab index width of the compute layouts: 0 automatic, or 1, 2, 4
        bytes (that form where it fits, else the next wider one)."""
        _check(lib().mf_matrix_set_index_width(self.h, width))

    def spmv(self, x, mixed=False):
        info = self.info()
        y = np.zeros(info["tilem"] * self.B, np.float64)
        xx = np.ascontiguousarray(x, np.float64)
        assert len(xx) >= self.colA
        _check(lib().mf_spmv(self.h, 1 if mixed else 0, _p(xx, C.c_double), _p(y, C.c_double)))
        return y[: self.rowA]

    def spmv_bench(self, reps=20, mixed=False):
        ms = C.c_double()
        _check(lib().mf_spmv_bench(self.h, 1 if mixed else 0, reps, C.byref(ms)))
        return ms.value

    def solve(self, b, **opts):
        o = default_opts(**opts)
        rep = Report()
        bb = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.rowA, np.float64)
        _check(lib().mf_solve(self.h, _p(bb, C.c_double), _p(x, C.c_double), C.byref(o), C.byref(rep)))
        return x, rep.as_dict()

    def vendor_solve(self, b, **opts):
        """The vendor-library baseline (rocSPARSE SpMV + rocBLAS, host loop)."""
        o = default_opts(**opts)
        rep = Report()
        bb = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.rowA, np.float64)
        _check(lib().mf_vendor_solve(self.h, _p(bb, C.c_double), _p(x, C.c_double), C.byref(o), C.byref(rep)))
        return x, rep.as_dict()

    def vendor_solve_device(self, d_b, d_x, **opts):
        o = default_opts(**opts)
        rep = Report()
        _check(lib().mf_vendor_solve_device(self.h, C.c_void_p(d_b), C.c_void_p(d_x), C.byref(o), C.byref(rep)))
        return rep.as_dict()

    def solve_device(self, d_b, d_x, **opts):
        """Solve with device pointers (tilem*B doubles each, b zero padded)."""
        o = default_opts(**opts)
        rep = Report()
        _check(lib().mf_solve_device(self.h, C.c_void_p(d_b), C.c_void_p(d_x), C.byref(o), C.byref(rep)))
        return rep.as_dict()

    def memory_report(self, tau=0.0, allow_mask=ALLOW_ALL_LOW):
        r = Mem()
        lib().mf_memory_report.argtypes = [C.c_void_p, C.c_double, C.c_int, C.POINTER(Mem)]
        _check(lib().mf_memory_report(self.h, tau, allow_mask, C.byref(r)))
        d = {k: getattr(r, k) for k, _ in Mem._fields_}
        d["nnz_per_code"] = list(r.nnz_per_code)
        return d

    def mc_classify(self):
        cnt = (C.c_longlong * 4)()
        mb = lib().mf_mc_classify(self.h, cnt)
        return mb, [cnt[i] for i in range(4)]


def _take_csr(n, rp, ci, va, nnz=None):
    nrows = n
    m = _copy(rp, nrows + 1, np.int32)
    nz = int(m[-1])
    out = (m, _copy(ci, nz, np.int32), _copy(va, nz, np.float64))
    for q in (rp, ci, va):
        lib().mf_free(C.cast(q, C.c_void_p))
    return out


def gen_laplace2d(nx, ny):
    n, rp, ci, va = C.c_int(), _i(), _i(), _d()
    _check(lib().mf_gen_laplace2d(nx, ny, C.byref(n), C.byref(rp), C.byref(ci), C.byref(va)))
    return _take_csr(n.value, rp, ci, va)


def gen_poisson3d(nx, ny, nz):
    n, rp, ci, va = C.c_int(), _i(), _i(), _d()
    _check(lib().mf_gen_poisson3d(nx, ny, nz, C.byref(n), C.byref(rp), C.byref(ci), C.byref(va)))
    return _take_csr(n.value, rp, ci, va)


def gen_convdiff3d(m, peclet=10.0, sigma=1.0, seed=0):
    n, rp, ci, va = C.c_int(), _i(), _i(), _d()
    _check(lib().mf_gen_convdiff3d(m, peclet, sigma, seed, C.byref(n), C.byref(rp), C.byref(ci), C.byref(va)))
    return _take_csr(n.value, rp, ci, va)


def gen_random_spd(n, zipf_a=2.0, lmin=2, lmax=2000, seed=0):
    nnz, rp, ci, va = C.c_int(), _i(), _i(), _d()
    _check(lib().mf_gen_random_spd(n, zipf_a, lmin, lmax, seed, C.byref(nnz), C.byref(rp), C.byref(ci),
                                   C.byref(va)))
    return _take_csr(n, rp, ci, va)


def gen_stencil_lognormal(dim, n_target, sigma=1.5, seed=0):
    n, rp, ci, va = C.c_int(), _i(), _i(), _d()
    _check(lib().mf_gen_stencil_lognormal(dim, n_target, sigma, seed, C.byref(n), C.byref(rp), C.byref(ci),
                                          C.byref(va)))
    return _take_csr(n.value, rp, ci, va)


def gen_uniform(seed, count, lo, hi):
    out = np.zeros(count, np.float64)
    _check(lib().mf_gen_uniform(seed, count, lo, hi, _p(out, C.c_double)))
    return out


def gen_normal(seed, count):
    out = np.zeros(count, np.float64)
    _check(lib().mf_gen_normal(seed, count, _p(out, C.c_double)))
    return out


def mmio(path):
    L = lib()
    m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = L.mmio_info(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym), path.encode())
    if rc:
        return rc, None
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    va = np.zeros(max(nnz.value, 1), np.float64)
    rc = L.mmio_data(_p(rp, C.c_int), _p(ci, C.c_int), _p(va, C.c_double), path.encode())
    return rc, dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value, rowptr=rp, col=ci[: nnz.value],
                    val=va[: nnz.value])


_i64p = C.POINTER(C.c_int64)


def _take64(r, z, rp, ci, va):
    out = dict(m=r, rowptr=np.ctypeslib.as_array(rp, (r + 1,)).astype(np.int64),
               col=np.ctypeslib.as_array(ci, (max(z, 1),))[:z].copy(),
               val=np.ctypeslib.as_array(va, (max(z, 1),))[:z].copy())
    for q in (rp, ci, va):
        lib().mf_free(C.cast(q, C.c_void_p))
    return out


def read_cbd(path):
    """binary_read_csr: (rc, dict(m, n, nnz, rowptr int64, col, val))."""
    L = lib()
    r, c, z = C.c_int(), C.c_int(), C.c_int64()
    rp, ci, va = _i64p(), _i(), _d()
    L.binary_read_csr.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(_i64p),
                                  C.POINTER(_i), C.POINTER(_d), C.c_char_p]
    rc = L.binary_read_csr(C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci), C.byref(va), path.encode())
    if rc:
        return rc, None
    d = _take64(r.value, z.value, rp, ci, va)
    d.update(n=c.value, nnz=z.value)
    return 0, d


def write_cbd(path, m, n, rowptr, col, val):
    L = lib()
    rp = np.ascontiguousarray(rowptr, np.int64)
    ci = np.ascontiguousarray(col, np.int32)
    va = np.ascontiguousarray(val, np.float64)
    L.binary_write_csr.argtypes = [C.c_int, C.c_int, C.c_int64, _i64p, _i, _d, C.c_char_p]
    L.binary_write_csr(m, n, int(rp[m]), rp.ctypes.data_as(_i64p), _p(ci, C.c_int), _p(va, C.c_double),
                       path.encode())


def read_matrix_32(path):
    """read_Dmatrix_32: dict(m, n, nnz, sym, rowptr, col, val) or None."""
    L = lib()
    r, c, z, sym = C.c_int(-1), C.c_int(), C.c_int(), C.c_int()
    rp, ci, va = _i(), _i(), _d()
    L.read_Dmatrix_32.argtypes = [C.POINTER(C.c_int)] * 3 + [C.POINTER(_i), C.POINTER(_i), C.POINTER(_d),
                                                              C.POINTER(C.c_int), C.c_char_p]
    L.read_Dmatrix_32(C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci), C.byref(va), C.byref(sym),
                      path.encode())
    if r.value < 0:
        return None
    d = _take64(r.value, z.value, rp, ci, va)
    d["rowptr"] = d["rowptr"].astype(np.int32)
    d.update(n=c.value, nnz=z.value, sym=sym.value)
    return d


def read_matrix_convert(path):
    """read_Dmatrix_convert: reads <path minus 3 chars>cbd when present, else
    the .mtx, writing that cache."""
    L = lib()
    r, c, z, sym = C.c_int(-1), C.c_int(), C.c_int64(), C.c_int()
    rp, ci, va = _i64p(), _i(), _d()
    L.read_Dmatrix_convert.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                       C.POINTER(_i64p), C.POINTER(_i), C.POINTER(_d), C.POINTER(C.c_int),
                                       C.c_char_p]
    L.read_Dmatrix_convert(C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci), C.byref(va),
                           C.byref(sym), path.encode())
    if r.value < 0:
        return None
    d = _take64(r.value, z.value, rp, ci, va)
    d.update(n=c.value, nnz=z.value)
    return d


def write_mtx(path, m, n, rowptr, col, val):
    L = lib()
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(col, np.int32)
    va = np.ascontiguousarray(val, np.float64)
    L.mf_write_mtx.argtypes = [C.c_char_p, C.c_int, C.c_int, _i, _i, _d]
    _check(L.mf_write_mtx(path.encode(), m, n, _p(rp, C.c_int), _p(ci, C.c_int), _p(va, C.c_double)))


def compact_rows(ref_struct):
    c = Compact()
    _check(lib().mf_compact_rows(C.byref(ref_struct), C.byref(c)))
    cnt = c.cnt_non_new
    out = dict(cnt_non_new=cnt,
               nonzero_row_new=_copy(c.nonzero_row_new, ref_struct.tilenum + 1, np.int32),
               blockrowid_new=_copy(c.blockrowid_new, cnt + 1, np.uint8),
               blockcsr_ptr_new=_copy(c.blockcsr_ptr_new, cnt + 1, np.uint8),
               block_signal=_copy(c.block_signal, ref_struct.tilem + 1, np.int32),
               tile_rowidx=_copy(c.tile_rowidx, ref_struct.tilenum, np.int32))
    for q in (c.nonzero_row_new, c.blockrowid_new, c.blockcsr_ptr_new, c.block_signal, c.tile_rowidx):
        lib().mf_free(C.cast(q, C.c_void_p))
    return out


def partition(ref_struct, nnzR, block_nnz, sm_num, mode):
    p = Partition()
    _check(lib().mf_partition_build(C.byref(ref_struct), nnzR, block_nnz, sm_num, mode, C.byref(p)))
    n1 = ref_struct.tilenum + 1
    out = dict(index=p.index, each_block_nnz=p.each_block_nnz, block_per_warp=p.block_per_warp,
               vector_each_warp_32=p.vector_each_warp_32, vector_total_32=p.vector_total_32,
               balance_tile_ptr=_copy(p.balance_tile_ptr, p.index + 1, np.int32),
               index_each_block=_copy(p.index_each_block, n1, np.int32),
               row_each_block=_copy(p.row_each_block, n1, np.int32),
               non_each_block_offset=_copy(p.non_each_block_offset, n1, np.int32),
               balance_tile_ptr_shared_end=_copy(p.balance_tile_ptr_shared_end, p.index + 1, np.int32))
    for q in (p.balance_tile_ptr, p.index_each_block, p.row_each_block, p.non_each_block_offset,
              p.balance_tile_ptr_shared_end):
        lib().mf_free(C.cast(q, C.c_void_p))
    return out


def encode(cols, offset=0):
    cols = np.ascontiguousarray(cols, np.uint8)
    res = np.zeros(len(cols) // 2 + 1 + offset, np.uint8)
    lib().encode(_p(cols, C.c_ubyte), _p(res, C.c_ubyte), len(cols), offset)
    return res[: offset + (len(cols) + 1) // 2]


def decode(packed, n):
    packed = np.ascontiguousarray(packed, np.uint8)
    res = np.zeros(n, np.uint8)
    lib().decode(_p(packed, C.c_ubyte), _p(res, C.c_ubyte), n)
    return res


def tile_create(rowA, colA, rowptr, col, val, val_low=None):
    """Tile_create (PC/csr2block.h:267) into a TileMatrix (16-row tiles); the
    caller releases it with tile_destroy."""
    L = lib()
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(col, np.int32)
    va = np.ascontiguousarray(val, np.float64)
    low = None if val_low is None else np.ascontiguousarray(val_low, np.float32)
    t = TileMatrix()
    L.Tile_create.argtypes = [C.POINTER(TileMatrix), C.c_int, C.c_int, C.c_int, _i, _i, _d, _f]
    L.Tile_create(C.byref(t), rowA, colA, int(rp[rowA]), _p(rp, C.c_int), _p(ci, C.c_int), _p(va, C.c_double),
                  _p(low, C.c_float) if low is not None else None)
    return t


def tile_destroy(t):
    L = lib()
    L.Tile_destroy.argtypes = [C.POINTER(TileMatrix)]
    L.Tile_destroy(C.byref(t))


def blockspmv_cpu(t, rowA, colA, x):
    """blockspmv_cpu (PC/blockspmv_cpu.h:2) over a TileMatrix: dict of the
    per-tile offsets, the row-block split and y (tilem * 16 entries)."""
    L = lib()
    n1 = max(t.tilenum, 1)
    p1 = np.zeros(n1, np.int32)
    p2 = np.zeros(n1, np.int32)
    rbb = C.c_int()
    br, cs, ce = C.POINTER(C.c_uint)(), _i(), _i()
    xx = np.zeros(max(t.tilen * 16, colA), np.float64)
    xx[: len(x)] = x
    y = np.zeros(t.tilem * 16, np.float64)
    yg = np.zeros(max(rowA, 1), np.float64)
    L.blockspmv_cpu.argtypes = [C.POINTER(TileMatrix), _i, _i, C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_uint)),
                                C.POINTER(_i), C.POINTER(_i), C.c_int, C.c_int, C.c_int, _i, _i, _d, _d, _d, _d]
    L.blockspmv_cpu(C.byref(t), _p(p1, C.c_int), _p(p2, C.c_int), C.byref(rbb), C.byref(br), C.byref(cs),
                    C.byref(ce), rowA, colA, 0, None, None, None, _p(xx, C.c_double), _p(y, C.c_double),
                    _p(yg, C.c_double))
    k = rbb.value
    out = dict(rowblkblock=k, ptroffset1=p1[: t.tilenum], ptroffset2=p2[: t.tilenum],
               blkrow=_copy(br, k, np.uint32), colstart=_copy(cs, k, np.int32), colstop=_copy(ce, k, np.int32),
               y=y, y_golden=yg)
    for q in (br, cs, ce):
        L.mf_free(C.cast(q, C.c_void_p))
    return out


def mmio_low(path):
    """mmio_data_Low (PC/mmio_highlevel.h:594): (rc, dict(rowptr, col, val float32))."""
    L = lib()
    m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = L.mmio_info(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym), path.encode())
    if rc:
        return rc, None
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    va = np.zeros(max(nnz.value, 1), np.float32)
    L.mmio_data_Low.argtypes = [_i, _i, _f, C.c_char_p]
    rc = L.mmio_data_Low(_p(rp, C.c_int), _p(ci, C.c_int), _p(va, C.c_float), path.encode())
    return rc, dict(rowptr=rp, col=ci[: nnz.value], val=va[: nnz.value])
