"""ctypes access to the oracle (oracle/liboracle.so) and to the reference's own
host C built in oracle/_ref/ -- the checkers used by the tests.

Test infrastructure only; the product never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_i = C.POINTER(C.c_int)
_u8 = C.POINTER(C.c_ubyte)
_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)


class Tiles(C.Structure):
    _fields_ = [
        ("B", C.c_int), ("rowA", C.c_int), ("colA", C.c_int), ("tilem", C.c_int),
        ("tilen", C.c_int), ("tilenum", C.c_int), ("csrsize", C.c_int), ("csrptrlen", C.c_int),
        ("tile_ptr", _i), ("tile_columnidx", _i), ("tile_nnz", _i), ("Format", C.POINTER(C.c_char)),
        ("blknnz", _i), ("blknnznnz", _u8), ("csr_offset", _i), ("csrptr_offset", _i),
        ("Blockcsr_Val", _d), ("Blockcsr_Val_Low", _f), ("Blockcsr_Ptr", _u8),
        ("Tile_csr_Col", _u8), ("csr_compressedIdx", _u8),
    ]


class Compact(C.Structure):
    _fields_ = [("cnt_non_new", C.c_int), ("nonzero_row_new", _i), ("blockrowid_new", _u8),
                ("blockcsr_ptr_new", _u8), ("block_signal", _i), ("tile_rowidx", _i)]


class Partition(C.Structure):
    _fields_ = [("index", C.c_int), ("each_block_nnz", C.c_int), ("block_per_warp", C.c_int),
                ("vector_each_warp_32", C.c_int), ("vector_total_32", C.c_int),
                ("balance_tile_ptr", _i), ("index_each_block", _i), ("row_each_block", _i),
                ("non_each_block_offset", _i), ("balance_tile_ptr_shared_end", _i)]


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("relres_true", C.c_double),
                ("relres_updated", C.c_double), ("converged", C.c_int),
                ("hist", _d), ("hist_cap", C.c_int), ("hist_len", C.c_int)]


def _report(hist):
    """A zeroed report; with hist > 0 it records that many residual norms
    (available afterwards as rep.history)."""
    rep = Report()
    rep.history = None
    if hist:
        buf = np.zeros(hist, np.float64)
        rep._buf = buf
        rep.hist = buf.ctypes.data_as(_d)
        rep.hist_cap = hist
    return rep


def _history(rep):
    if getattr(rep, "_buf", None) is not None:
        rep.history = rep._buf[: rep.hist_len].copy()
    return rep


def _arr(ptr, n, dtype):
    if n <= 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def tiles_to_dict(t):
    """Copy every reference-layout array out of a Tiles struct."""
    n1, m1 = t.tilenum + 1, t.tilem + 1
    csz = t.csrsize // 2 + (t.csrsize % 2)
    fmt = np.ctypeslib.as_array(C.cast(t.Format, _u8), shape=(max(t.tilenum, 1),))[: t.tilenum]
    return {
        "tilem": t.tilem, "tilen": t.tilen, "tilenum": t.tilenum,
        "csrsize": t.csrsize, "csrptrlen": t.csrptrlen,
        "tile_ptr": _arr(t.tile_ptr, m1, np.int32),
        "tile_columnidx": _arr(t.tile_columnidx, t.tilenum, np.int32),
        "tile_nnz": _arr(t.tile_nnz, n1, np.int32),
        "Format": fmt.astype(np.uint8),
        "blknnz": _arr(t.blknnz, n1, np.int32),
        "blknnznnz": _arr(t.blknnznnz, n1, np.uint8),
        "csr_offset": _arr(t.csr_offset, n1, np.int32),
        "csrptr_offset": _arr(t.csrptr_offset, n1, np.int32),
        "Blockcsr_Val": _arr(t.Blockcsr_Val, t.csrsize, np.float64),
        "Blockcsr_Val_Low": _arr(t.Blockcsr_Val_Low, t.csrsize, np.float32),
        "Blockcsr_Ptr": _arr(t.Blockcsr_Ptr, t.csrptrlen, np.uint8),
        "Tile_csr_Col": _arr(t.Tile_csr_Col, t.csrsize, np.uint8),
        "csr_compressedIdx": _arr(t.csr_compressedIdx, csz, np.uint8),
    }


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle")
        L = self.L = C.CDLL(path)
        L.orc_tile_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i, _i, _d, _f, C.POINTER(Tiles)]
        L.orc_tiles_free.argtypes = [C.POINTER(Tiles)]
        L.orc_mmio_info.argtypes = [_i, _i, _i, _i, C.c_char_p]
        L.orc_mmio_data.argtypes = [_i, _i, _d, C.c_char_p]
        L.orc_encode.argtypes = [_u8, _u8, C.c_int, C.c_int]
        L.orc_decode.argtypes = [_u8, _u8, C.c_int]
        L.orc_compact_rows.argtypes = [C.POINTER(Tiles), C.POINTER(Compact)]
        L.orc_compact_free.argtypes = [C.POINTER(Compact)]
        L.orc_partition.argtypes = [C.POINTER(Tiles), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Partition)]
        L.orc_partition_free.argtypes = [C.POINTER(Partition)]
        L.orc_mc_classify.argtypes = [C.POINTER(Tiles), C.POINTER(Compact), C.POINTER(C.c_longlong)]
        L.orc_mc_classify.restype = C.c_double
        L.orc_tile_spmv.argtypes = [C.POINTER(Tiles), _d, _d]
        L.orc_tile_spmv_mixed.argtypes = [C.POINTER(Tiles), _u8, _d, _d]
        L.orc_csr_spmv.argtypes = [C.c_int, _i, _i, _d, _d, _d]
        L.orc_e4m3_from_double.argtypes = [C.c_double]
        L.orc_e4m3_from_double.restype = C.c_ubyte
        L.orc_e4m3_to_double.argtypes = [C.c_ubyte]
        L.orc_e4m3_to_double.restype = C.c_double
        L.orc_f16_from_double.argtypes = [C.c_double]
        L.orc_f16_from_double.restype = C.c_ushort
        L.orc_f16_to_double.argtypes = [C.c_ushort]
        L.orc_f16_to_double.restype = C.c_double
        L.orc_tile_classes.argtypes = [C.POINTER(Tiles), C.c_double, C.c_int, _u8]
        L.orc_blockspmv_offsets.argtypes = [C.POINTER(Tiles), _i, _i, C.POINTER(C.POINTER(C.c_uint)),
                                            C.POINTER(_i), C.POINTER(_i)]
        for name in ("orc_cg", "orc_bicgstab"):
            getattr(L, name).argtypes = [C.c_int, _i, _i, _d, C.POINTER(Tiles), _d, _d, C.c_double,
                                         C.c_int, C.c_int, C.POINTER(Report)]
        for name in ("orc_petsc_cg", "orc_petsc_bcgs"):
            getattr(L, name).argtypes = [C.c_int, _i, _i, _d, _d, _d, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(Report)]
        L.orc_free.argtypes = [C.c_void_p]

    # -- A1
    def mmio(self, path):
        m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rc = self.L.orc_mmio_info(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym), path.encode())
        if rc:
            return rc, None
        rp = np.zeros(m.value + 1, np.int32)
        ci = np.zeros(max(nnz.value, 1), np.int32)
        va = np.zeros(max(nnz.value, 1), np.float64)
        rc = self.L.orc_mmio_data(_p(rp, C.c_int), _p(ci, C.c_int), _p(va, C.c_double), path.encode())
        return rc, dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value, rowptr=rp,
                        col=ci[: nnz.value], val=va[: nnz.value])

    # -- A2/A3 (keeps the struct alive in the returned holder)
    def tiles(self, B, rowA, colA, rowptr, col, val, val_low=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        if val_low is None:
            val_low = val.astype(np.float32)
        val_low = np.ascontiguousarray(val_low, np.float32)
        t = Tiles()
        self.L.orc_tile_create(B, rowA, colA, int(rowptr[rowA]), _p(rowptr, C.c_int), _p(col, C.c_int),
                               _p(val, C.c_double), _p(val_low, C.c_float), C.byref(t))
        return OTiles(self, t)

    def encode(self, cols, offset=0):
        cols = np.ascontiguousarray(cols, np.uint8)
        res = np.zeros(len(cols) // 2 + 1 + offset, np.uint8)
        self.L.orc_encode(_p(cols, C.c_ubyte), _p(res, C.c_ubyte), len(cols), offset)
        return res[: offset + (len(cols) + 1) // 2]

    def decode(self, packed, n):
        packed = np.ascontiguousarray(packed, np.uint8)
        res = np.zeros(n, np.uint8)
        self.L.orc_decode(_p(packed, C.c_ubyte), _p(res, C.c_ubyte), n)
        return res

    def csr_spmv(self, rowptr, col, val, x):
        n = len(rowptr) - 1
        y = np.zeros(n, np.float64)
        self.L.orc_csr_spmv(n, _p(rowptr, C.c_int), _p(col, C.c_int), _p(val, C.c_double),
                            _p(np.ascontiguousarray(x, np.float64), C.c_double), _p(y, C.c_double))
        return y

    def _solve(self, name, rowptr, col, val, b, tol, maxit, tiles=None, nthreads=1, hist=0):
        n = len(b)
        x = np.zeros(n, np.float64)
        rep = _report(hist)
        tp = C.byref(tiles.t) if tiles is not None else None
        getattr(self.L, name)(n, _p(rowptr, C.c_int), _p(col, C.c_int), _p(val, C.c_double), tp,
                              _p(np.ascontiguousarray(b, np.float64), C.c_double), _p(x, C.c_double),
                              tol, maxit, nthreads, C.byref(rep))
        return x, _history(rep)

    def cg(self, *a, **k):
        return self._solve("orc_cg", *a, **k)

    def bicgstab(self, *a, **k):
        return self._solve("orc_bicgstab", *a, **k)

    def _psolve(self, name, rowptr, col, val, b, rtol, maxit, nthreads=1, hist=0):
        n = len(b)
        x = np.zeros(n, np.float64)
        rep = _report(hist)
        getattr(self.L, name)(n, _p(rowptr, C.c_int), _p(col, C.c_int), _p(val, C.c_double),
                              _p(np.ascontiguousarray(b, np.float64), C.c_double), _p(x, C.c_double),
                              rtol, maxit, nthreads, C.byref(rep))
        return x, _history(rep)

    def petsc_cg(self, *a, **k):
        return self._psolve("orc_petsc_cg", *a, **k)

    def petsc_bcgs(self, *a, **k):
        return self._psolve("orc_petsc_bcgs", *a, **k)


class OTiles:
    """Owns an oracle Tiles struct."""

    def __init__(self, orc, t):
        self.orc, self.t = orc, t
        self.d = tiles_to_dict(t)

    def __del__(self):
        try:
            self.orc.L.orc_tiles_free(C.byref(self.t))
        except Exception:
            pass

    def compact(self):
        c = Compact()
        self.orc.L.orc_compact_rows(C.byref(self.t), C.byref(c))
        cnt = c.cnt_non_new
        out = dict(cnt_non_new=cnt,
                   nonzero_row_new=_arr(c.nonzero_row_new, self.t.tilenum + 1, np.int32),
                   blockrowid_new=_arr(c.blockrowid_new, cnt + 1, np.uint8),
                   blockcsr_ptr_new=_arr(c.blockcsr_ptr_new, cnt + 1, np.uint8),
                   block_signal=_arr(c.block_signal, self.t.tilem + 1, np.int32),
                   tile_rowidx=_arr(c.tile_rowidx, self.t.tilenum, np.int32))
        counts = (C.c_longlong * 4)()
        out["mc_memory_MB"] = self.orc.L.orc_mc_classify(C.byref(self.t), C.byref(c), counts)
        out["mc_counts"] = [counts[i] for i in range(4)]
        self.orc.L.orc_compact_free(C.byref(c))
        return out

    def partition(self, nnzR, block_nnz, sm_num, mode):
        p = Partition()
        self.orc.L.orc_partition(C.byref(self.t), nnzR, block_nnz, sm_num, mode, C.byref(p))
        n1 = self.t.tilenum + 1
        out = dict(index=p.index, each_block_nnz=p.each_block_nnz, block_per_warp=p.block_per_warp,
                   vector_each_warp_32=p.vector_each_warp_32, vector_total_32=p.vector_total_32,
                   balance_tile_ptr=_arr(p.balance_tile_ptr, p.index + 1, np.int32),
                   index_each_block=_arr(p.index_each_block, n1, np.int32),
                   row_each_block=_arr(p.row_each_block, n1, np.int32),
                   non_each_block_offset=_arr(p.non_each_block_offset, n1, np.int32),
                   balance_tile_ptr_shared_end=_arr(p.balance_tile_ptr_shared_end, p.index + 1, np.int32))
        self.orc.L.orc_partition_free(C.byref(p))
        return out

    def offsets(self):
        n = self.t.tilenum
        p1 = np.zeros(max(n, 1), np.int32)
        p2 = np.zeros(max(n, 1), np.int32)
        br = C.POINTER(C.c_uint)()
        cs, ce = _i(), _i()
        rbb = self.orc.L.orc_blockspmv_offsets(C.byref(self.t), _p(p1, C.c_int), _p(p2, C.c_int),
                                               C.byref(br), C.byref(cs), C.byref(ce))
        out = dict(rowblkblock=rbb, ptroffset1=p1[:n], ptroffset2=p2[:n],
                   blkrow=_arr(br, rbb, np.uint32), colstart=_arr(cs, rbb, np.int32),
                   colstop=_arr(ce, rbb, np.int32))
        for q in (br, cs, ce):
            self.orc.L.orc_free(C.cast(q, C.c_void_p))
        return out

    def spmv(self, x):
        y = np.zeros(self.t.tilem * self.t.B, np.float64)
        self.orc.L.orc_tile_spmv(C.byref(self.t), _p(np.ascontiguousarray(x, np.float64), C.c_double),
                                 _p(y, C.c_double))
        return y[: self.t.rowA]

    def classes(self, tau=0.0, allow_mask=(1 << 2) | (1 << 3) | (1 << 4)):
        codes = np.zeros(max(self.t.tilenum, 1), np.uint8)
        self.orc.L.orc_tile_classes(C.byref(self.t), tau, allow_mask, _p(codes, C.c_ubyte))
        return codes[: self.t.tilenum]

    def spmv_mixed(self, codes, x):
        y = np.zeros(self.t.tilem * self.t.B, np.float64)
        codes = np.ascontiguousarray(codes, np.uint8)
        self.orc.L.orc_tile_spmv_mixed(C.byref(self.t), _p(codes, C.c_ubyte),
                                       _p(np.ascontiguousarray(x, np.float64), C.c_double),
                                       _p(y, C.c_double))
        return y[: self.t.rowA]


class Ref:
    """The reference's own host C (oracle/_ref), one library per BLOCK_SIZE."""

    def __init__(self, B=16):
        path = os.path.join(REF_DIR, f"libref_b{B}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.L = C.CDLL(path)
        self.B = B
        L.ref_tile_create.argtypes = [C.c_int, C.c_int, C.c_int, _i, _i, _d, _f, C.POINTER(Tiles)]
        L.ref_tile_create.restype = C.c_void_p
        L.ref_tile_free.argtypes = [C.c_void_p]
        L.ref_mmio_info.argtypes = [_i, _i, _i, _i, C.c_char_p]
        L.ref_mmio_data.argtypes = [_i, _i, _d, C.c_char_p]
        L.ref_encode.argtypes = [_u8, _u8, C.c_int, C.c_int]
        L.ref_decode.argtypes = [_u8, _u8, C.c_int]
        L.ref_blockspmv_offsets.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _i, _i, _d, _i, _i,
                                            C.POINTER(C.c_uint), _i, _i, C.c_int]

    def mmio(self, path):
        m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rc = self.L.ref_mmio_info(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym), path.encode())
        if rc:
            return rc, None
        rp = np.zeros(m.value + 1, np.int32)
        ci = np.zeros(max(nnz.value, 1), np.int32)
        va = np.zeros(max(nnz.value, 1), np.float64)
        rc = self.L.ref_mmio_data(_p(rp, C.c_int), _p(ci, C.c_int), _p(va, C.c_double), path.encode())
        return rc, dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value, rowptr=rp,
                        col=ci[: nnz.value], val=va[: nnz.value])

    def tiles(self, rowA, colA, rowptr, col, val, val_low=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        if val_low is None:
            val_low = val.astype(np.float32)
        val_low = np.ascontiguousarray(val_low, np.float32)
        t = Tiles()
        h = self.L.ref_tile_create(rowA, colA, int(rowptr[rowA]), _p(rowptr, C.c_int), _p(col, C.c_int),
                                   _p(val, C.c_double), _p(val_low, C.c_float), C.byref(t))
        d = tiles_to_dict(t)
        n = t.tilenum
        p1 = np.zeros(max(n, 1), np.int32)
        p2 = np.zeros(max(n, 1), np.int32)
        cap = 4 * (n + t.tilem) + 16
        br = np.zeros(cap, np.uint32)
        cs = np.zeros(cap, np.int32)
        ce = np.zeros(cap, np.int32)
        rbb = self.L.ref_blockspmv_offsets(h, rowA, colA, int(rowptr[rowA]), _p(rowptr, C.c_int),
                                           _p(col, C.c_int), _p(val, C.c_double), _p(p1, C.c_int),
                                           _p(p2, C.c_int), _p(br, C.c_uint), _p(cs, C.c_int),
                                           _p(ce, C.c_int), cap)
        d["offsets"] = dict(rowblkblock=rbb, ptroffset1=p1[:n], ptroffset2=p2[:n], blkrow=br[:rbb],
                            colstart=cs[:rbb], colstop=ce[:rbb])
        self.L.ref_tile_free(h)
        return d

    def encode(self, cols, offset=0):
        cols = np.ascontiguousarray(cols, np.uint8)
        res = np.zeros(len(cols) // 2 + 1 + offset, np.uint8)
        self.L.ref_encode(_p(cols, C.c_ubyte), _p(res, C.c_ubyte), len(cols), offset)
        return res[: offset + (len(cols) + 1) // 2]

    def decode(self, packed, n):
        packed = np.ascontiguousarray(packed, np.uint8)
        res = np.zeros(n, np.uint8)
        self.L.ref_decode(_p(packed, C.c_ubyte), _p(res, C.c_ubyte), n)
        return res


REF_ARRAYS = ("tile_ptr", "tile_columnidx", "tile_nnz", "Format", "blknnz", "blknnznnz",
              "csr_offset", "csrptr_offset", "Blockcsr_Val", "Blockcsr_Val_Low", "Blockcsr_Ptr",
              "Tile_csr_Col", "csr_compressedIdx")


class RefBiio:
    """The reference's binary CSR reader/writer (oracle/_ref/libref_biio.so)."""

    def __init__(self):
        path = os.path.join(REF_DIR, "libref_biio.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = L = C.CDLL(path)
        i64p = C.POINTER(C.c_int64)
        L.ref_binary_read_csr.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                          C.POINTER(i64p), C.POINTER(_i), C.POINTER(_d), C.c_char_p]
        L.ref_binary_write_csr.argtypes = [C.c_int, C.c_int, C.c_int64, i64p, _i, _d, C.c_char_p]
        L.ref_read_Dmatrix_32.argtypes = [C.POINTER(C.c_int)] * 3 + [C.POINTER(_i), C.POINTER(_i), C.POINTER(_d),
                                                                     C.POINTER(C.c_int), C.c_char_p]
        L.ref_biio_free.argtypes = [C.c_void_p]

    def _take(self, r, z, rp, ci, va, rp_dtype):
        out = dict(m=r, rowptr=np.ctypeslib.as_array(rp, (r + 1,)).astype(rp_dtype),
                   col=np.ctypeslib.as_array(ci, (max(z, 1),))[:z].copy(),
                   val=np.ctypeslib.as_array(va, (max(z, 1),))[:z].copy())
        for q in (rp, ci, va):
            self.L.ref_biio_free(C.cast(q, C.c_void_p))
        return out

    def read_cbd(self, path):
        r, c, z = C.c_int(), C.c_int(), C.c_int64()
        rp, ci, va = C.POINTER(C.c_int64)(), _i(), _d()
        rc = self.L.ref_binary_read_csr(C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci),
                                        C.byref(va), path.encode())
        if rc:
            return rc, None
        d = self._take(r.value, z.value, rp, ci, va, np.int64)
        d.update(n=c.value, nnz=z.value)
        return 0, d

    def write_cbd(self, path, m, n, rowptr, col, val):
        rp = np.ascontiguousarray(rowptr, np.int64)
        ci = np.ascontiguousarray(col, np.int32)
        va = np.ascontiguousarray(val, np.float64)
        self.L.ref_binary_write_csr(m, n, int(rp[m]), rp.ctypes.data_as(C.POINTER(C.c_int64)), _p(ci, C.c_int),
                                    _p(va, C.c_double), path.encode())

    def read_matrix_32(self, path):
        r, c, z, sym = C.c_int(-1), C.c_int(), C.c_int(), C.c_int()
        rp, ci, va = _i(), _i(), _d()
        self.L.ref_read_Dmatrix_32(C.byref(r), C.byref(c), C.byref(z), C.byref(rp), C.byref(ci), C.byref(va),
                                   C.byref(sym), path.encode())
        if r.value < 0:
            return None
        d = self._take(r.value, z.value, rp, ci, va, np.int32)
        d.update(n=c.value, nnz=z.value, sym=sym.value)
        return d
