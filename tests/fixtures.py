"""Shared test helpers: golden fixtures, small synthetic matrices, and a
builder of the library's mf_ref_tiles struct from numpy arrays."""
import ctypes as C
import glob
import os

import numpy as np

from orc import REF_ARRAYS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_cases():
    return sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(GOLDEN, "ref_*.npz")))


def load_golden(name):
    with np.load(os.path.join(GOLDEN, f"ref_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_variants(g):
    """(key, B, rowA) for every stored tile build of a fixture."""
    out = []
    for B in (16, 8):
        for tr in ("full", "trunc"):
            key = f"b{B}_{tr}"
            if f"{key}_rowA" in g:
                out.append((key, B, int(g[f"{key}_rowA"])))
    return out


def golden_tiles(g, key):
    return {k: g[f"{key}_{k}"] for k in REF_ARRAYS}


def to_ref_struct(mf, B, rowA, colA, d):
    """mf_ref_tiles pointing into the arrays of dict d (kept alive on the struct)."""
    keep = {k: np.ascontiguousarray(v) for k, v in d.items() if isinstance(v, np.ndarray)}
    t = mf.RefTiles()
    t.B, t.rowA, t.colA = B, rowA, colA
    t.tilem = len(keep["tile_ptr"]) - 1
    t.tilen = (colA + B - 1) // B
    t.tilenum = len(keep["tile_columnidx"])
    t.csrsize = len(keep["Blockcsr_Val"])
    t.csrptrlen = len(keep["Blockcsr_Ptr"])
    ptr = lambda k, ct: keep[k].ctypes.data_as(C.POINTER(ct))  # noqa: E731
    t.tile_ptr = ptr("tile_ptr", C.c_int)
    t.tile_columnidx = ptr("tile_columnidx", C.c_int)
    t.tile_nnz = ptr("tile_nnz", C.c_int)
    t.Format = keep["Format"].ctypes.data_as(C.POINTER(C.c_char))
    t.blknnz = ptr("blknnz", C.c_int)
    t.blknnznnz = ptr("blknnznnz", C.c_ubyte)
    t.csr_offset = ptr("csr_offset", C.c_int)
    t.csrptr_offset = ptr("csrptr_offset", C.c_int)
    t.Blockcsr_Val = ptr("Blockcsr_Val", C.c_double)
    t.Blockcsr_Val_Low = ptr("Blockcsr_Val_Low", C.c_float)
    t.Blockcsr_Ptr = ptr("Blockcsr_Ptr", C.c_ubyte)
    t.Tile_csr_Col = ptr("Tile_csr_Col", C.c_ubyte)
    t.csr_compressedIdx = ptr("csr_compressedIdx", C.c_ubyte)
    t._keep = keep
    return t


def csr_scipy(rp, ci, va, m, n):
    import scipy.sparse as sp
    return sp.csr_matrix((va, ci, rp), shape=(m, n))


def to_tile_matrix(mf, d):
    """The reference's Tile_matrix record pointing into the arrays of dict d
    (CSR-tile fields only; kept alive on the struct)."""
    keep = {k: np.ascontiguousarray(v) for k, v in d.items() if isinstance(v, np.ndarray)}
    t = mf.TileMatrix()
    t.tilem = len(keep["tile_ptr"]) - 1
    t.tilenum = len(keep["tile_columnidx"])
    t.tilen = int(keep["tile_columnidx"].max()) + 1 if t.tilenum else 0
    ptr = lambda k, ct: keep[k].ctypes.data_as(C.POINTER(ct))  # noqa: E731
    t.tile_ptr = ptr("tile_ptr", C.c_int)
    t.tile_columnidx = ptr("tile_columnidx", C.c_int)
    t.tile_nnz = ptr("tile_nnz", C.c_int)
    t.Format = keep["Format"].ctypes.data_as(C.POINTER(C.c_char))
    t.blknnz = ptr("blknnz", C.c_int)
    t.blknnznnz = ptr("blknnznnz", C.c_ubyte)
    t.csr_offset = ptr("csr_offset", C.c_int)
    t.csrptr_offset = ptr("csrptr_offset", C.c_int)
    t.Blockcsr_Val = ptr("Blockcsr_Val", C.c_double)
    t.Blockcsr_Val_Low = ptr("Blockcsr_Val_Low", C.c_float)
    t.Blockcsr_Ptr = ptr("Blockcsr_Ptr", C.c_ubyte)
    t.Tile_csr_Col = ptr("Tile_csr_Col", C.c_ubyte)
    t.csr_compressedIdx = ptr("csr_compressedIdx", C.c_ubyte)
    t.csrsize = len(keep["Blockcsr_Val"])
    t.csrptrlen = len(keep["Blockcsr_Ptr"])
    t._keep = keep
    return t


def tile_matrix_arrays(mf, t):
    """Copies of the CSR-tile arrays of a Tile_matrix filled by Tile_create."""
    n1, m1 = t.tilenum + 1, t.tilem + 1
    cp = lambda ptr, n, dt: np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dt, copy=True)  # noqa: E731
    fmt = np.ctypeslib.as_array(C.cast(t.Format, C.POINTER(C.c_ubyte)), shape=(max(t.tilenum, 1),))[: t.tilenum]
    return {
        "tile_ptr": cp(t.tile_ptr, m1, np.int32), "tile_columnidx": cp(t.tile_columnidx, t.tilenum, np.int32),
        "tile_nnz": cp(t.tile_nnz, n1, np.int32), "Format": fmt.astype(np.uint8),
        "blknnz": cp(t.blknnz, n1, np.int32), "blknnznnz": cp(t.blknnznnz, n1, np.uint8),
        "csr_offset": cp(t.csr_offset, n1, np.int32), "csrptr_offset": cp(t.csrptr_offset, n1, np.int32),
        "Blockcsr_Val": cp(t.Blockcsr_Val, t.csrsize, np.float64),
        "Blockcsr_Val_Low": cp(t.Blockcsr_Val_Low, t.csrsize, np.float32),
        "Blockcsr_Ptr": cp(t.Blockcsr_Ptr, t.csrptrlen, np.uint8),
        "Tile_csr_Col": cp(t.Tile_csr_Col, t.csrsize, np.uint8),
        "csr_compressedIdx": cp(t.csr_compressedIdx, t.csrsize // 2 + t.csrsize % 2, np.uint8),
    }
