"""CPU tests of the product's host side: the C-ABI library loads and exports
every declared symbol, and the host functions (Matrix Market reader, nibble
codec, scans, reference-schedule preprocessing, generators) agree with the
oracle and the golden fixtures.  No GPU call is made here."""
import os
import re

import numpy as np
import pytest

from fixtures import (GOLDEN, golden_cases, golden_tiles, golden_variants, load_golden, to_ref_struct,
                      to_tile_matrix)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "mf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "for", "while")))


def test_header_declarations_are_exported(mf):
    L = mf.lib()
    decl = declared_functions()
    assert len(decl) >= 35
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(mf.EXPORTS) == sorted(decl)


@pytest.mark.parametrize("name", ["1138_bus", "ani1_nonsymm", "LFAT5"])
def test_mmio_matches_reference_reader(mf, name):
    rc, d = mf.mmio(os.path.join(GOLDEN, f"{name}.mtx"))
    g = load_golden(name)
    assert rc == 0 and d["nnz"] == int(g["nnz"]) and d["sym"] == int(g["sym"])
    for k in ("rowptr", "col", "val"):
        assert np.array_equal(d[k], g[k]), k


def test_mmio_variants_and_errors(mf, orc, tmp_path):
    p = tmp_path / "m.mtx"
    # pattern, integer, complex (real part), comments, blank tail, hermitian
    cases = [
        "%%MatrixMarket matrix coordinate pattern general\n% c\n3 3 4\n1 1\n2 3\n3 1\n3 3\n",
        "%%MatrixMarket matrix coordinate integer symmetric\n3 3 3\n1 1 7\n3 1 -2\n2 2 5\n\n",
        "%%MatrixMarket matrix coordinate complex hermitian\n2 2 2\n1 1 1.5 0\n2 1 2.25e-1 -1\n",
        "%%MatrixMarket MATRIX COORDINATE REAL GENERAL\n2 3 3\n1 3 1e300\n2 1 -0.1\n1 1 3.0000000000000004\n",
    ]
    for txt in cases:
        p.write_text(txt)
        rc1, a = mf.mmio(str(p))
        rc2, b = orc.mmio(str(p))
        assert rc1 == rc2 == 0
        for k in ("rowptr", "col", "val"):
            assert np.array_equal(a[k], b[k]), (txt, k)
    # the banner token itself is compared case-sensitively (PC/mmio.h:449)
    p.write_text("%%MATRIXMARKET matrix coordinate real general\n1 1 1\n1 1 1\n")
    assert mf.mmio(str(p))[0] == orc.mmio(str(p))[0] == -2
    assert mf.mmio(str(tmp_path / "missing.mtx"))[0] == -1
    p.write_text("%%MatrixMarket matrix coordinate real weird\n1 1 1\n1 1 1\n")
    assert mf.mmio(str(p))[0] == -2
    p.write_text("%%MatrixMarket matrix coordinate real general\n")
    assert mf.mmio(str(p))[0] == -4


@pytest.mark.parametrize("name", golden_cases())
def test_encode_decode_match_reference(mf, name):
    g = load_golden(name)
    for key, B, rowA in golden_variants(g):
        cols = g[f"{key}_Tile_csr_Col"]
        packed = mf.encode(cols)
        assert np.array_equal(packed, g[f"{key}_encode"]), key
        if B == 16:
            assert np.array_equal(mf.decode(packed, len(cols)), cols)


def test_scans_match_oracle(mf, orc):
    import ctypes as C
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 17, 1000):
        a = rng.integers(0, 300, n).astype(np.int32)
        b, c = a.copy(), a.copy()
        mf.lib().exclusive_scan(b.ctypes.data_as(C.POINTER(C.c_int)), n)
        orc.L.orc_exclusive_scan(c.ctypes.data_as(C.POINTER(C.c_int)), n)
        assert np.array_equal(b, c)
        u = rng.integers(0, 256, n).astype(np.uint8)
        v, w = u.copy(), u.copy()
        mf.lib().exclusive_scan_char(v.ctypes.data_as(C.POINTER(C.c_ubyte)), n)
        orc.L.orc_exclusive_scan_char(w.ctypes.data_as(C.POINTER(C.c_ubyte)), n)
        assert np.array_equal(v, w)


@pytest.mark.parametrize("name", golden_cases())
def test_partition_and_compact_match_oracle(mf, orc, name):
    g = load_golden(name)
    for key, B, rowA in golden_variants(g):
        if B != 16:
            continue
        d = golden_tiles(g, key)
        ref = to_ref_struct(mf, B, rowA, int(g["n"]), d)
        ot = orc.tiles(B, rowA, int(g["n"]), g["rowptr"], g["col"], g["val"])
        oc = ot.compact()
        pc = mf.compact_rows(ref)
        for k in ("nonzero_row_new", "blockrowid_new", "blockcsr_ptr_new", "block_signal", "tile_rowidx"):
            assert np.array_equal(pc[k], oc[k]), (key, k)
        for mode, bnnz, sm in ((0, 16, 6912), (1, 16, 6912), (1, 16, 8192), (0, 64, 6912)):
            pp = mf.partition(ref, len(g["col"]), bnnz, sm, mode)
            op = ot.partition(len(g["col"]), bnnz, sm, mode)
            for k in op:
                assert np.array_equal(np.asarray(pp[k]), np.asarray(op[k])), (key, mode, k)


def test_partition_known_answer_1138_bus(mf):
    g = load_golden("1138_bus")
    d = golden_tiles(g, "b16_trunc")
    ref = to_ref_struct(mf, 16, int(g["b16_trunc_rowA"]), int(g["n"]), d)
    assert mf.partition(ref, int(g["nnz"]), 16, 6912, 0)["index"] == 149


def test_generators_shapes_and_properties(mf):
    import scipy.sparse as sp
    rp, ci, va = mf.gen_laplace2d(256, 256)
    assert len(rp) - 1 == 65536 and rp[-1] == 326656  # 2D 5-point, 256^2
    rp, ci, va = mf.gen_poisson3d(16, 16, 16)
    assert rp[-1] == 16 ** 3 * 7 - 6 * 16 ** 2
    A = sp.csr_matrix((va, ci, rp))
    assert abs(A - A.T).max() == 0 and np.all(A.diagonal() == 6.0)
    m = 12
    rp, ci, va = mf.gen_convdiff3d(m, 10.0, 1.0, 0)
    assert rp[-1] == (3 * m - 2) ** 3  # 27-point: nnz = (3m-2)^3
    A = sp.csr_matrix((va, ci, rp))
    off = abs(A).sum(axis=1).A1 - abs(A.diagonal())
    assert np.all(A.diagonal() >= off * (1 - 1e-14))  # weakly diagonally dominant M-matrix
    assert abs(A - A.T).max() > 0  # convection makes it nonsymmetric
    rp, ci, va = mf.gen_random_spd(2000, seed=0)
    A = sp.csr_matrix((va, ci, rp))
    assert abs(A - A.T).max() == 0
    off = abs(A).sum(axis=1).A1 - abs(A.diagonal())
    assert np.allclose(A.diagonal(), 1.05 * off)
    for dim in (2, 3):
        rp, ci, va = mf.gen_stencil_lognormal(dim, 1000, 1.5, 3)
        A = sp.csr_matrix((va, ci, rp))
        assert A.shape[0] >= 1000 and abs(A - A.T).max() < 1e-12
    # the portable stream is deterministic and uniform in [lo, hi)
    u1, u2 = mf.gen_uniform(0, 1000, -1, 1), mf.gen_uniform(0, 1000, -1, 1)
    assert np.array_equal(u1, u2) and u1.min() >= -1 and u1.max() < 1 and abs(u1.mean()) < 0.1
    z = mf.gen_normal(0, 20000)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03


def test_cli_binaries_built():
    for exe in ("mf_cg", "mf_bicgstab"):
        assert os.access(os.path.join(ROOT, "mille-feillue_amd", exe), os.X_OK)


def _ref_biio():
    from orc import RefBiio
    try:
        return RefBiio()
    except FileNotFoundError as e:
        pytest.skip(f"reference biio build missing: {e}")


def test_mtx_writer_round_trip(mf, tmp_path):
    # what mf_write_mtx writes, mmio_data reads back bit for bit
    for rp, ci, va in (mf.gen_poisson3d(7, 6, 5), mf.gen_random_spd(800, seed=4),
                       mf.gen_stencil_lognormal(2, 500, 1.5, 3)):
        n = len(rp) - 1
        p = str(tmp_path / "w.mtx")
        mf.write_mtx(p, n, n, rp, ci, va)
        rc, d = mf.mmio(p)
        assert rc == 0 and d["m"] == n and d["n"] == n and d["nnz"] == len(ci)
        assert np.array_equal(d["rowptr"], rp) and np.array_equal(d["col"], ci)
        assert np.array_equal(d["val"].view(np.uint64), np.asarray(va).view(np.uint64))


@pytest.mark.parametrize("name", ["1138_bus", "ani1_nonsymm", "LFAT5"])
def test_binary_csr_matches_reference(mf, tmp_path, name):
    """.cbd files written by the library and by the reference's biio are the
    same bytes and each reader reads the other's file; read_Dmatrix_32 equals
    the reference's on the shipped matrices."""
    ref = _ref_biio()
    src = os.path.join(GOLDEN, f"{name}.mtx")
    mine, theirs = mf.read_matrix_32(src), ref.read_matrix_32(src)
    assert mine is not None and theirs is not None
    for k in ("m", "n", "nnz", "sym"):
        assert mine[k] == theirs[k], k
    for k in ("rowptr", "col", "val"):
        assert np.array_equal(mine[k], theirs[k]), k
    a, b = str(tmp_path / "a.cbd"), str(tmp_path / "b.cbd")
    m, n = mine["m"], mine["n"]
    rp64 = mine["rowptr"].astype(np.int64)
    mf.write_cbd(a, m, n, rp64, mine["col"], mine["val"])
    ref.write_cbd(b, m, n, rp64, mine["col"], mine["val"])
    assert open(a, "rb").read() == open(b, "rb").read()
    for reader in (mf.read_cbd, ref.read_cbd):
        for path in (a, b):
            rc, d = reader(path)
            assert rc == 0 and d["m"] == m and d["n"] == n and d["nnz"] == mine["nnz"]
            for k in ("rowptr", "col", "val"):
                assert np.array_equal(d[k], rp64 if k == "rowptr" else mine[k]), (reader, path, k)
    # read_Dmatrix_32 takes the .cbd by its extension
    d = mf.read_matrix_32(a)
    assert np.array_equal(d["rowptr"], mine["rowptr"]) and np.array_equal(d["val"], mine["val"])


def test_binary_csr_cache_and_errors(mf, tmp_path):
    import shutil
    src = tmp_path / "g.mtx"
    shutil.copy(os.path.join(GOLDEN, "1138_bus.mtx"), src)
    cache = tmp_path / "g.cbd"
    assert not cache.exists()
    d1 = mf.read_matrix_convert(str(src))
    assert d1 is not None and cache.exists()
    rc, dm = mf.mmio(str(src))
    assert np.array_equal(d1["rowptr"], dm["rowptr"]) and np.array_equal(d1["val"], dm["val"])
    # with the cache present the Matrix Market file is not needed
    src.write_text("not a matrix\n")
    d2 = mf.read_matrix_convert(str(src))
    assert all(np.array_equal(d1[k], d2[k]) for k in ("rowptr", "col", "val"))
    # a missing or truncated binary file is reported as -1
    assert mf.read_cbd(str(tmp_path / "none.cbd"))[0] == -1
    cache.write_bytes(cache.read_bytes()[:100])
    assert mf.read_cbd(str(cache))[0] == -1
    # other extensions leave the outputs untouched
    assert mf.read_matrix_32(str(tmp_path / "m.txt")) is None


def test_binding_structs_match_library(mf):
    # every ctypes mirror has the size of the struct compiled into the library
    sizes = mf.abi_sizes()
    bad = {k: v for k, v in sizes.items() if v[0] != v[1]}
    assert not bad, bad


@pytest.mark.parametrize("name", golden_cases())
def test_blockspmv_cpu_matches_reference(mf, orc, name):
    """blockspmv_cpu over the reference-built Tile_matrix arrays (golden
    fixtures): the per-tile offsets and the row-block split equal the
    reference's own outputs bit for bit; y equals the oracle's tile product
    over the same arrays, and the CSR product where no tile holds more than
    255 entries (beyond, the reference's uint8 row starts wrap and the tile
    arrays no longer hold every entry).  The reference's own y decodes an
    unfilled column array and is not comparable."""
    import scipy.sparse as sp
    g = load_golden(name)
    n = int(g["n"])
    rng = np.random.default_rng(7)
    for key, B, rowA in golden_variants(g):
        if B != 16:
            continue
        t = to_tile_matrix(mf, golden_tiles(g, key))
        x = rng.standard_normal(n)
        out = mf.blockspmv_cpu(t, rowA, n, x)
        assert out["rowblkblock"] == int(g[f"{key}_off_rowblkblock"]), key
        for k in ("ptroffset1", "ptroffset2", "blkrow", "colstart", "colstop"):
            assert np.array_equal(np.asarray(out[k]), g[f"{key}_off_{k}"]), (key, k)
        ot = orc.tiles(B, rowA, n, g["rowptr"], g["col"], g["val"])
        xx = np.zeros(ot.d["tilen"] * B)
        xx[:n] = x
        want = ot.spmv(xx)[:rowA]
        np.testing.assert_allclose(out["y"][:rowA], want, rtol=1e-12, atol=1e-12 * np.abs(want).max(), err_msg=key)
        if np.diff(g[f"{key}_blknnz"]).max() <= 255:
            A = sp.csr_matrix((g["val"], g["col"], g["rowptr"]), shape=(int(g["m"]), n))[:rowA]
            np.testing.assert_allclose(out["y"][:rowA], A @ x, rtol=1e-12, atol=1e-12 * np.abs(want).max(),
                                       err_msg=key)
        assert not out["y"][rowA:].any()
        assert not out["y_golden"].any()  # left untouched, as in the reference


@pytest.mark.parametrize("name", ["1138_bus", "ani1_nonsymm", "LFAT5"])
def test_mmio_data_low_matches_reference_reader(mf, name):
    # PC/mmio_highlevel.h:594: the same CSR, values converted to float
    rc, d = mf.mmio_low(os.path.join(GOLDEN, f"{name}.mtx"))
    g = load_golden(name)
    assert rc == 0
    assert np.array_equal(d["rowptr"], g["rowptr"]) and np.array_equal(d["col"], g["col"])
    assert np.array_equal(d["val"], g["val"].astype(np.float32))
    assert mf.mmio_low(os.path.join(GOLDEN, "missing.mtx"))[0] == -1
