"""Pins the oracle (CPU restatement) before it is trusted as the checker:
against the reference's own host C (golden fixtures written from oracle/_ref),
the reference's published known answers, and what the mathematics fixes."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from fixtures import GOLDEN, golden_cases, golden_tiles, golden_variants, load_golden
from orc import REF_ARRAYS, REF_DIR, Ref


@pytest.mark.parametrize("name", ["1138_bus", "ani1_nonsymm", "LFAT5"])
def test_mmio_matches_reference_reader(orc, name):
    rc, d = orc.mmio(os.path.join(GOLDEN, f"{name}.mtx"))
    g = load_golden(name)
    assert rc == 0
    assert d["nnz"] == int(g["nnz"]) and d["sym"] == int(g["sym"])
    for k in ("rowptr", "col", "val"):
        assert np.array_equal(d[k], g[k]), k


def test_mmio_known_answer_1138_bus(orc):
    # nnzR = 4054 after symmetric expansion (PC/data/cg_solver_mi200.csv, 1138_bus row)
    rc, d = orc.mmio(os.path.join(GOLDEN, "1138_bus.mtx"))
    assert rc == 0 and d["nnz"] == 4054 and d["m"] == 1138


def test_mmio_errors(orc, tmp_path):
    assert orc.mmio(str(tmp_path / "missing.mtx"))[0] == -1
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%NotMatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n")
    assert orc.mmio(str(bad))[0] == -2
    bad.write_text("%%MatrixMarket matrix coordinate real general\n% only a comment\n")
    assert orc.mmio(str(bad))[0] == -4


@pytest.mark.parametrize("name", golden_cases())
def test_tile_create_matches_reference(orc, name):
    g = load_golden(name)
    for key, B, rowA in golden_variants(g):
        t = orc.tiles(B, rowA, int(g["n"]), g["rowptr"], g["col"], g["val"])
        ref = golden_tiles(g, key)
        for k in REF_ARRAYS:
            assert np.array_equal(t.d[k], ref[k]), (key, k)
        off = t.offsets()
        for k in ("rowblkblock", "ptroffset1", "ptroffset2", "blkrow", "colstart", "colstop"):
            assert np.array_equal(np.asarray(off[k]), g[f"{key}_off_{k}"]), (key, k)


@pytest.mark.parametrize("name", golden_cases())
def test_encode_matches_reference(orc, name):
    g = load_golden(name)
    for key, B, rowA in golden_variants(g):
        cols = g[f"{key}_Tile_csr_Col"]
        packed = orc.encode(cols)
        assert np.array_equal(packed, g[f"{key}_encode"]), key
        if B == 16:
            assert np.array_equal(orc.decode(packed, len(cols)), cols)


def test_tiles_reconstruct_matrix(orc):
    """Brute force: every tile expanded to a dense BxB block gives back the CSR."""
    rng = np.random.default_rng(3)
    A = sp.random(77, 90, density=0.08, random_state=4, format="csr")
    A.data = rng.standard_normal(A.nnz)
    for B in (8, 16):
        t = orc.tiles(B, 77, 90, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
        d = t.d
        dense = np.zeros((d["tilem"] * B, d["tilen"] * B))
        for bi in range(d["tilem"]):
            rowlen = B if bi < d["tilem"] - 1 else 77 - (d["tilem"] - 1) * B
            for tid in range(d["tile_ptr"][bi], d["tile_ptr"][bi + 1]):
                off, po = d["csr_offset"][tid], d["csrptr_offset"][tid]
                nz = d["blknnz"][tid + 1] - d["blknnz"][tid]
                assert nz == d["tile_nnz"][tid + 1] - d["tile_nnz"][tid]
                for ri in range(rowlen):
                    s = d["Blockcsr_Ptr"][po + ri]
                    e = nz if ri == rowlen - 1 else d["Blockcsr_Ptr"][po + ri + 1]
                    for k in range(s, e):
                        dense[bi * B + ri, d["tile_columnidx"][tid] * B + d["Tile_csr_Col"][off + k]] += \
                            d["Blockcsr_Val"][off + k]
        assert np.array_equal(dense[:77, :90], A.toarray())
        assert np.all(np.diff(d["tile_ptr"]) >= 0) and d["tile_nnz"][-1] == A.nnz


def test_partition_known_answer_1138_bus(orc):
    # MP/data/cg_syncfree_reduce_2757.csv: 1138_bus, each_nnz 16 -> index 149
    g = load_golden("1138_bus")
    rowA = int(g["b16_trunc_rowA"])
    t = orc.tiles(16, rowA, int(g["n"]), g["rowptr"], g["col"], g["val"])
    p = t.partition(int(g["nnz"]), 16, 6912, 0)
    assert p["index"] == 149 and p["each_block_nnz"] == 16


def test_partition_invariants(orc):
    g = load_golden("nontrivial_mc64_example")
    t = orc.tiles(16, int(g["m"]), int(g["n"]), g["rowptr"], g["col"], g["val"])
    n = t.d["tilenum"]
    for mode, bnnz in ((0, 16), (0, 64), (1, 16)):
        p = t.partition(len(g["col"]), bnnz, 6912, mode)
        bt = p["balance_tile_ptr"]
        assert bt[0] == 0 and np.all(np.diff(bt) >= 0) and bt[-1] == n
        # every tile appears exactly once in the recorded order
        assert sorted(p["index_each_block"][:n].tolist()) == list(range(n))


def test_compact_rows_reproduce_rows(orc):
    g = load_golden("dup_wrap")
    t = orc.tiles(16, int(g["m"]), int(g["n"]), g["rowptr"], g["col"], g["val"])
    c = t.compact()
    d = t.d
    for tid in range(d["tilenum"]):
        seg = slice(c["nonzero_row_new"][tid], c["nonzero_row_new"][tid + 1])
        rows, stops = c["blockrowid_new"][seg], c["blockcsr_ptr_new"][seg]
        assert rows[0] == 0 and stops[0] == 0  # sentinel
        assert np.all(np.diff(rows[1:].astype(int)) > 0)
    assert c["block_signal"][: d["tilem"]].tolist() == np.diff(d["tile_ptr"]).tolist()


@pytest.mark.parametrize("name", golden_cases())
def test_compact_rows_stops_against_reference_tiles(orc, name):
    """Every (row, stop) pair of orc_compact_rows, brute-forced from the tile
    arrays the reference's own csr2block.h built (golden fixtures): per tile a
    (0, 0) sentinel, then each row whose stop differs from its uint8 start,
    where the stop is the next row's uint8 start, or the tile's nonzero count
    for the tile's last row (PC/Mille-feuille_CG_AMD.cu:1459-1480), stored
    as uint8.  The golden row starts are first checked against the CSR: each
    row's entries in the tile's column range, mod 256."""
    g = load_golden(name)
    n = int(g["n"])
    wrapped = 0
    for key, B, rowA in golden_variants(g):
        ref = golden_tiles(g, key)
        tp, tcol, bn, po, ptr = (ref[k] for k in ("tile_ptr", "tile_columnidx", "blknnz", "csrptr_offset",
                                                   "Blockcsr_Ptr"))
        c = orc.tiles(B, rowA, n, g["rowptr"], g["col"], g["val"]).compact()
        tilem, tilenum = len(tp) - 1, len(tcol)
        assert len(c["nonzero_row_new"]) == tilenum + 1
        cols = g["col"]
        for bi in range(tilem):
            rowlen = B if bi < tilem - 1 else rowA - (tilem - 1) * B
            for tid in range(tp[bi], tp[bi + 1]):
                cc, nz = int(po[tid]), int(bn[tid + 1] - bn[tid])
                lo, hi = int(tcol[tid]) * B, int(tcol[tid]) * B + B
                cnt = [int(np.count_nonzero((cols[g["rowptr"][bi * B + ri]:g["rowptr"][bi * B + ri + 1]] >= lo)
                                            & (cols[g["rowptr"][bi * B + ri]:g["rowptr"][bi * B + ri + 1]] < hi)))
                       for ri in range(rowlen)]
                assert sum(cnt) == nz, (key, tid)
                if rowlen > 1:  # scanned starts: S_ri mod 256
                    assert [int(ptr[cc + ri]) for ri in range(rowlen)] == \
                        [sum(cnt[:ri]) % 256 for ri in range(rowlen)], (key, tid)
                expect = [(0, 0)]
                for ri in range(rowlen):
                    start = int(ptr[cc + ri])
                    stop = nz if ri == rowlen - 1 else int(ptr[cc + ri + 1])
                    if stop != start:
                        expect.append((ri, stop % 256))
                    wrapped += stop != sum(cnt[:ri + 1])
                seg = slice(int(c["nonzero_row_new"][tid]), int(c["nonzero_row_new"][tid + 1]))
                got = list(zip(c["blockrowid_new"][seg].tolist(), c["blockcsr_ptr_new"][seg].tolist()))
                assert got == expect, (key, tid)
        assert c["cnt_non_new"] == int(c["nonzero_row_new"][-1])
    if name == "dup_wrap":
        assert wrapped > 0  # tiles of more than 255 entries: the uint8 wrap is exercised


def test_mc_classifier_closed_form(orc):
    # MC/memory_Mille_feillue.cu:866-918 bins |v| <= 3.9375, 60, 6000
    vals = np.array([3.9375, -3.9375, 3.94, 60.0, -60.0000001, 6000.0, 6000.5, 1e-300, -1e9])
    n = len(vals)
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    t = orc.tiles(16, n, n, rp, ci, vals)
    c = t.compact()
    assert c["mc_counts"] == [3, 2, 2, 2]
    assert c["mc_memory_MB"] == pytest.approx((8 * 2 + 2 * 2 + 4 * 2 + 3) / 1048576.0)


def _all_e4m3(orc):
    codes = [c for c in range(256) if (c & 0x7F) != 0x7F]
    return {c: orc.L.orc_e4m3_to_double(c) for c in codes}


def _brute_round(v, table, mantissa_bits_of):
    """Nearest representable value, ties to the even mantissa, saturating."""
    best = None
    for code, val in table.items():
        d = abs(val - v)
        if best is None or d < best[0] or (d == best[0] and mantissa_bits_of(code) % 2 == 0 and
                                           mantissa_bits_of(best[1]) % 2 == 1):
            best = (d, code)
    return table[best[1]]


def test_e4m3_against_brute_force(orc):
    table = _all_e4m3(orc)
    assert max(table.values()) == 448.0 and min(v for v in table.values() if v > 0) == 2.0 ** -9
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.uniform(-500, 500, 300), rng.uniform(-1, 1, 300) * 2.0 ** rng.integers(-12, 9, 300),
                           [0.0, -0.0, 448.0, 449.0, 464.0, 1e9, 2.0 ** -10, 3 * 2.0 ** -11, 0.0625 + 2.0 ** -8]])
    for v in vals:
        got = orc.L.orc_e4m3_to_double(orc.L.orc_e4m3_from_double(float(v)))
        want = _brute_round(min(max(v, -448.0), 448.0), table, lambda c: c & 7)
        assert got == want, (v, got, want)


def test_f16_against_numpy(orc):
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.standard_normal(2000) * 2.0 ** rng.integers(-26, 17, 2000),
                           [65504.0, 65519.9, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 0.0]])
    for v in vals:
        # numpy's float64 -> float16 conversion rounds once, to nearest even
        want = float(np.float16(v))
        got = orc.L.orc_f16_to_double(orc.L.orc_f16_from_double(float(v)))
        assert got == want or (np.isinf(want) and np.isinf(got)), (v, got, want)


def test_tile_classes_policy(orc):
    # tau = 0: lossless only.  6 and -1 are exact in E4M3, 0.1 nowhere below FP64,
    # 1000 exceeds E4M3's range but is exact in FP16, 0.5 + 2^-12 needs FP32.
    rows = [[(0, 6.0), (1, -1.0)], [(16, 0.1)], [(32, 1000.0)], [(48, 0.5 + 2.0 ** -12)], [(64, 2.0 ** -8)]]
    n = 80
    rp, ci, va = [0], [], []
    for r in range(n):
        row = rows[r] if r < len(rows) else []
        for c, v in row:
            ci.append(c)
            va.append(v)
        rp.append(len(ci))
    t = orc.tiles(16, n, n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(va))
    assert t.classes(0.0).tolist() == [2, 0, 3, 4, 3]
    # 2^-8 is below E4M3's normal range (2^-6): FP16.  With tau = 2^-4 0.1 fits E4M3.
    assert t.classes(2.0 ** -4).tolist()[1] == 2
    assert t.classes(0.0, allow_mask=1 << 4).tolist() == [4, 0, 4, 4, 4]


def test_mixed_spmv_with_fp64_codes_is_tile_spmv(orc):
    g = load_golden("nontrivial_mc64_example")
    t = orc.tiles(16, int(g["m"]), int(g["n"]), g["rowptr"], g["col"], g["val"])
    x = np.random.default_rng(2).standard_normal(t.d["tilen"] * 16)
    codes = np.zeros(t.d["tilenum"], np.uint8)
    assert np.array_equal(t.spmv_mixed(codes, x), t.spmv(x))
    y = t.spmv(x)
    A = sp.csr_matrix((g["val"], g["col"], g["rowptr"]), shape=(int(g["m"]), int(g["n"])))
    assert np.allclose(y, A @ x[: int(g["n"])], rtol=1e-13, atol=1e-12)


def test_mixed_spmv_low_codes_against_scipy(orc):
    """Lossless low-precision tiles: the mixed product differs from the exact
    one only by FP32 rounding of x and FP32 accumulation, so it must agree
    with scipy's FP64 product to ~1e-6 relative -- a dropped term, a wrong
    sign or a wrong column would be off by O(1)."""
    rng = np.random.default_rng(5)
    n = 200
    A = sp.random(n, n, density=0.05, random_state=6, format="csr")
    A.data = rng.choice([-2.0, -1.0, -0.5, 0.25, 1.0, 3.0, 6.0], A.nnz)  # exact in E4M3
    A = (A + sp.diags(np.full(n, 8.0))).tocsr()
    A.data[: n // 3] = 1000.0 + np.arange(n // 3)  # forces FP16/FP32 in some tiles
    t = orc.tiles(16, n, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    codes = t.classes(0.0)
    assert set(codes.tolist()) >= {2, 3}
    x = rng.standard_normal(t.d["tilen"] * 16)
    x[n:] = 0.0
    y = t.spmv_mixed(codes, x)[:n]
    ref = A @ x[:n]
    scale = np.abs(A) @ np.abs(x[:n])
    assert np.all(np.abs(y - ref) <= 1e-6 * scale + 1e-300)
    assert not np.array_equal(y, ref)  # the FP32 path really ran


def test_cg_oracle_against_direct_solve(orc):
    g = load_golden("isai_spd")
    A = sp.csr_matrix((g["val"], g["col"], g["rowptr"]), shape=(int(g["m"]), int(g["n"])))
    b = A @ np.ones(A.shape[0])
    x, rep = orc.cg(g["rowptr"], g["col"], g["val"], b, 1e-12, 1000)
    xd = spl.spsolve(A.tocsc(), b)
    assert rep.converged and rep.relres_true <= 1e-12
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_cg_oracle_finite_termination(orc):
    # a diagonal SPD matrix with 3 distinct eigenvalues: CG is exact after 3 steps
    n = 48
    d = np.array([1.0, 4.0, 9.0] * (n // 3))
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    b = np.linspace(1, 2, n)
    x, rep = orc.cg(rp, ci, d, b, 1e-13, 100)
    assert rep.iterations == 3 and np.allclose(x, b / d, rtol=1e-12)


def test_bicgstab_oracle_against_direct_solve(orc):
    g = load_golden("ani1_nonsymm")
    A = sp.csr_matrix((g["val"], g["col"], g["rowptr"]), shape=(int(g["m"]), int(g["n"])))
    b = A @ np.linspace(-1, 1, A.shape[0])
    x, rep = orc.bicgstab(g["rowptr"], g["col"], g["val"], b, 1e-11, 2000)
    xd = spl.spsolve(A.tocsc(), b)
    assert rep.converged and rep.relres_true <= 1e-11
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7


def _csr(A):
    A = sp.csr_matrix(A)
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def _spd_cases():
    g = load_golden("isai_spd")
    yield "isai_spd", (g["rowptr"], g["col"], g["val"])
    # 2D 5-point Laplacian 24 x 24
    k = 24
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(k, k))
    yield "laplace24", _csr(sp.kron(sp.eye(k), T) + sp.kron(T, sp.eye(k)))


def _nonsym_cases():
    g = load_golden("ani1_nonsymm")
    yield "ani1_nonsymm", (g["rowptr"], g["col"], g["val"])
    # well conditioned nonsymmetric: upwinded 1D convection-diffusion blocks
    rng = np.random.default_rng(5)
    n = 300
    A = sp.diags([-1.3, 3.0, -0.4], [-1, 0, 1], shape=(n, n)) + sp.random(n, n, density=0.01, random_state=6) * 0.2
    A = A.tocsr()
    A.data += 0.0 * rng.standard_normal(A.nnz)
    yield "convdiff1d", _csr(A)


@pytest.mark.parametrize("case", list(range(2)))
def test_petsc_cg_iterates_match_textbook_cg(orc, case):
    """KSPSolve_CG (cg.c:112, unpreconditioned) and Hestenes-Stiefel CG have the
    same iterates in exact arithmetic: the residual histories agree to
    rounding, iteration by iteration (a dropped or wrong-signed term in
    either restatement changes them from the second iteration on)."""
    name, (rp, ci, va) = list(_spd_cases())[case]
    n = len(rp) - 1
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    b = A @ np.linspace(-1.0, 2.0, n)
    xp, rp_ = orc.petsc_cg(rp, ci, va, b, 1e-10, 2000, hist=2001)
    xc, rc_ = orc.cg(rp, ci, va, b, 1e-10, 2000, hist=2001)
    hp, hc = rp_.history, rc_.history
    k = min(len(hp), len(hc), 11)
    assert k >= 10, (name, len(hp), len(hc))
    np.testing.assert_allclose(hp[:k], hc[:k], rtol=1e-8, err_msg=name)
    assert rp_.converged and rc_.converged
    assert abs(rp_.iterations - rc_.iterations) <= 1, name
    xd = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(xp - xd) / np.linalg.norm(xd) < 1e-8, name


@pytest.mark.parametrize("case", list(range(2)))
def test_petsc_bcgs_iterates_match_bicgstab(orc, case):
    """KSPSolve_BCGS (bcgs.c:19-150) and van der Vorst BiCGSTAB are the same
    recurrences (p = r + beta (p - omega v) written as r - omega beta v +
    beta p): their first ten residual norms agree to 1e-8 relative on well
    conditioned nonsymmetric systems, and both reach the direct solution."""
    name, (rp, ci, va) = list(_nonsym_cases())[case]
    n = len(rp) - 1
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    b = A @ np.linspace(-1.0, 1.0, n)
    xp, rp_ = orc.petsc_bcgs(rp, ci, va, b, 1e-10, 2000, hist=2001)
    xb, rb_ = orc.bicgstab(rp, ci, va, b, 1e-10, 2000, hist=2001)
    hp, hb = rp_.history, rb_.history
    k = min(len(hp), len(hb), 11)
    assert k >= 10, (name, len(hp), len(hb))
    np.testing.assert_allclose(hp[:k], hb[:k], rtol=1e-8, err_msg=name)
    assert rp_.converged and rb_.converged
    xd = spl.spsolve(A.tocsc(), b)
    for x in (xp, xb):
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7, name


def test_bicgstab_restatements_finite_termination(orc):
    """A diagonal matrix with two distinct eigenvalues: BiCG's polynomial
    annihilates r0 at step 2, so both BiCGSTAB restatements are exact after
    two iterations (PETSc's through its d2 = (t, t) = 0 branch, bcgs.c:104)."""
    n = 40
    d = np.array([1.0, 3.0] * (n // 2))
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    b = np.linspace(1.0, 2.0, n)
    for solve in (orc.petsc_bcgs, orc.bicgstab):
        x, rep = solve(rp, ci, d, b, 1e-12, 50, hist=51)
        assert rep.converged and rep.iterations == 2, (solve, rep.iterations)
        assert rep.history[-1] <= 1e-13 * rep.history[0]
        np.testing.assert_allclose(x, b / d, rtol=1e-13)
    # CG on three distinct eigenvalues: KSPCG exact after three iterations
    d3 = np.array([1.0, 4.0, 9.0] * 16)
    rp3 = np.arange(49, dtype=np.int32)
    x, rep = orc.petsc_cg(rp3, rp3[:-1].copy(), d3, np.linspace(1, 2, 48), 1e-13, 100)
    assert rep.converged and rep.iterations == 3
    np.testing.assert_allclose(x, np.linspace(1, 2, 48) / d3, rtol=1e-12)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "libref_b16.so")), reason="oracle/_ref not built")
def test_oracle_matches_reference_on_random_matrices(orc):
    R = Ref(16)
    rng = np.random.default_rng(11)
    for trial in range(6):
        m, n = int(rng.integers(17, 300)), int(rng.integers(17, 300))
        A = sp.random(m, n, density=float(rng.uniform(0.01, 0.2)), random_state=trial, format="csr")
        # shuffle columns inside rows (file order is kept by the layout)
        ci = A.indices.copy()
        for i in range(m):
            s, e = A.indptr[i], A.indptr[i + 1]
            ci[s:e] = rng.permutation(ci[s:e])
        rp = A.indptr.astype(np.int32)
        va = rng.standard_normal(A.nnz)
        t = orc.tiles(16, m, n, rp, ci.astype(np.int32), va)
        r = R.tiles(m, n, rp, ci.astype(np.int32), va)
        for k in REF_ARRAYS:
            assert np.array_equal(t.d[k], r[k]), (trial, k)
