"""GPU tile builder and compute-layout parity (needs an MI355X).

The reference tile layout built on the GPU must equal the reference's own
Tile_create output (golden fixtures) and the oracle bit-for-bit; the nibble
codec, the per-tile precision codes and the Memory-Cost classifier must match
the oracle exactly."""
import numpy as np
import pytest

from fixtures import golden_cases, golden_tiles, golden_variants, load_golden, tile_matrix_arrays
from orc import REF_ARRAYS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_cases())
def test_gpu_tile_build_matches_reference(gpu, name):
    g = load_golden(name)
    for key, B, rowA in golden_variants(g):
        print(f"case {name} {key}", flush=True)
        A = gpu.Matrix(g["rowptr"], g["col"], g["val"], rowA=rowA, colA=int(g["n"]), B=B)
        d = A.export()
        ref = golden_tiles(g, key)
        for k in REF_ARRAYS:
            assert np.array_equal(d[k], ref[k]), (key, k)
        assert np.array_equal(A.encode_cols(), g[f"{key}_encode"]), key
        A.close()


@pytest.mark.parametrize("name", golden_cases())
def test_gpu_tile_create_drop_in(gpu, name):
    """Tile_create (PC/csr2block.h:267) through the C-ABI fills the caller's
    Tile_matrix with arrays equal to the reference's own (golden fixtures),
    blockspmv_cpu reads that struct, and Tile_destroy releases and zeroes it."""
    g = load_golden(name)
    for key, B, rowA in golden_variants(g):
        if B != 16:
            continue
        print(f"case {name} {key}", flush=True)
        t = gpu.tile_create(rowA, int(g["n"]), g["rowptr"], g["col"], g["val"])
        assert t.tilem == (rowA + 15) // 16 and t.tilen == (int(g["n"]) + 15) // 16
        d = tile_matrix_arrays(gpu, t)
        ref = golden_tiles(g, key)
        for k in REF_ARRAYS:
            assert np.array_equal(d[k], ref[k]), (key, k)
        assert t.csrsize == len(ref["Blockcsr_Val"]) and t.csrptrlen == len(ref["Blockcsr_Ptr"])
        off = gpu.blockspmv_cpu(t, rowA, int(g["n"]), np.ones(int(g["n"])))
        for k in ("ptroffset1", "ptroffset2", "blkrow", "colstart", "colstop"):
            assert np.array_equal(np.asarray(off[k]), g[f"{key}_off_{k}"]), (key, k)
        gpu.tile_destroy(t)
        assert t.tilem == 0 and not t.tile_ptr and not t.Blockcsr_Val


def _configs(mf):
    out = {}
    out["laplace2d_256"] = mf.gen_laplace2d(256, 256)
    out["poisson3d_32"] = mf.gen_poisson3d(32, 32, 32)
    out["convdiff27_20"] = mf.gen_convdiff3d(20, 10.0, 1.0, 0)
    out["random_spd_30k"] = mf.gen_random_spd(30000, seed=0)
    out["stencil_ln_3d"] = mf.gen_stencil_lognormal(3, 5000, 1.5, 1)
    # one tile row with 16 x 400 entries: more than the LDS sort holds
    rng = np.random.default_rng(9)
    rows = [np.sort(rng.choice(20000, 400, replace=False)) if i < 16 else np.array([i]) for i in range(64)]
    rp = np.zeros(65, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32)
    out["big_tile_row"] = (rp, ci, rng.standard_normal(len(ci)), 20000)
    return out


@pytest.mark.parametrize("B", [16, 8, 32])
def test_gpu_tile_build_matches_oracle_on_configs(gpu, orc, B):
    for name, c in _configs(gpu).items():
        rp, ci, va = c[:3]
        ncol = c[3] if len(c) > 3 else len(rp) - 1
        m = len(rp) - 1
        print(f"case {name} B={B} n={m} nnz={len(ci)}", flush=True)
        A = gpu.Matrix(rp, ci, va, rowA=m, colA=ncol, B=B)
        d = A.export()
        t = orc.tiles(B, m, ncol, rp, ci, va)
        for k in REF_ARRAYS:
            assert np.array_equal(d[k], t.d[k]), (name, B, k)
        A.close()


@pytest.mark.parametrize("tau,mask", [(0.0, 28), (2.0 ** -11, 28), (2.0 ** -4, 28), (2.0 ** -24, 16), (0.5, 4)])
def test_gpu_classes_match_oracle(gpu, orc, tau, mask):
    for name in ("convdiff27_20", "poisson3d_32", "random_spd_30k"):
        rp, ci, va = _configs(gpu)[name][:3]
        m = len(rp) - 1
        A = gpu.Matrix(rp, ci, va)
        t = orc.tiles(16, m, m, rp, ci, va)
        assert np.array_equal(A.classes(tau, mask), t.classes(tau, mask)), (name, tau, mask)
        A.close()


def test_gpu_mc_classifier_matches_oracle(gpu, orc):
    rp, ci, va = gpu.gen_convdiff3d(16, 10.0, 1.0, 0)
    va = va * np.exp(np.random.default_rng(4).uniform(0.0, 9.0, len(va)))  # spread over all four bins
    m = len(rp) - 1
    A = gpu.Matrix(rp, ci, va)
    mb, counts = A.mc_classify()
    c = orc.tiles(16, m, m, rp, ci, va).compact()
    assert counts == c["mc_counts"] and mb == c["mc_memory_MB"]
    assert all(x > 0 for x in counts)


def test_gpu_layout_info(gpu):
    rp, ci, va = gpu.gen_poisson3d(32, 32, 32)
    A = gpu.Matrix(rp, ci, va)
    A.layout(0.0)
    info = A.info()
    # 6 and -1 are exact in E4M3: the lossless mixed layout is all FP8
    assert info["tiles_per_code"][2] == info["tilenum"]
    assert info["nnz_per_code"][2] == len(ci)
    assert info["bytes_alg_mixed"] < info["bytes_alg_fp64"]


def test_gpu_memory_report(gpu, orc):
    """Memory-Cost figure of the reference format from the oracle's compact
    rows and magnitude bins, beside the slab layouts' bytes: the FP64 slabs
    hold at least 8 bytes per nonzero, the E4M3 slabs of the Poisson matrix
    (values 6 and -1 exact) at least 1 byte and less than half the FP64 ones
    (the stencil's slabs are pattern slabs: one offset table per slab, no
    index per entry)."""
    rp, ci, va = gpu.gen_poisson3d(20, 20, 20)
    m = len(rp) - 1
    A = gpu.Matrix(rp, ci, va)
    r = A.memory_report(0.0)
    t = orc.tiles(16, m, m, rp, ci, va)
    c = t.compact()
    tm, tn, nnz = t.d["tilem"], t.d["tilenum"], len(ci)
    meta = 4 * (tm + 1) + 4 * tn + (tn + 1) + 3 * 4 * (tn + 1) + 8 + 2 * (c["cnt_non_new"] + 1) + tn + nnz + 4 * tn
    MB = 1048576.0
    assert abs(r["ref_memory_MB"] - (meta / MB + c["mc_memory_MB"])) < 1e-9
    assert abs(r["ref_memory_CP_MB"] - 8.0 * c["nonzero_row_new"][-1] / MB) < 1e-12
    assert abs(r["csr_fp64_MB"] - (4 * (m + 1) + 12 * nnz) / MB) < 1e-12
    iw = r["index_width"]
    print(r, flush=True)
    assert r["nnz_per_code"][2] == nnz  # every value exact in E4M3
    assert r["layout_fp64_MB"] * MB >= nnz * 8 and r["layout_mixed_MB"] * MB >= nnz * 1
    assert r["layout_mixed_MB"] < 0.5 * r["layout_fp64_MB"]
    # the same width asked for explicitly gives indexed slabs, which are larger
    A.set_index_width(iw)
    ri = A.memory_report(0.0)
    assert ri["layout_fp64_MB"] * MB >= nnz * (8 + iw) and ri["layout_mixed_MB"] * MB >= nnz * (1 + iw)
    assert r["layout_fp64_MB"] < ri["layout_fp64_MB"] and r["layout_mixed_MB"] < ri["layout_mixed_MB"]
    A.close()
