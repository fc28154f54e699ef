"""Batched persistent CG (many systems x several right-hand sides in one
graph-captured launch) against the oracle's FP64 CG per system and column
(needs an MI355X).

Every column of every system must reach true relative residual 1e-10
(reported by the kernel's FP64 pass and re-checked here with scipy), and a
system's iteration count (its slowest column) must be within 5% (+2) of the
oracle's CG on that column."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _systems(mf, sizes, seed=0):
    out = []
    for i, n in enumerate(sizes):
        dim = 2 if i % 2 == 0 else 3
        out.append(mf.gen_stencil_lognormal(dim, n, 1.5, seed + i))
    return out


def _rhs(systems, nrhs, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((nrhs, len(s[0]) - 1)) for s in systems]


def _check_batch(orc, systems, b, xs, it, rel, rep, check_iters=True):
    assert rep["status"] == 0, rep
    for s, (rp, ci, va) in enumerate(systems):
        n = len(rp) - 1
        A = sp.csr_matrix((va, ci, rp), shape=(n, n))
        bs = np.reshape(b[s], (-1, n))
        x = np.reshape(xs[s], (-1, n))
        worst = max(np.linalg.norm(bs[k] - A @ x[k]) / np.linalg.norm(bs[k]) for k in range(bs.shape[0]))
        assert rel[s] <= TOL and worst <= 1.5 * TOL, (s, rel[s], worst)
        if check_iters:
            its = [orc.cg(rp, ci, va, bs[k], TOL, 20000)[1].iterations for k in range(bs.shape[0])]
            print(f"system {s} n={n} gpu it={it[s]} oracle it={max(its)} rel={rel[s]:.2e}", flush=True)
            assert abs(it[s] - max(its)) <= 0.05 * max(its) + 2, (s, it[s], its)


def test_gpu_batched_fp64_eight_rhs(gpu, orc):
    systems = _systems(gpu, [300, 1200, 900, 2500, 4096, 700])
    b = _rhs(systems, 8, 1)
    bt = gpu.Batch(systems, nrhs=8, tol=TOL)
    info = bt.info()
    print(info, flush=True)
    xs, it, rel, rep = bt.solve(b)
    print(rep, flush=True)
    _check_batch(orc, systems, b, xs, it, rel, rep)
    # the same graph again: identical results
    xs2, it2, rel2, _ = bt.solve(b)
    assert np.array_equal(it, it2) and all(np.array_equal(a, c) for a, c in zip(xs, xs2))
    bt.close()


def test_gpu_batched_single_rhs_and_tile_sizes(gpu, orc):
    systems = _systems(gpu, [500, 2000, 1500], seed=7)
    b = _rhs(systems, 1, 2)
    b1 = [bb[0] for bb in b]
    ref_it = None
    for B in (16, 8, 32):
        bt = gpu.Batch(systems, nrhs=1, tol=TOL, block=B)
        xs, it, rel, rep = bt.solve(b1)
        print(f"B={B} it={it.tolist()} rel={rel.tolist()}", flush=True)
        _check_batch(orc, systems, b1, xs, it, rel, rep, check_iters=(B == 16))
        if ref_it is None:
            ref_it = it
        assert np.all(np.abs(it - ref_it) <= 0.05 * ref_it + 2)
        bt.close()


def test_gpu_batched_multi_workgroup_groups(gpu, orc):
    # a large grid target splits the bigger systems over several workgroups
    # that synchronise through the tagged-word group allreduce
    systems = _systems(gpu, [4000, 9000, 600], seed=3)
    b = _rhs(systems, 8, 4)
    bt = gpu.Batch(systems, nrhs=8, tol=TOL, grid=96)
    info = bt.info()
    print(info, flush=True)
    assert info["G"] > len(systems)
    xs, it, rel, rep = bt.solve(b)
    _check_batch(orc, systems, b, xs, it, rel, rep)
    bt.close()


def test_gpu_batched_common_index_width(gpu, orc):
    # the 27-point system (18 tiles per tile row) does not fit the one-byte
    # slot form, so a request for one byte gives the two-byte form for all;
    # the four-byte column form (rebased per system when the layouts are
    # joined) gives the same iterates bit for bit; the Laplacians alone run
    # in the slot form
    cd = gpu.gen_convdiff3d(32, 0.0, 0.0, 0)  # symmetric (no convection, constant diffusion)
    lap = [gpu.gen_laplace2d(30, 30)] + _systems(gpu, [1500], seed=11)
    systems = [lap[0], cd, lap[1]]
    b = _rhs(systems, 8, 12)
    out = {}
    for w, want in ((0, 2), (1, 2), (4, 4)):
        bt = gpu.Batch(systems, nrhs=8, tol=TOL, index_width=w)
        info = bt.info()
        print(f"width {w}: {info['index_width']}", flush=True)
        assert info["index_width"] == want
        xs, it, rel, rep = bt.solve(b)
        _check_batch(orc, systems, b, xs, it, rel, rep, check_iters=(w == 0))
        out[w] = (xs, it)
        bt.close()
    for w in (1, 4):
        assert np.array_equal(out[0][1], out[w][1])
        assert all(np.array_equal(a, c) for a, c in zip(out[0][0], out[w][0]))
    bl = [b[0], b[2]]
    bt = gpu.Batch(lap, nrhs=8, tol=TOL, index_width=1)
    assert bt.info()["index_width"] == 1
    xs, it, rel, rep = bt.solve(bl)
    bt.close()
    _check_batch(orc, lap, bl, xs, it, rel, rep)


def test_gpu_batched_mixed_lossy_tiles(gpu):
    # FP16-rounded tiles (tau = 2^-11) as the inner operator; the FP64
    # true-residual passes still bring every column to 1e-10
    systems = [gpu.gen_poisson3d(12, 12, 12), gpu.gen_laplace2d(40, 40)] + _systems(gpu, [800], seed=5)
    b = _rhs(systems, 8, 6)
    bt = gpu.Batch(systems, nrhs=8, tol=TOL, precision=gpu.MF_MIXED, tau=2.0 ** -11, rr_factor=1e-2)
    info = bt.info()
    print(info, flush=True)
    assert sum(info["nnz_per_code"][2:]) > 0  # some low-precision tiles
    xs, it, rel, rep = bt.solve(b)
    _check_batch(None, systems, b, xs, it, rel, rep, check_iters=False)
    bt.close()


def test_gpu_batched_on_chip_matches_hbm(gpu, orc):
    """The on-chip kernel (vectors = MF_VECTORS_ONCHIP: each system on a gang
    of workgroups with its vectors in LDS) and the default HBM-vector kernel
    on the same batch: every column at 1e-10 in both, iterations within 5%
    (+2) of each other and of the oracle; gangs of up to 18 workgroups (9000
    rows at 512 rows each)."""
    systems = _systems(gpu, [4000, 9000, 600, 1500], seed=3)
    b = _rhs(systems, 8, 4)
    out = {}
    for vec in (gpu.MF_VECTORS_ONCHIP, gpu.MF_VECTORS_AUTO):
        bt = gpu.Batch(systems, nrhs=8, tol=TOL, vectors=vec)
        info = bt.info()
        print(f"vectors={vec} on_chip={info['on_chip']} G={info['G']}", flush=True)
        assert info["on_chip"] == (1 if vec == gpu.MF_VECTORS_ONCHIP else 0)
        xs, it, rel, rep = bt.solve(b)
        _check_batch(orc, systems, b, xs, it, rel, rep, check_iters=(vec == gpu.MF_VECTORS_ONCHIP))
        out[vec] = it
        bt.close()
    a, h = out[gpu.MF_VECTORS_ONCHIP], out[gpu.MF_VECTORS_AUTO]
    assert np.all(np.abs(a - h) <= 0.05 * h + 2), (a, h)


def test_gpu_batched_on_chip_queued_gangs(gpu):
    """More gang slots than CUs (30 systems of ~9000 rows need ~540
    workgroup tasks on 256 workgroups): the scheduled order runs every gang
    without a cycle of waits, twice in a row on the captured graph."""
    systems = _systems(gpu, [9000 + 37 * i for i in range(30)], seed=21)
    b = _rhs(systems, 8, 22)
    bt = gpu.Batch(systems, nrhs=8, tol=TOL, vectors=gpu.MF_VECTORS_ONCHIP)
    assert bt.info()["on_chip"] == 1
    xs, it, rel, rep = bt.solve(b)
    print(f"queued gangs: it max {it.max()} ms {rep['ms_solve']:.1f}", flush=True)
    _check_batch(None, systems, b, xs, it, rel, rep, check_iters=False)
    xs2, it2, rel2, rep2 = bt.solve(b)
    assert rep2["status"] == 0 and np.array_equal(it, it2)
    assert all(np.array_equal(p, q) for p, q in zip(xs, xs2))
    bt.close()


def test_gpu_batched_on_chip_refusals(gpu):
    """MF_VECTORS_ONCHIP needs two or more right-hand sides (the kernel's 8
    columns); a single system that fits on chip solves with it, and an
    out-of-range vectors value is refused."""
    systems = _systems(gpu, [900, 1200], seed=5)
    with pytest.raises(gpu.MFError):
        gpu.Batch(systems, nrhs=1, tol=TOL, vectors=gpu.MF_VECTORS_ONCHIP)
    with pytest.raises(gpu.MFError):
        gpu.Batch(systems, nrhs=8, tol=TOL, vectors=7)
    rp, ci, va = systems[0]
    A = gpu.Matrix(rp, ci, va)
    x, rep = A.solve(np.ones(len(rp) - 1), method=gpu.MF_CG, tol=TOL, vectors=gpu.MF_VECTORS_ONCHIP)
    assert rep["status"] == 0 and rep["streamed"] == 0
    with pytest.raises(gpu.MFError):
        A.solve(np.ones(len(rp) - 1), method=gpu.MF_CG, tol=TOL, vectors=7)
    A.close()
