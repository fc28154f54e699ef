"""The allocation-free solve (mf_solve_device_noalloc) through the C-ABI,
against mf_solve_device on the same inputs.

Both run the same persistent kernels on the same grid (the allocation-free
one publishes BiCGSTAB's vectors through grid barriers, which give the same
iterates as neighbour waits), so iterations and x must be identical bit for
bit.  Two calls in a row use the one static array, the second on a smaller
system; device free memory is unchanged across the allocation-free calls;
sizes beyond the limit and missing layouts are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from fixtures import load_golden

pytestmark = pytest.mark.gpu

TOL = 1e-10
NA_MAX_ROWS = 1 << 21
MF_E_ARG = -10


def _hip_free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _dev(v, npad, fill=0.0):
    d = torch.full((npad,), fill, dtype=torch.float64, device="cuda")
    d[: len(v)] = torch.from_numpy(np.ascontiguousarray(v, np.float64)).to("cuda")
    return d


def _noalloc(gpu, A, d_b, d_x, **opts):
    o = gpu.default_opts(**opts)
    rep = gpu.Report()
    rc = gpu.lib().mf_solve_device_noalloc(A.h, C.c_void_p(d_b.data_ptr()), C.c_void_p(d_x.data_ptr()), C.byref(o),
                                           C.byref(rep))
    return rc, rep.as_dict()


def test_gpu_noalloc_matches_solve_device(gpu, orc):
    g = load_golden("1138_bus")
    systems = [  # larger first, then smaller systems on the same array
        ("poisson3d_40 cg fp64", gpu.gen_poisson3d(40, 40, 40), dict(method=gpu.MF_CG, precision=gpu.MF_FP64)),
        ("poisson3d_40 cg mixed", gpu.gen_poisson3d(40, 40, 40),
         dict(method=gpu.MF_CG, precision=gpu.MF_MIXED, tau=0.0, rr_factor=1e-4)),
        ("convdiff27_24 bicgstab mixed lossy", gpu.gen_convdiff3d(24, 10.0, 1.0, 0),
         dict(method=gpu.MF_BICGSTAB, precision=gpu.MF_MIXED, tau=2.0 ** -11, rr_factor=1e-2)),
        ("1138_bus cg fp64", (g["rowptr"], g["col"], g["val"]), dict(method=gpu.MF_CG, precision=gpu.MF_FP64)),
        ("convdiff27_10 bicgstab fp64", gpu.gen_convdiff3d(10, 10.0, 1.0, 0),
         dict(method=gpu.MF_BICGSTAB, precision=gpu.MF_FP64)),
        # Zipf rows: solved on the row-reordered copy with long rows split
        ("random_spd_30001 cg fp64 reordered", gpu.gen_random_spd(30001, seed=3),
         dict(method=gpu.MF_CG, precision=gpu.MF_FP64)),
    ]
    runs = []
    for name, (rp, ci, va), opts in systems:
        n = len(rp) - 1
        b = orc.csr_spmv(rp, ci, va, gpu.gen_uniform(21, n, -1.0, 1.0))
        A = gpu.Matrix(rp, ci, va)
        info = A.info()
        npad = info["tilem"] * info["B"]
        d_b = _dev(b, npad)
        d_x = _dev([], npad)
        print(f"case {name} n={n}: mf_solve_device", flush=True)
        ref = A.solve_device(d_b.data_ptr(), d_x.data_ptr(), tol=TOL, **opts)  # builds the layouts
        assert ref["status"] == 0 and ref["relres_true"] <= TOL
        runs.append((name, rp, ci, va, b, A, d_b, d_x[:n].cpu().numpy(), ref, _dev([], npad, np.nan), opts))
    torch.cuda.synchronize()
    free0 = _hip_free_bytes()
    for name, rp, ci, va, b, A, d_b, x_ref, ref, d_x2, opts in runs:
        print(f"case {name}: mf_solve_device_noalloc", flush=True)
        rc, rep = _noalloc(gpu, A, d_b, d_x2, tol=TOL, **opts)
        assert rc == 0, gpu.lib().mf_last_error().decode()
        x = d_x2[: len(b)].cpu().numpy()
        print(f"  it={rep['iterations']} (ref {ref['iterations']}) relres={rep['relres_true']:.2e} "
              f"grid={rep['grid']} ms={rep['ms_solve']:.3f}", flush=True)
        assert rep["status"] == 0 and rep["grid"] == ref["grid"] and rep["reordered"] == ref["reordered"]
        assert rep["iterations"] == ref["iterations"] and rep["replacements"] == ref["replacements"]
        assert np.array_equal(x, x_ref)
        assert 0.0 < rep["relres_true"] <= TOL
        A_ = sp.csr_matrix((va, ci, rp), shape=(len(b), len(b)))
        assert np.linalg.norm(b - A_ @ x) / np.linalg.norm(b) <= 1.5 * TOL
    torch.cuda.synchronize()
    assert _hip_free_bytes() == free0  # nothing allocated (nor freed) on the device
    for r in runs:
        r[5].close()


def test_gpu_noalloc_rejects_what_it_cannot_hold(gpu):
    # padded rows beyond the static array
    n = NA_MAX_ROWS + 1
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    va = np.full(n, 2.0)
    A = gpu.Matrix(rp, ci, va)
    d_b = torch.ones(A.info()["tilem"] * 16, dtype=torch.float64, device="cuda")
    d_x = torch.zeros_like(d_b)
    rc, _ = _noalloc(gpu, A, d_b, d_x)
    assert rc == MF_E_ARG and "exceed" in gpu.lib().mf_last_error().decode()
    A.close()
    # layouts not built yet
    rp, ci, va = gpu.gen_poisson3d(8, 8, 8)
    A = gpu.Matrix(rp, ci, va)
    d_b = torch.ones(512, dtype=torch.float64, device="cuda")
    d_x = torch.zeros_like(d_b)
    rc, _ = _noalloc(gpu, A, d_b, d_x)
    assert rc == MF_E_ARG and "layout" in gpu.lib().mf_last_error().decode()
    A.layout(0.0)  # builds the FP64 and the lossless mixed layouts
    rc, rep = _noalloc(gpu, A, d_b, d_x, precision=gpu.MF_MIXED, tau=2.0 ** -11)
    assert rc == MF_E_ARG  # mixed layout of another tau
    rc, rep = _noalloc(gpu, A, d_b, d_x, precision=gpu.MF_MIXED, tau=0.0)
    assert rc == 0 and rep["status"] == 0 and rep["relres_true"] <= 1e-10
    A.close()
