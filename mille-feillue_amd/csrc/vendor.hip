// Vendor-library baseline: CG and BiCGSTAB run the way the reference's
// hipSPARSE drivers run them (PC/hipSPARSE_CG.cu:455-500,
// PC/hipSPARSE_BiCGSTAB.cu:600-710): a host loop issuing one library SpMV
// per product (rocSPARSE generic SpMV, CSR, FP64, adaptive algorithm), library
// dot products (rocBLAS, results left on the device), small kernels for the
// scalar and vector updates, and a copy of ||r||^2 to the host every
// iteration for the stopping test.  It stops on the updated residual and
// then checks the true residual b - A x with the same SpMV (FP64), replacing
// r by it and continuing when the recurrence has drifted.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "mf_matrix.h"

namespace mf {

struct VendorCtx {
    rocsparse_handle sp = nullptr;
    rocblas_handle bl = nullptr;
    int n = 0;
    int *rp = nullptr, *ci = nullptr;  // CSR with the columns of each row sorted
    double *va = nullptr;
    rocsparse_const_spmat_descr mat = nullptr;
    rocsparse_spmv_descr sd = nullptr;
    // work vectors and device scalars
    double *r = nullptr, *p = nullptr, *q = nullptr, *s = nullptr, *t = nullptr, *rh = nullptr, *x = nullptr;
    double *sc = nullptr;
    void *buf = nullptr;
    size_t buf_bytes = 0;
    ~VendorCtx();
};

VendorCtx::~VendorCtx()
{
    if (sd) rocsparse_destroy_spmv_descr(sd);
    if (mat) rocsparse_destroy_spmat_descr(mat);
    if (sp) rocsparse_destroy_handle(sp);
    if (bl) rocblas_destroy_handle(bl);
    void *ps[] = {rp, ci, va, r, p, q, s, t, rh, x, sc, buf};
    for (void *ptr : ps)
        if (ptr) (void)hipFree(ptr);
}

void vendor_release(mf_matrix m)
{
    delete m->vendor;
    m->vendor = nullptr;
}

#define MF_SP(call)                                                          \
    do {                                                                     \
        if ((call) != rocsparse_status_success) return fail(MF_E_HIP, #call); \
    } while (0)

// a library call that can describe its failure (rocsparse_error)
template <typename F> static int sp_call(const char *what, F f)
{
    rocsparse_error err = nullptr;
    const rocsparse_status st = f(&err);
    if (st == rocsparse_status_success) return MF_OK;
    std::string msg = std::string(what) + " (status " + std::to_string((int)st) + ")";
    if (err) {
        const char *m = rocsparse_error_get_message(err);
        if (m) msg += std::string(": ") + m;
        rocsparse_destroy_error(err);
    }
    return fail(MF_E_HIP, msg.c_str());
}
#define MF_BL(call)                                                          \
    do {                                                                     \
        if ((call) != rocblas_status_success) return fail(MF_E_HIP, #call);   \
    } while (0)

// y = A x through the library (dense-vector descriptors made per call)
static int vspmv(VendorCtx &V, const double *x, double *y)
{
    rocsparse_const_dnvec_descr dx = nullptr;
    rocsparse_dnvec_descr dy = nullptr;
    MF_SP(rocsparse_create_const_dnvec_descr(&dx, V.n, x, rocsparse_datatype_f64_r));
    MF_SP(rocsparse_create_dnvec_descr(&dy, V.n, y, rocsparse_datatype_f64_r));
    const double one = 1.0, zero = 0.0;
    const int rc = sp_call("spmv", [&](rocsparse_error *e) {
        return rocsparse_v2_spmv(V.sp, V.sd, &one, V.mat, dx, &zero, dy, rocsparse_v2_spmv_stage_compute, V.buf_bytes,
                                 V.buf, e);
    });
    rocsparse_destroy_dnvec_descr(dx);
    rocsparse_destroy_dnvec_descr(dy);
    return rc;
}

static int vendor_setup(mf_matrix m)
{
    if (m->vendor) return MF_OK;
    const int n = m->rowA, nnz = m->nnz;
    // the library's CSR wants sorted columns; the Matrix Market order is the
    // file order, so sort each row (duplicates stay and are summed)
    std::vector<int> rp(n + 1), ci(std::max(nnz, 1));
    std::vector<double> va(std::max(nnz, 1));
    MF_HIP(hipMemcpy(rp.data(), m->d_rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (nnz) {
        MF_HIP(hipMemcpy(ci.data(), m->d_col, sizeof(int) * nnz, hipMemcpyDeviceToHost));
        MF_HIP(hipMemcpy(va.data(), m->d_val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < n; i++) {
        const int a = rp[i], b = rp[i + 1];
        std::vector<int> idx(b - a);
        std::iota(idx.begin(), idx.end(), a);
        std::stable_sort(idx.begin(), idx.end(), [&](int u, int w) { return ci[u] < ci[w]; });
        std::vector<int> c2(b - a);
        std::vector<double> v2(b - a);
        for (int k = 0; k < b - a; k++) {
            c2[k] = ci[idx[k]];
            v2[k] = va[idx[k]];
        }
        std::copy(c2.begin(), c2.end(), ci.begin() + a);
        std::copy(v2.begin(), v2.end(), va.begin() + a);
    }
    // attached to the matrix only once every step has succeeded
    std::unique_ptr<VendorCtx> owner(new VendorCtx());
    VendorCtx *V = owner.get();
    V->n = n;
    const size_t vb = sizeof(double) * (size_t)std::max(n, 1);
    MF_HIP(hipMalloc((void **)&V->rp, sizeof(int) * (n + 1)));
    MF_HIP(hipMalloc((void **)&V->ci, sizeof(int) * ci.size()));
    MF_HIP(hipMalloc((void **)&V->va, sizeof(double) * va.size()));
    MF_HIP(hipMemcpy(V->rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(V->ci, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(V->va, va.data(), sizeof(double) * va.size(), hipMemcpyHostToDevice));
    double **vecs[] = {&V->r, &V->p, &V->q, &V->s, &V->t, &V->rh, &V->x};
    for (double **v : vecs) MF_HIP(hipMalloc((void **)v, vb));
    MF_HIP(hipMalloc((void **)&V->sc, sizeof(double) * 16));
    MF_SP(rocsparse_create_handle(&V->sp));
    MF_SP(rocsparse_set_stream(V->sp, stream()));
    MF_BL(rocblas_create_handle(&V->bl));
    MF_BL(rocblas_set_stream(V->bl, stream()));
    MF_BL(rocblas_set_pointer_mode(V->bl, rocblas_pointer_mode_device));
    MF_SP(rocsparse_create_const_csr_descr(&V->mat, n, n, nnz, V->rp, V->ci, V->va, rocsparse_indextype_i32,
                                           rocsparse_indextype_i32, rocsparse_index_base_zero,
                                           rocsparse_datatype_f64_r));
    MF_SP(rocsparse_create_spmv_descr(&V->sd));
    const rocsparse_spmv_alg alg = rocsparse_spmv_alg_csr_adaptive;
    const rocsparse_operation op = rocsparse_operation_none;
    const rocsparse_datatype dt = rocsparse_datatype_f64_r;
    MF_SP(rocsparse_spmv_set_input(V->sp, V->sd, rocsparse_spmv_input_alg, &alg, sizeof(alg), nullptr));
    MF_SP(rocsparse_spmv_set_input(V->sp, V->sd, rocsparse_spmv_input_operation, &op, sizeof(op), nullptr));
    MF_SP(rocsparse_spmv_set_input(V->sp, V->sd, rocsparse_spmv_input_scalar_datatype, &dt, sizeof(dt), nullptr));
    MF_SP(rocsparse_spmv_set_input(V->sp, V->sd, rocsparse_spmv_input_compute_datatype, &dt, sizeof(dt), nullptr));
    // analysis once (the adaptive algorithm's row binning), then the compute buffer
    rocsparse_const_dnvec_descr dx = nullptr;
    rocsparse_dnvec_descr dy = nullptr;
    MF_SP(rocsparse_create_const_dnvec_descr(&dx, n, V->p, dt));
    MF_SP(rocsparse_create_dnvec_descr(&dy, n, V->q, dt));
    size_t ba = 0, bc = 0;
    int rc = sp_call("spmv buffer size (analysis)", [&](rocsparse_error *e) {
        return rocsparse_v2_spmv_buffer_size(V->sp, V->sd, V->mat, dx, dy, rocsparse_v2_spmv_stage_analysis, &ba, e);
    });
    if (!rc)
        rc = sp_call("spmv buffer size (compute)", [&](rocsparse_error *e) {
            return rocsparse_v2_spmv_buffer_size(V->sp, V->sd, V->mat, dx, dy, rocsparse_v2_spmv_stage_compute, &bc, e);
        });
    if (rc) return rc;
    V->buf_bytes = std::max<size_t>(std::max(ba, bc), 256);
    MF_HIP(hipMalloc(&V->buf, V->buf_bytes));
    const double one = 1.0, zero = 0.0;
    rc = sp_call("spmv analysis", [&](rocsparse_error *e) {
        return rocsparse_v2_spmv(V->sp, V->sd, &one, V->mat, dx, &zero, dy, rocsparse_v2_spmv_stage_analysis,
                                 V->buf_bytes, V->buf, e);
    });
    if (rc) return rc;
    rocsparse_destroy_dnvec_descr(dx);
    rocsparse_destroy_dnvec_descr(dy);
    MF_HIP(hipStreamSynchronize(stream()));
    m->vendor = owner.release();
    return MF_OK;
}

// ---- scalar and vector kernels (device scalars, no host round trip)
// sc: 0 rho/gamma, 1 (p,q) or (rh,v), 2 alpha, 3 (t,s), 4 (t,t), 5 omega,
//     6 ||r||^2, 7 rho_new, 8 beta
__global__ void kv_cg_alpha(double *sc) { sc[2] = sc[0] / sc[1]; }
__global__ void kv_cg_beta(double *sc)
{
    sc[8] = sc[6] / sc[0];
    sc[0] = sc[6];
}
__global__ void kv_bcg_alpha(double *sc) { sc[2] = sc[0] / sc[1]; }
__global__ void kv_bcg_omega(double *sc) { sc[5] = sc[4] > 0.0 ? sc[3] / sc[4] : 0.0; }
__global__ void kv_bcg_beta(double *sc)
{
    sc[8] = (sc[7] / sc[0]) * (sc[2] / sc[5]);
    sc[0] = sc[7];
}

// CG: x += alpha p; r -= alpha q
__global__ void kv_cg_xr(int n, const double *__restrict__ sc, const double *__restrict__ p,
                         const double *__restrict__ q, double *__restrict__ x, double *__restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double a = sc[2];
        x[i] += a * p[i];
        r[i] -= a * q[i];
    }
}
// CG: p = r + beta p
__global__ void kv_cg_p(int n, const double *__restrict__ sc, const double *__restrict__ r, double *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + sc[8] * p[i];
}
// BiCGSTAB: s = r - alpha v
__global__ void kv_bcg_s(int n, const double *__restrict__ sc, const double *__restrict__ r,
                         const double *__restrict__ v, double *__restrict__ s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s[i] = r[i] - sc[2] * v[i];
}
// BiCGSTAB: x += alpha p + omega s; r = s - omega t
__global__ void kv_bcg_xr(int n, const double *__restrict__ sc, const double *__restrict__ p,
                          const double *__restrict__ s, const double *__restrict__ t, double *__restrict__ x,
                          double *__restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double a = sc[2], w = sc[5];
        x[i] += a * p[i] + w * s[i];
        r[i] = s[i] - w * t[i];
    }
}
// BiCGSTAB: p = r + beta (p - omega v)
__global__ void kv_bcg_p(int n, const double *__restrict__ sc, const double *__restrict__ r,
                         const double *__restrict__ v, double *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + sc[8] * (p[i] - sc[5] * v[i]);
}
// r = b - r (r holds A x on entry)
__global__ void kv_resid(int n, const double *__restrict__ b, double *__restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - r[i];
}

// True ||b - A x||^2 into sc[idx] (r overwritten with the residual).
static int true_resid(VendorCtx &V, const double *b, int idx)
{
    const unsigned nb = (unsigned)((V.n + 255) / 256);
    int rc = vspmv(V, V.x, V.r);
    if (rc) return rc;
    kv_resid<<<nb, 256, 0, stream()>>>(V.n, b, V.r);
    MF_BL(rocblas_ddot(V.bl, V.n, V.r, 1, V.r, 1, V.sc + idx));
    return MF_OK;
}

int vendor_solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep)
{
    std::memset(rep, 0, sizeof(*rep));
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    double t0 = now_ms();
    int rc = vendor_setup(m);
    if (rc) return rc;
    rep->ms_setup = now_ms() - t0;
    VendorCtx &V = *m->vendor;
    hipStream_t st = stream();
    const int n = V.n;
    const unsigned nb = (unsigned)((n + 255) / 256);
    const size_t vb = sizeof(double) * (size_t)n;
    hipEvent_t e0, e1;
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    MF_HIP(hipEventRecord(e0, st));
    // x = 0, r = b
    MF_HIP(hipMemsetAsync(V.x, 0, vb, st));
    MF_HIP(hipMemcpyAsync(V.r, d_b, vb, hipMemcpyDeviceToDevice, st));
    MF_BL(rocblas_ddot(V.bl, n, d_b, 1, d_b, 1, V.sc + 9));
    double h[2] = {0, 0};
    MF_HIP(hipMemcpyAsync(h, V.sc + 9, sizeof(double), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    const double bn2 = h[0], tol2 = o.tol * o.tol * bn2;
    int iters = 0, repl = 0, status = MF_E_MAXIT;
    double rr = bn2;
    bool fresh = true;
    if (bn2 == 0.0) status = MF_OK;
    while (status == MF_E_MAXIT && iters < o.maxit) {
        if (fresh) {
            // new direction from the current r
            MF_HIP(hipMemcpyAsync(V.p, V.r, vb, hipMemcpyDeviceToDevice, st));
            MF_BL(rocblas_ddot(V.bl, n, V.r, 1, V.r, 1, V.sc + 0));
            if (o.method == MF_BICGSTAB) MF_HIP(hipMemcpyAsync(V.rh, V.r, vb, hipMemcpyDeviceToDevice, st));
            fresh = false;
        }
        if (o.method == MF_CG) {
            rc = vspmv(V, V.p, V.q);  // q = A p
            if (rc) return rc;
            MF_BL(rocblas_ddot(V.bl, n, V.p, 1, V.q, 1, V.sc + 1));
            kv_cg_alpha<<<1, 1, 0, st>>>(V.sc);
            kv_cg_xr<<<nb, 256, 0, st>>>(n, V.sc, V.p, V.q, V.x, V.r);
            MF_BL(rocblas_ddot(V.bl, n, V.r, 1, V.r, 1, V.sc + 6));
            kv_cg_beta<<<1, 1, 0, st>>>(V.sc);
            kv_cg_p<<<nb, 256, 0, st>>>(n, V.sc, V.r, V.p);
        } else {
            rc = vspmv(V, V.p, V.q);  // v = A p
            if (rc) return rc;
            MF_BL(rocblas_ddot(V.bl, n, V.rh, 1, V.q, 1, V.sc + 1));
            kv_bcg_alpha<<<1, 1, 0, st>>>(V.sc);
            kv_bcg_s<<<nb, 256, 0, st>>>(n, V.sc, V.r, V.q, V.s);
            rc = vspmv(V, V.s, V.t);  // t = A s
            if (rc) return rc;
            MF_BL(rocblas_ddot(V.bl, n, V.t, 1, V.s, 1, V.sc + 3));
            MF_BL(rocblas_ddot(V.bl, n, V.t, 1, V.t, 1, V.sc + 4));
            kv_bcg_omega<<<1, 1, 0, st>>>(V.sc);
            kv_bcg_xr<<<nb, 256, 0, st>>>(n, V.sc, V.p, V.s, V.t, V.x, V.r);
            MF_BL(rocblas_ddot(V.bl, n, V.r, 1, V.r, 1, V.sc + 6));
            MF_BL(rocblas_ddot(V.bl, n, V.rh, 1, V.r, 1, V.sc + 7));
            kv_bcg_beta<<<1, 1, 0, st>>>(V.sc);
            kv_bcg_p<<<nb, 256, 0, st>>>(n, V.sc, V.r, V.q, V.p);
        }
        ++iters;
        // the stopping test reads ||r||^2 on the host, as the reference's driver does
        MF_HIP(hipMemcpyAsync(h, V.sc + 6, sizeof(double), hipMemcpyDeviceToHost, st));
        MF_HIP(hipStreamSynchronize(st));
        rr = h[0];
        if (!std::isfinite(rr)) { status = MF_E_NONFINITE; break; }
        if (rr <= tol2) {
            rc = true_resid(V, d_b, 10);
            if (rc) return rc;
            MF_HIP(hipMemcpyAsync(h + 1, V.sc + 10, sizeof(double), hipMemcpyDeviceToHost, st));
            MF_HIP(hipStreamSynchronize(st));
            ++repl;
            if (h[1] <= tol2) status = MF_OK;
            else fresh = true;  // drifted: continue from the true residual
        }
    }
    MF_HIP(hipMemcpyAsync(d_x, V.x, vb, hipMemcpyDeviceToDevice, st));
    MF_HIP(hipEventRecord(e1, st));
    MF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    MF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // independent check with the library SpMV
    rc = true_resid(V, d_b, 11);
    if (rc) return rc;
    MF_HIP(hipMemcpyAsync(h, V.sc + 11, sizeof(double), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    rep->relres_true = bn2 > 0 ? std::sqrt(h[0] / bn2) : 0.0;
    rep->status = status;
    rep->iterations = iters;
    rep->outer = 1;
    rep->replacements = repl;
    rep->ms_solve = ms;
    return status == MF_OK || status == MF_E_MAXIT ? MF_OK : fail(status, "vendor solve failed");
}

}  // namespace mf
