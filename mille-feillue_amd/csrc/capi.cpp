// C-ABI of libmfamd.so: device set-up, matrix handles, exports, SpMV and solve
// entry points, and the reference drop-ins (Tile_create, cg_solve_inc,
// bicgstab) that sit on top of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "mf_matrix.h"

namespace mf {

int solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep);
int noalloc_setup(int dev);  // solver.hip
void noalloc_release();      // solver.hip
int solve_noalloc(mf_matrix m, int dev, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep);
int spmv_device(mf_matrix m, const Layout &L, const double *d_x, double *d_y, bool convert);

int hip_fail(hipError_t e, const char *what)
{
    return fail(MF_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static std::mutex g_mu;
static int g_device = -1, g_cus = 0;
static hipStream_t g_stream = nullptr;

hipStream_t stream() { return g_stream; }

static std::mutex g_live_mu;
static std::set<void *> g_live[2];
void live_add(void *h, int kind)
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    g_live[kind].insert(h);
}
bool live_remove(void *h, int kind)
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    return g_live[kind].erase(h) > 0;
}
bool live_has(void *h, int kind)
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    return g_live[kind].count(h) > 0;
}
int device_cu_count() { return g_cus; }

static int init_device(int dev)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_device >= 0 && g_device == dev) return MF_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(MF_E_NODEV, "no HIP device visible");
    if (dev < 0 || dev >= count) return fail(MF_E_NODEV, "device index out of range");
    MF_HIP(hipSetDevice(dev));
    hipDeviceProp_t p;
    MF_HIP(hipGetDeviceProperties(&p, dev));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(MF_E_NODEV, std::string("built for gfx950, found ") + p.gcnArchName);
    int coop = 0;
    MF_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) return fail(MF_E_NODEV, "device lacks cooperative launch");
    g_cus = p.multiProcessorCount;
    if (g_stream) (void)hipStreamDestroy(g_stream);
    MF_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // the static buffers of the allocation-free solve on this device
    const int rc = noalloc_setup(dev);
    if (rc) return rc;
    g_device = dev;
    return MF_OK;
}

int ensure_device()
{
    if (g_device >= 0) {
        MF_HIP(hipSetDevice(g_device));
        return MF_OK;
    }
    const char *env = std::getenv("MF_DEVICE");
    return init_device(env ? std::atoi(env) : 0);
}

}  // namespace mf

using namespace mf;

static void free_dev(mf_matrix m)
{
    void *ps[] = {m->d_rowptr, m->d_col, m->d_val, m->d_val_low, m->d_tile_ptr, m->d_tile_columnidx,
                  m->d_tile_start, m->d_csrptr_offset, m->d_blknnznnz, m->d_Blockcsr_Ptr,
                  m->d_Tile_csr_Col, m->d_rowtmp, m->d_Blockcsr_Val, m->d_Blockcsr_Val_Low};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    void *qs[] = {m->d_perm, m->d_pb, m->d_px, m->d_split};
    for (void *p : qs)
        if (p) (void)hipFree(p);
    if (m->child) mf_matrix_destroy(m->child);
    m->child = nullptr;
    mf::vendor_release(m);
    m->L64.release();
    m->Lmix.release();
    m->ws.release();
}

extern "C" {

int mf_init(int device)
{
    clear_error();
    return init_device(device);
}

void mf_opts_default(mf_opts *o)
{
    std::memset(o, 0, sizeof(*o));
    o->method = MF_CG;
    o->precision = MF_FP64;
    o->tol = 1e-10;
    o->maxit = 10000;
    o->tau = 0.0;
    o->allow_mask = (1 << 2) | (1 << 3) | (1 << 4);
    o->rr_factor = 0.0;
    o->ir_max_outer = 0;
    o->ir_inner_tol = 1e-6;
    o->grid = 0;
    o->block = 0;
    o->spin_limit = 1LL << 22;  // a few seconds of polling
    o->index_width = 0;
    o->reorder = MF_REORDER_AUTO;
    o->sync = MF_SYNC_AUTO;
}

int mf_matrix_create(int rowA, int colA, int nnz, const int *rowptr, const int *colidx,
                     const double *val, const float *val_low, int B, mf_matrix *out)
{
    clear_error();
    *out = nullptr;
    if (B != 8 && B != 16 && B != 32) return fail(MF_E_ARG, "tile size must be 8, 16 or 32");
    if (rowA <= 0 || colA <= 0 || !rowptr) return fail(MF_E_ARG, "empty matrix");
    if (rowptr[0] != 0) return fail(MF_E_ARG, "rowptr[0] must be 0");
    const int nz = rowptr[rowA];
    if (nz < 0 || (nnz >= 0 && nz > nnz)) return fail(MF_E_ARG, "rowptr inconsistent with nnz");
    if (nz > 0 && (!colidx || !val)) return fail(MF_E_ARG, "null column or value array");
    // host-side bounds check: every index the kernels will derive stays inside
    // the arrays they read
    int bad = 0;
    omp_quiet();
#pragma omp parallel for reduction(| : bad)
    for (int i = 0; i < rowA; i++)
        if (rowptr[i + 1] < rowptr[i]) bad |= 1;
    if (bad) return fail(MF_E_ARG, "rowptr not non-decreasing");
#pragma omp parallel for reduction(| : bad)
    for (long long k = 0; k < nz; k++)
        if (colidx[k] < 0 || colidx[k] >= colA) bad |= 1;
    if (bad) return fail(MF_E_ARG, "column index outside [0, colA)");
    // the kernels address vectors with 32-bit byte offsets
    const long long npad = ((long long)std::max(rowA, colA) + B - 1) / B * B;
    if (npad * 8 >= (1ll << 32)) return fail(MF_E_UNSUPPORTED, "vectors of 4 GB or more");
    int rc = ensure_device();
    if (rc) return rc;
    mf_matrix m = new mf_matrix_s();
    m->B = B;
    m->rowA = rowA;
    m->colA = colA;
    m->nnz = nz;
    m->csrsize = nz;
    m->tilem = (rowA + B - 1) / B;
    m->tilen = (colA + B - 1) / B;
    double t0 = now_ms();
    auto bail = [&](int code) {
        free_dev(m);
        delete m;
        return code;
    };
    const size_t nzs = (size_t)std::max(nz, 1);
    if (hipMalloc((void **)&m->d_rowptr, sizeof(int) * (rowA + 1)) != hipSuccess ||
        hipMalloc((void **)&m->d_col, sizeof(int) * nzs) != hipSuccess ||
        hipMalloc((void **)&m->d_val, sizeof(double) * nzs) != hipSuccess)
        return bail(fail(MF_E_NOMEM, "device allocation of the CSR"));
    if (hipMemcpy(m->d_rowptr, rowptr, sizeof(int) * (rowA + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (nz && hipMemcpy(m->d_col, colidx, sizeof(int) * nz, hipMemcpyHostToDevice) != hipSuccess) ||
        (nz && hipMemcpy(m->d_val, val, sizeof(double) * nz, hipMemcpyHostToDevice) != hipSuccess))
        return bail(fail(MF_E_HIP, "upload of the CSR"));
    if (val_low) {
        if (hipMalloc((void **)&m->d_val_low, sizeof(float) * nzs) != hipSuccess ||
            (nz && hipMemcpy(m->d_val_low, val_low, sizeof(float) * nz, hipMemcpyHostToDevice) != hipSuccess))
            return bail(fail(MF_E_HIP, "upload of the low values"));
    }
    rc = build_reference_tiles(m);
    if (rc) return bail(rc);
    m->ms_build = now_ms() - t0;
    live_add(m, 0);
    *out = m;
    return MF_OK;
}

void mf_matrix_destroy(mf_matrix m)
{
    if (!m || !live_remove(m, 0)) return;
    free_dev(m);
    delete m;
}

// Releases every matrix and batch still held, the library's device buffers
// and its stream, before the process tears the runtime down (bindings call it
// at interpreter exit).  Handles released here may still be passed to the
// destroy functions, which then do nothing.  The next entry point sets the
// device up again.
void mf_finalize(void)
{
    std::vector<void *> bs, ms;
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        bs.assign(g_live[1].begin(), g_live[1].end());
        ms.assign(g_live[0].begin(), g_live[0].end());
    }
    for (void *b : bs) mf_batch_destroy((mf_batch)b);
    for (void *m : ms) mf_matrix_destroy((mf_matrix)m);  // children go with their parents
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_device >= 0) {
        (void)hipSetDevice(g_device);
        mf::release_phase_buffer();
        mf::noalloc_release();
        if (g_stream) (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
        g_device = -1;
    }
}

int mf_matrix_export(mf_matrix m, mf_ref_tiles *t)
{
    clear_error();
    if (!m || !t) return fail(MF_E_ARG, "null");
    std::memset(t, 0, sizeof(*t));
    t->B = m->B; t->rowA = m->rowA; t->colA = m->colA; t->tilem = m->tilem; t->tilen = m->tilen;
    t->tilenum = m->tilenum; t->csrsize = m->csrsize; t->csrptrlen = m->csrptrlen;
    const size_t n1 = (size_t)m->tilenum + 1, nz = (size_t)m->csrsize;
    t->tile_ptr = (int *)std::malloc(sizeof(int) * (m->tilem + 1));
    t->tile_columnidx = (int *)std::malloc(sizeof(int) * n1);
    t->tile_nnz = (int *)std::malloc(sizeof(int) * n1);
    t->Format = (char *)std::calloc(n1, 1);
    t->blknnz = (int *)std::malloc(sizeof(int) * n1);
    t->blknnznnz = (unsigned char *)std::malloc(n1);
    t->csr_offset = (int *)std::malloc(sizeof(int) * n1);
    t->csrptr_offset = (int *)std::malloc(sizeof(int) * n1);
    t->Blockcsr_Val = (double *)std::malloc(sizeof(double) * (nz + 1));
    t->Blockcsr_Val_Low = (float *)std::malloc(sizeof(float) * (nz + 1));
    t->Blockcsr_Ptr = (unsigned char *)std::malloc((size_t)m->csrptrlen + 1);
    t->Tile_csr_Col = (unsigned char *)std::malloc(nz + 1);
    t->csr_compressedIdx = (unsigned char *)std::calloc(nz / 2 + 2, 1);
    MF_HIP(hipMemcpy(t->tile_ptr, m->d_tile_ptr, sizeof(int) * (m->tilem + 1), hipMemcpyDeviceToHost));
    MF_HIP(hipMemcpy(t->tile_columnidx, m->d_tile_columnidx, sizeof(int) * m->tilenum, hipMemcpyDeviceToHost));
    MF_HIP(hipMemcpy(t->tile_nnz, m->d_tile_start, sizeof(int) * n1, hipMemcpyDeviceToHost));
    std::memcpy(t->blknnz, t->tile_nnz, sizeof(int) * n1);
    std::memcpy(t->csr_offset, t->tile_nnz, sizeof(int) * n1);
    MF_HIP(hipMemcpy(t->blknnznnz, m->d_blknnznnz, n1, hipMemcpyDeviceToHost));
    MF_HIP(hipMemcpy(t->csrptr_offset, m->d_csrptr_offset, sizeof(int) * n1, hipMemcpyDeviceToHost));
    if (nz) {
        MF_HIP(hipMemcpy(t->Blockcsr_Val, m->d_Blockcsr_Val, sizeof(double) * nz, hipMemcpyDeviceToHost));
        MF_HIP(hipMemcpy(t->Blockcsr_Val_Low, m->d_Blockcsr_Val_Low, sizeof(float) * nz, hipMemcpyDeviceToHost));
        MF_HIP(hipMemcpy(t->Tile_csr_Col, m->d_Tile_csr_Col, nz, hipMemcpyDeviceToHost));
    }
    if (m->csrptrlen)
        MF_HIP(hipMemcpy(t->Blockcsr_Ptr, m->d_Blockcsr_Ptr, m->csrptrlen, hipMemcpyDeviceToHost));
    // The reference copies a tile row by row from the uint8 row starts
    // (PC/csr2block.h:239-253): row ri covers [start[ri], next start or nnz).
    // When the starts wrapped or were never scanned, some positions are not
    // covered and keep the zero of the reference's memset.
    for (int bi = 0; bi < m->tilem; bi++) {
        const int rowlen = bi == m->tilem - 1 ? m->rowA - (m->tilem - 1) * m->B : m->B;
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            const int nz = t->tile_nnz[tid + 1] - t->tile_nnz[tid];
            if (nz <= 255 && rowlen > 1) continue;
            const int off = t->csr_offset[tid], po = t->csrptr_offset[tid];
            for (int k = 0; k < nz; k++) {
                bool cov = false;
                for (int ri = 0; ri < rowlen && !cov; ri++) {
                    const int a = t->Blockcsr_Ptr[po + ri];
                    const int b = ri == rowlen - 1 ? nz : t->Blockcsr_Ptr[po + ri + 1];
                    cov = a <= k && k < b;
                }
                if (!cov) {
                    t->Blockcsr_Val[off + k] = 0.0;
                    t->Blockcsr_Val_Low[off + k] = 0.0f;
                    t->Tile_csr_Col[off + k] = 0;
                }
            }
        }
    }
    return MF_OK;
}

void mf_ref_tiles_free(mf_ref_tiles *t)
{
    if (!t) return;
    void *ps[] = {t->tile_ptr, t->tile_columnidx, t->tile_nnz, t->Format, t->blknnz, t->blknnznnz,
                  t->csr_offset, t->csrptr_offset, t->Blockcsr_Val, t->Blockcsr_Val_Low,
                  t->Blockcsr_Ptr, t->Tile_csr_Col, t->csr_compressedIdx};
    for (void *p : ps) std::free(p);
    std::memset(t, 0, sizeof(*t));
}

int mf_matrix_encode_cols(mf_matrix m, unsigned char *packed)
{
    clear_error();
    if (!m || !packed) return fail(MF_E_ARG, "null");
    const size_t nb = ((size_t)m->csrsize + 1) / 2;
    if (!nb) return MF_OK;
    uint8_t *d = nullptr;
    MF_HIP(hipMalloc((void **)&d, nb));
    int rc = encode_cols_device(m, d);
    if (!rc) {
        if (hipMemcpy(packed, d, nb, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(MF_E_HIP, "copy");
    }
    (void)hipFree(d);
    return rc;
}

int mf_matrix_classes(mf_matrix m, double tau, int allow_mask, unsigned char *codes)
{
    clear_error();
    if (!m || !codes) return fail(MF_E_ARG, "null");
    if (!m->tilenum) return MF_OK;
    uint8_t *d = nullptr;
    MF_HIP(hipMalloc((void **)&d, m->tilenum));
    int rc = classes_device(m, tau, allow_mask, d);
    if (!rc && hipMemcpy(codes, d, m->tilenum, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(MF_E_HIP, "copy");
    (void)hipFree(d);
    return rc;
}

int mf_matrix_layout(mf_matrix m, double tau, int allow_mask)
{
    clear_error();
    if (!m) return fail(MF_E_ARG, "null");
    int rc = MF_OK;
    if (!m->L64.built) rc = build_layout(m, m->L64, 0.0, 0);
    if (rc) return rc;
    return build_layout(m, m->Lmix, tau, allow_mask);
}

int mf_matrix_set_index_width(mf_matrix m, int width)
{
    clear_error();
    if (!m) return fail(MF_E_ARG, "null");
    if (width != 0 && width != 1 && width != 2 && width != 4) return fail(MF_E_ARG, "index width must be 0, 1, 2 or 4");
    if (width != m->want_iw) {
        m->want_iw = width;
        m->L64.release();
        m->Lmix.release();
    }
    return MF_OK;
}

int mf_matrix_get_info(mf_matrix m, mf_matrix_info *info)
{
    clear_error();
    if (!m || !info) return fail(MF_E_ARG, "null");
    std::memset(info, 0, sizeof(*info));
    info->rowA = m->rowA; info->colA = m->colA; info->nnz = m->nnz; info->B = m->B;
    info->tilem = m->tilem; info->tilen = m->tilen; info->tilenum = m->tilenum;
    if (!m->L64.built) {
        int rc = build_layout(m, m->L64, 0.0, 0);
        if (rc) return rc;
    }
    info->bytes_alg_fp64 = m->L64.bytes_alg;
    const Layout &L = m->Lmix.built ? m->Lmix : m->L64;
    for (int c = 0; c < 5; c++) {
        info->tiles_per_code[c] = L.tiles_per_code[c];
        info->nnz_per_code[c] = L.nnz_per_code[c];
    }
    info->bytes_alg_mixed = L.bytes_alg;
    info->bytes_layout_mixed = L.bytes_actual;
    info->index_width = L.iw;
    info->pattern_slabs_fp64 = m->L64.pattern_slabs;
    info->pattern_slabs_mixed = L.pattern_slabs;
    return MF_OK;
}

int mf_spmv(mf_matrix m, int mixed, const double *x, double *y)
{
    clear_error();
    if (!m || !x || !y) return fail(MF_E_ARG, "null");
    if (!m->L64.built) {
        int rc = build_layout(m, m->L64, 0.0, 0);
        if (rc) return rc;
    }
    const Layout &L = mixed ? m->Lmix : m->L64;
    if (!L.built) return fail(MF_E_ARG, "mixed layout not built (mf_matrix_layout)");
    const size_t nc = (size_t)m->tilen * m->B, nr = (size_t)m->tilem * m->B;
    double *dx = nullptr, *dy = nullptr;
    MF_HIP(hipMalloc((void **)&dx, sizeof(double) * nc));
    MF_HIP(hipMalloc((void **)&dy, sizeof(double) * nr));
    MF_HIP(hipMemset(dx, 0, sizeof(double) * nc));
    MF_HIP(hipMemcpy(dx, x, sizeof(double) * m->colA, hipMemcpyHostToDevice));
    int rc = spmv_device(m, L, dx, dy, true);
    if (!rc) {
        if (hipStreamSynchronize(stream()) != hipSuccess ||
            hipMemcpy(y, dy, sizeof(double) * nr, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(MF_E_HIP, "spmv copy");
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int mf_spmv_bench(mf_matrix m, int mixed, int reps, double *ms)
{
    clear_error();
    if (!m || !ms || reps <= 0) return fail(MF_E_ARG, "bad args");
    if (!m->L64.built) {
        int rc = build_layout(m, m->L64, 0.0, 0);
        if (rc) return rc;
    }
    const Layout &L = mixed ? m->Lmix : m->L64;
    if (!L.built) return fail(MF_E_ARG, "mixed layout not built (mf_matrix_layout)");
    const size_t nc = (size_t)m->tilen * m->B, nr = (size_t)m->tilem * m->B;
    double *dx = nullptr, *dy = nullptr;
    MF_HIP(hipMalloc((void **)&dx, sizeof(double) * nc));
    MF_HIP(hipMalloc((void **)&dy, sizeof(double) * nr));
    std::vector<double> ones(nc, 1.0);
    MF_HIP(hipMemcpy(dx, ones.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    int rc = spmv_device(m, L, dx, dy, true);  // also fills the FP32 copy
    if (rc) return rc;
    hipStream_t st = stream();
    hipEvent_t e0, e1;
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    MF_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps && !rc; i++) rc = spmv_device(m, L, dx, dy, false);
    MF_HIP(hipEventRecord(e1, st));
    MF_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    MF_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int mf_solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts *o, mf_report *rep)
{
    clear_error();
    if (!m || !d_b || !d_x || !o || !rep) return fail(MF_E_ARG, "null");
    return solve_device(m, d_b, d_x, *o, rep);
}

int mf_solve_device_noalloc(mf_matrix m, const double *d_b, double *d_x, const mf_opts *o, mf_report *rep)
{
    clear_error();
    if (!m || !d_b || !d_x || !o || !rep) return fail(MF_E_ARG, "null");
    const int rc = ensure_device();
    if (rc) return rc;
    return mf::solve_noalloc(m, g_device, d_b, d_x, *o, rep);
}

int mf_vendor_solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts *o, mf_report *rep)
{
    clear_error();
    if (!m || !d_b || !d_x || !o || !rep) return fail(MF_E_ARG, "null");
    return mf::vendor_solve_device(m, d_b, d_x, *o, rep);
}

int mf_vendor_solve(mf_matrix m, const double *b, double *x, const mf_opts *o, mf_report *rep)
{
    clear_error();
    if (!m || !b || !x || !o || !rep) return fail(MF_E_ARG, "null");
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    const size_t n = (size_t)m->rowA;
    double *db = nullptr, *dx = nullptr;
    MF_HIP(hipMalloc((void **)&db, sizeof(double) * n));
    MF_HIP(hipMalloc((void **)&dx, sizeof(double) * n));
    MF_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
    int rc = mf::vendor_solve_device(m, db, dx, *o, rep);
    if (hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess && !rc)
        rc = fail(MF_E_HIP, "copy x");
    (void)hipFree(db);
    (void)hipFree(dx);
    return rc;
}

int mf_solve(mf_matrix m, const double *b, double *x, const mf_opts *o, mf_report *rep)
{
    clear_error();
    if (!m || !b || !x || !o || !rep) return fail(MF_E_ARG, "null");
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    const size_t n = (size_t)std::max(m->tilem, m->tilen) * m->B;
    double *db = nullptr, *dx = nullptr;
    MF_HIP(hipMalloc((void **)&db, sizeof(double) * n));
    MF_HIP(hipMalloc((void **)&dx, sizeof(double) * n));
    MF_HIP(hipMemset(db, 0, sizeof(double) * n));
    MF_HIP(hipMemset(dx, 0, sizeof(double) * n));
    MF_HIP(hipMemcpy(db, b, sizeof(double) * m->rowA, hipMemcpyHostToDevice));
    int rc = solve_device(m, db, dx, *o, rep);
    if (hipMemcpy(x, dx, sizeof(double) * m->rowA, hipMemcpyDeviceToHost) != hipSuccess && !rc)
        rc = fail(MF_E_HIP, "copy x");
    (void)hipFree(db);
    (void)hipFree(dx);
    return rc;
}

double mf_mc_classify(mf_matrix m, long long counts[4])
{
    clear_error();
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    mf_ref_tiles t;
    if (mf_matrix_export(m, &t)) return -1.0;
    mf_compact c;
    if (mf_compact_rows(&t, &c)) { mf_ref_tiles_free(&t); return -1.0; }
    // MC/memory_Mille_feillue.cu:866-918: bins by |v| <= 3.9375, 60, 6000
    for (int tid = 0; tid < t.tilenum; tid++) {
        const int off = t.csr_offset[tid];
        for (int ri = c.nonzero_row_new[tid]; ri < c.nonzero_row_new[tid + 1]; ri++)
            for (int rj = c.blockcsr_ptr_new[ri]; rj < c.blockcsr_ptr_new[ri + 1]; rj++) {
                const double v = t.Blockcsr_Val[off + rj];
                const int bin = (v >= -3.9375 && v <= 3.9375) ? 0 : (v >= -60.0 && v <= 60.0) ? 1
                                : (v >= -6000.0 && v <= 6000.0) ? 2 : 3;
                counts[bin]++;
            }
    }
    std::free(c.nonzero_row_new); std::free(c.blockrowid_new); std::free(c.blockcsr_ptr_new);
    std::free(c.block_signal); std::free(c.tile_rowidx);
    mf_ref_tiles_free(&t);
    return (double)(8 * counts[3] + 2 * counts[1] + 4 * counts[2] + counts[0]) / 1048576.0;
}

// Device bytes of a compute layout: slabs + headers (trh 32 B per tile row,
// stc 4 B per tile, tile_ptr).
static double layout_bytes(const mf_matrix m, const mf::Layout &L)
{
    return (double)L.blob_bytes + 32.0 * m->tilem + 4.0 * m->tilenum + 4.0 * (m->tilem + 1);
}

int mf_abi_sizes(size_t *sizes, int n)
{
    const size_t all[] = {sizeof(mf_ref_tiles), sizeof(mf_matrix_info), sizeof(mf_opts), sizeof(mf_report),
                          sizeof(mf_batch_info_t), sizeof(mf_memory), sizeof(mf_compact), sizeof(mf_partition),
                          sizeof(Tile_matrix)};
    const int k = (int)(sizeof(all) / sizeof(all[0]));
    for (int i = 0; i < k && i < n && sizes; i++) sizes[i] = all[i];
    return k;
}

int mf_memory_report(mf_matrix m, double tau, int allow_mask, mf_memory *out)
{
    clear_error();
    if (!m || !out) return fail(MF_E_ARG, "null");
    std::memset(out, 0, sizeof(*out));
    mf_ref_tiles t;
    int rc = mf_matrix_export(m, &t);
    if (rc) return rc;
    mf_compact c;
    rc = mf_compact_rows(&t, &c);
    if (rc) { mf_ref_tiles_free(&t); return rc; }
    // MC/memory_Mille_feillue.cu:659-671 (tile metadata), :859-864 (compact
    // lists, Tile_csr_Col, row index), :866-918 (values by magnitude bin),
    // :920-922 (memory_CP: 8 B per row segment); sums in 64 bits
    const double MB = 1048576.0;
    const long long tm = t.tilem, tn = t.tilenum;
    const double a = 4.0 * (tm + 1) + 4.0 * tn + (tn + 1) + 4.0 * (tn + 1) + 4.0 * (tn + 1) + 4.0 * (tn + 3);
    const double b = 2.0 * ((double)c.cnt_non_new + 1) + tn + (double)t.csrsize + 4.0 * tn;
    long long cnt[4] = {0, 0, 0, 0};
    for (int tid = 0; tid < t.tilenum; tid++) {
        const int off = t.csr_offset[tid];
        for (int ri = c.nonzero_row_new[tid]; ri < c.nonzero_row_new[tid + 1]; ri++)
            for (int rj = c.blockcsr_ptr_new[ri]; rj < c.blockcsr_ptr_new[ri + 1]; rj++) {
                const double v = t.Blockcsr_Val[off + rj];
                const int bin = (v >= -3.9375 && v <= 3.9375) ? 0 : (v >= -60.0 && v <= 60.0) ? 1
                                : (v >= -6000.0 && v <= 6000.0) ? 2 : 3;
                cnt[bin]++;
            }
    }
    const double cv = 8.0 * cnt[3] + 2.0 * cnt[1] + 4.0 * cnt[2] + 1.0 * cnt[0];
    out->ref_memory_MB = (a + b + cv) / MB;
    out->ref_memory_CP_MB = 8.0 * (double)c.nonzero_row_new[t.tilenum] / MB;
    std::free(c.nonzero_row_new); std::free(c.blockrowid_new); std::free(c.blockcsr_ptr_new);
    std::free(c.block_signal); std::free(c.tile_rowidx);
    mf_ref_tiles_free(&t);
    out->csr_fp64_MB = (4.0 * (m->rowA + 1) + 12.0 * m->nnz) / MB;
    if (!m->L64.built) {
        rc = mf::build_layout(m, m->L64, 0.0, 0);
        if (rc) return rc;
    }
    rc = mf::build_layout(m, m->Lmix, tau, allow_mask);
    if (rc) return rc;
    out->layout_fp64_MB = layout_bytes(m, m->L64) / MB;
    out->layout_mixed_MB = layout_bytes(m, m->Lmix) / MB;
    out->index_width = m->Lmix.iw;
    for (int k = 0; k < 5; k++) out->nnz_per_code[k] = m->Lmix.nnz_per_code[k];
    return MF_OK;
}

int mf_solve_batched(int nsys, const int *n, const int *const *rowptr, const int *const *colidx,
                     const double *const *val, int nrhs, const double *const *b, double *const *x,
                     const mf_opts *o, int *iters, double *relres, mf_report *rep);

// ------------------------------------------------------------ drop-ins
// PC/csr2block.h:267
void Tile_create(Tile_matrix *matrix, int rowA, int colA, int nnzA, int *csrRowPtrA, int *csrColIdxA,
                 double *csrValA, float *csrValA_Low)
{
    std::memset(matrix, 0, sizeof(*matrix));
    mf_matrix m = nullptr;
    if (mf_matrix_create(rowA, colA, nnzA, csrRowPtrA, csrColIdxA, csrValA, csrValA_Low,
                         MF_TILE_BLOCK_SIZE, &m))
        return;
    mf_ref_tiles t;
    int rc = mf_matrix_export(m, &t);
    mf_matrix_destroy(m);
    if (rc) return;
    matrix->tilem = t.tilem;
    matrix->tilen = t.tilen;
    matrix->tilenum = t.tilenum;
    matrix->tile_ptr = t.tile_ptr;
    matrix->tile_columnidx = t.tile_columnidx;
    matrix->tile_nnz = t.tile_nnz;
    matrix->Format = t.Format;
    matrix->blknnz = t.blknnz;
    matrix->blknnznnz = t.blknnznnz;
    matrix->tilewidth = (char *)std::calloc((size_t)t.tilenum + 1, 1);
    matrix->csr_offset = t.csr_offset;
    matrix->csrptr_offset = t.csrptr_offset;
    matrix->Blockcsr_Val = t.Blockcsr_Val;
    matrix->Blockcsr_Val_Low = t.Blockcsr_Val_Low;
    matrix->Tile_csr_Col = t.Tile_csr_Col;
    matrix->Blockcsr_Ptr = t.Blockcsr_Ptr;
    matrix->csr_compressedIdx = t.csr_compressedIdx;
    matrix->csrsize = t.csrsize;
    matrix->csrptrlen = t.csrptrlen;
}

// PC/format.h:88
void Tile_destroy(Tile_matrix *matrix)
{
    void *ps[] = {matrix->tile_ptr, matrix->tile_columnidx, matrix->tile_nnz, matrix->Format,
                  matrix->blknnz, matrix->blknnznnz, matrix->tilewidth, matrix->csr_offset,
                  matrix->csrptr_offset, matrix->Blockcsr_Val, matrix->Blockcsr_Val_Low,
                  matrix->Tile_csr_Col, matrix->Blockcsr_Ptr, matrix->csr_compressedIdx};
    for (void *p : ps) std::free(p);
    std::memset(matrix, 0, sizeof(*matrix));
}

}  // extern "C"

// Shared body of the two solver drop-ins: truncate rows to a multiple of 16
// (PC/Mille-feuille_CG_AMD.cu:2122-2124), keep the leading square block,
// build tiles on the GPU, compute the reference partition for the CSV, solve.
struct DropinResult {
    int rc = MF_OK, iterations = 0, index = 0, each_block_nnz = 0;
    double ms_solve = 0, ms_format = 0, relres = 1, rnorm = 0;
};

static DropinResult dropin_solve(int method, const int *RowPtr, const int *ColIdx, const double *Val,
                                 const double *b, double *x, int n_rows, int nnzR, int block_nnz,
                                 int maxit)
{
    DropinResult R;
    const int B = MF_TILE_BLOCK_SIZE;
    const int rowA = (n_rows / B) * B;
    if (rowA <= 0) { R.rc = fail(MF_E_ARG, "fewer than 16 rows"); return R; }
    // leading rowA x rowA block (columns beyond the truncated rows dropped)
    std::vector<int> rp(rowA + 1, 0), ci;
    std::vector<double> vv;
    ci.reserve(RowPtr[rowA]);
    vv.reserve(RowPtr[rowA]);
    for (int i = 0; i < rowA; i++) {
        for (int k = RowPtr[i]; k < RowPtr[i + 1]; k++)
            if (ColIdx[k] < rowA) { ci.push_back(ColIdx[k]); vv.push_back(Val[k]); }
        rp[i + 1] = (int)ci.size();
    }
    double t0 = now_ms();
    mf_matrix m = nullptr;
    R.rc = mf_matrix_create(rowA, rowA, (int)ci.size(), rp.data(), ci.data(), vv.data(), nullptr, B, &m);
    if (R.rc) return R;
    mf_ref_tiles t;
    R.rc = mf_matrix_export(m, &t);
    if (!R.rc) {
        mf_partition p;
        const char *env = std::getenv("MF_SM_NUM");
        const int sm_num = env ? std::atoi(env) : device_cu_count() * 32;
        if (!mf_partition_build(&t, nnzR, block_nnz, sm_num, 1, &p)) {
            R.index = p.index;
            R.each_block_nnz = p.each_block_nnz;
            std::free(p.balance_tile_ptr); std::free(p.index_each_block); std::free(p.row_each_block);
            std::free(p.non_each_block_offset); std::free(p.balance_tile_ptr_shared_end);
        }
        mf_ref_tiles_free(&t);
    }
    R.ms_format = now_ms() - t0;
    mf_opts o;
    mf_opts_default(&o);
    o.method = method;
    o.maxit = std::max(maxit, 10000);
    mf_report rep;
    std::vector<double> xb(rowA, 0.0);
    if (!R.rc) R.rc = mf_solve(m, b, xb.data(), &o, &rep);
    mf_matrix_destroy(m);
    for (int i = 0; i < n_rows; i++) x[i] = i < rowA ? xb[i] : 0.0;
    if (!R.rc) {
        R.iterations = rep.iterations;
        R.ms_solve = rep.ms_solve;
        R.relres = rep.relres_true;
        double bn = 0;
        for (int i = 0; i < rowA; i++) bn += b[i] * b[i];
        R.rnorm = rep.relres_true * std::sqrt(bn);
        if (rep.status != MF_OK) R.rc = rep.status;
    }
    return R;
}

extern "C" {

// PC/Mille-feuille_CG_AMD.cu:1240
void cg_solve_inc(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b, int n,
                  int *iter, int maxiter, double threshold, char *filename, int nnzR, int ori,
                  int block_nnz)
{
    (void)Val_Low;
    (void)threshold;
    (void)ori;
    std::memset(x, 0, sizeof(double) * n);
    DropinResult R = dropin_solve(MF_CG, RowPtr, ColIdx, Val, b, x, n, nnzR, block_nnz, maxiter);
    if (iter) *iter = R.iterations;
    if (R.rc && R.rc != MF_E_MAXIT) {
        std::fprintf(stderr, "cg_solve_inc: %s (%s)\n", mf_status_str(R.rc), mf_last_error());
        return;
    }
    csv_append_cg("cg_syncfree_amd_mi200.csv", filename, R.iterations, R.ms_solve, nnzR, R.relres,
                  R.ms_format, R.index, R.each_block_nnz, R.rnorm);
}

// PC/Mille-feuille_BiCGSTAB_AMD.cu:837
int bicgstab(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename,
             int n, int block_nnz)
{
    // rhs = A * 1 (:1343-1350)
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = RowPtr[i]; j < RowPtr[i + 1]; j++) s += Val[j];
        rhs[i] = s;
    }
    std::memset(x, 0, sizeof(double) * n);
    DropinResult R = dropin_solve(MF_BICGSTAB, RowPtr, ColIdx, Val, rhs, x, n, nnzR, block_nnz, 10000);
    if (R.rc && R.rc != MF_E_MAXIT) {
        std::fprintf(stderr, "bicgstab: %s (%s)\n", mf_status_str(R.rc), mf_last_error());
        return 1;
    }
    csv_append_bicg("bicg_syncfree_amd_mi200.csv", filename, R.iterations, R.ms_solve, nnzR, R.relres,
                    R.relres, R.each_block_nnz);
    return R.rc == MF_OK ? 0 : 1;
}

}  // extern "C"
