// Tiled SpMV and the persistent, single-launch CG / BiCGSTAB kernels.
//
// Persistent kernels: one workgroup per CU (cooperative launch, so every
// workgroup is resident).  Workgroup g owns a contiguous chunk of tile rows
// (XCD-aware chunk order); thread (group, lane) owns row lane of tile rows
// chunk_start + i * groups + group for i < RPT, and keeps every Krylov vector
// of those rows in registers or LDS for the whole solve.  Only the vector that
// the SpMV gathers is published to global memory each iteration.  Reductions
// are deterministic: every workgroup sums all workgroups' partials in the same
// fixed order (Reducer), so all take identical control decisions.
//
// CG is Hestenes-Stiefel CG with s = A p carried by recurrence (one SpMV per
// iteration, w = A r, s = w + beta s; (p, s) is reduced after it and (r, r)
// together with the publication of r: two grid syncs per iteration).  The
// reference's recurrences (PC/Mille-feuille_CG_AMD.cu:607-912: q = A d,
// alpha = snew/(d.q), x += alpha d, r -= alpha q, beta = snew/sold,
// d = r + beta d) are the same method in exact arithmetic.
// BiCGSTAB follows PC/Mille-feuille_BiCGSTAB_AMD.cu:61-425 (van der Vorst),
// with the omega reduction of :735 computed as intended; its publications of
// p and s wait only for the workgroups whose rows the next SpMV gathers.
//
// Both kernels stop on the true residual: when the updated residual reaches
// the tolerance (or has dropped by rr_factor since the last replacement) they
// publish x, compute b - A x with FP64 accumulation over the stored values,
// and either finish or replace r by it (reliable update) and continue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mf_matrix.h"
#include "persistent.h"

namespace mf {

// ------------------------------------------------------------ SpMV kernel
// 256 threads = 256/B lane groups, lane = row inside the tile row.  A block
// covers NR * 256/B consecutive tile rows; lane group q handles tile rows
// base + k * 256/B + q (k < NR) together.  NR = 2 for layouts with narrow
// tiles and one- or two-byte indices, 1 otherwise (measured fastest at 128^3
// Poisson and the 96^3 27-point stencil).  With 16-row tiles the kernel is
// built per class mask of the layout (CM; see layout_cm), as the solvers are.
constexpr int SPMV_TPB = 256;  // 128, 512 and 1024 measured within 1 % (or slower)
template <int B, int IW, int NR, int CM = CM_ALL>
__global__ __launch_bounds__(SPMV_TPB) void k_spmv(DevL L, int tilem, const double *__restrict__ g64,
                                                   const float *__restrict__ g32, double *__restrict__ y)
{
    constexpr int GPB = SPMV_TPB / B, TRB = GPB * NR;
    __shared__ int4 s_h[2 * TRB];
    __shared__ int s_t0[TRB + 1];
    const int tb = xcd_chunk(blockIdx.x, gridDim.x) * TRB;
    for (int q = threadIdx.x; q <= TRB; q += SPMV_TPB) {
        const int tr = tb + q;
        if (q < TRB) {
            const int4 *h = reinterpret_cast<const int4 *>(L.trh + (size_t)min(tr, tilem) * 4);
            s_h[2 * q] = tr < tilem ? h[0] : make_int4(0, 0, 0, 0);
            s_h[2 * q + 1] = tr < tilem ? h[1] : make_int4(0, 0, 0, 0);
        }
        s_t0[q] = L.tile_ptr[min(tr, tilem)];
    }
    __syncthreads();
    const int grp = threadIdx.x / B, r = threadIdx.x % B;
    const int lanebase = (threadIdx.x & 63) & ~(B - 1);
    double v[NR];
    tile_rows_spmv<B, IW, NR, GPB, 2, 4, CM>(L, s_h, s_t0, grp, tb + grp, r, lanebase, g64, g32, v);
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int tr = tb + k * GPB + grp;
        if (tr < tilem) y[(size_t)tr * B + r] = v[k];
    }
}

__global__ void k_to_f32(const double *__restrict__ x, float *__restrict__ y, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

// r = b - A x (FP64 accumulation over stored values); per-block partial ||r||^2
template <int B, int IW>
__global__ __launch_bounds__(256) void k_residual(DevL L, int tilem, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *__restrict__ r,
                                                  double *__restrict__ part)
{
    __shared__ double s_w[4];
    const int tr = blockIdx.x * (256 / B) + threadIdx.x / B, lane = threadIdx.x % B;
    const int lanebase = (threadIdx.x & 63) & ~(B - 1);
    double v = 0.0;
    if (tr < tilem) {
        const double *g[1] = {x};
        double acc[1] = {0.0};
        tile_row_spmv_f64<B, IW, 1>(L, tr, lane, lanebase, g, acc);
        const size_t row = (size_t)tr * B + lane;
        const double rv = b[row] - acc[0];
        r[row] = rv;
        v = rv * rv;
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((s_w[0] + s_w[1]) + s_w[2]) + s_w[3];
}

__global__ void k_axpy(double *__restrict__ x, const double *__restrict__ d, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// Class mask of a layout's tiles (CM_F64, CM_E4M3, CM_LOW or CM_ALL; codes
// 2, 3 and 4 are E4M3, FP16 and FP32 tiles).
static int layout_cm(const Layout &L)
{
    const bool f64 = L.tiles_per_code[0] > 0;
    const bool low = L.tiles_per_code[2] + L.tiles_per_code[3] + L.tiles_per_code[4] > 0;
    const bool e4m3_only = !f64 && L.tiles_per_code[3] + L.tiles_per_code[4] == 0;
    return !low ? CM_F64 : e4m3_only ? CM_E4M3 : !f64 ? CM_LOW : CM_ALL;
}

static KernelChoice pick(int B, int iw, int method, long long rows_per_wg, bool stream, int cm = CM_ALL, int G = 0)
{
    if (B == 16 && !stream && cm != CM_F64) {
        const KernelChoice k = pick_b16_acc64(iw, method, rows_per_wg, G);
        if (k.fn) return k;
    }
    if (B == 16 && !stream && cm != CM_ALL) {
        const KernelChoice k = cm == CM_F64    ? pick_b16_f64(iw, method, rows_per_wg)
                               : cm == CM_E4M3 ? pick_b16_e4m3(iw, method, rows_per_wg)
                                               : pick_b16_low(iw, method, rows_per_wg);
        if (k.fn) return k;
    }
    switch (B) {
    case 16: return pick_tile<16>(iw, method, rows_per_wg, stream);
    case 8: return pick_b8(iw, method, rows_per_wg, stream);
    case 32: return pick_b32(iw, method, rows_per_wg, stream);
    default: return KernelChoice{};
    }
}

void Workspace::release()
{
    void *ps[] = {g64, h64, q64, g32, slots, ctl, out, r_outer, d, tmp, red, dep_ptr, dep_idx, flags, vec};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    *this = Workspace();
}

static int ensure_ws(mf_matrix m, int G)
{
    Workspace &w = m->ws;
    const size_t ncol = (size_t)std::max(m->tilem, m->tilen) * m->B;
    if (w.ncol >= ncol && w.G >= G && w.g64) return MF_OK;
    w.release();
    w.ncol = ncol;
    w.G = std::max(G, 1);
    MF_HIP(hipMalloc((void **)&w.g64, sizeof(double) * ncol));
    MF_HIP(hipMalloc((void **)&w.h64, sizeof(double) * ncol));
    MF_HIP(hipMalloc((void **)&w.q64, sizeof(double) * ncol));
    MF_HIP(hipMalloc((void **)&w.g32, sizeof(float) * ncol));
    MF_HIP(hipMalloc((void **)&w.slots, sizeof(unsigned long long) * SLOT_BUFS * SLOT_WORDS * w.G));
    MF_HIP(hipMalloc((void **)&w.ctl, sizeof(int) * 8));
    MF_HIP(hipMalloc((void **)&w.out, sizeof(double) * 4));
    MF_HIP(hipMalloc((void **)&w.r_outer, sizeof(double) * ncol));
    MF_HIP(hipMalloc((void **)&w.d, sizeof(double) * ncol));
    MF_HIP(hipMalloc((void **)&w.tmp, sizeof(double) * ncol));
    MF_HIP(hipMalloc((void **)&w.red, sizeof(double) * (ncol / 16 + 1024)));
    // gather vectors are read up to tilen*B; entries beyond the rows stay zero
    MF_HIP(hipMemset(w.g64, 0, sizeof(double) * ncol));
    MF_HIP(hipMemset(w.g32, 0, sizeof(float) * ncol));
    MF_HIP(hipMemset(w.h64, 0, sizeof(double) * ncol));
    MF_HIP(hipMemset(w.q64, 0, sizeof(double) * ncol));
    return MF_OK;
}

// Krylov vectors of the streamed kernels (nvec vectors of ncol rows), kept
// with the workspace once allocated.
static int ensure_vec(mf_matrix m, int nvec)
{
    Workspace &w = m->ws;
    const size_t need = (size_t)nvec * w.ncol;
    if (w.vec && w.vec_n >= need) return MF_OK;
    if (w.vec) MF_HIP(hipFree(w.vec));
    w.vec = nullptr;
    w.vec_n = 0;
    MF_HIP(hipMalloc((void **)&w.vec, sizeof(double) * need));
    w.vec_n = need;
    return MF_OK;
}

struct InnerResult {
    int status = MF_OK, iterations = 0, replacements = 0;
    double bnorm = 0, rt2 = 0, g2 = 0;
    double phase_us[5] = {0, 0, 0, 0, 0};  // CG, block 0: per iteration x4, passes total
    int sync_deps = -1;
    int streamed = 0;
};

// Empties the sync words, clears the chunk epoch words (epoch 0 never occurs
// in a launch) and the control words before every persistent launch (a fill
// kernel, so the launch sequence can be captured in a graph).
__global__ void k_reset_ctl(unsigned long long *slots, int nslots, unsigned *flags, int nflags, int *ctl)
{
    for (int i = threadIdx.x; i < nslots; i += blockDim.x) slots[i] = SLOT_EMPTY;
    for (int i = threadIdx.x; i < nflags; i += blockDim.x) flags[i] = 0u;
    if (threadIdx.x < 8) ctl[threadIdx.x] = 0;
}

// bits[c * G + d] = 1 when a tile of chunk c (tile rows [c trw, (c + 1) trw))
// lies in a tile column of chunk d, i.e. chunk c's SpMV gathers rows of d
__global__ void k_chunk_deps(const int *__restrict__ tile_ptr, const int *__restrict__ tile_col, int tilem, int trw,
                             int G, uint8_t *__restrict__ bits)
{
    const int tr = blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= tilem) return;
    const size_t c = (size_t)(tr / trw);
    for (int t = tile_ptr[tr]; t < tile_ptr[tr + 1]; t++) bits[c * G + min(tile_col[t] / trw, G - 1)] = 1;
}

// The neighbour lists of the grid (G chunks of trw tile rows), built once per
// grid shape: the solvers wait only for these chunks after a publish when no
// list is longer than one wave (64).
static int ensure_deps(mf_matrix m, int G, int trw)
{
    Workspace &w = m->ws;
    if (w.flags && w.dep_G == G && w.dep_trw == trw) return MF_OK;
    hipStream_t st = stream();
    for (void *p : {(void *)w.dep_ptr, (void *)w.dep_idx, (void *)w.flags})
        if (p) MF_HIP(hipFree(p));
    w.dep_ptr = w.dep_idx = nullptr;
    w.flags = nullptr;
    w.dep_G = w.dep_trw = 0;
    w.dep_max = -1;
    const size_t GG = (size_t)G * G;
    uint8_t *d_bits = nullptr;
    MF_HIP(hipMalloc((void **)&d_bits, GG));
    MF_HIP(hipMemsetAsync(d_bits, 0, GG, st));
    k_chunk_deps<<<(unsigned)((m->tilem + 255) / 256), 256, 0, st>>>(m->d_tile_ptr, m->d_tile_columnidx, m->tilem, trw,
                                                                        G, d_bits);
    MF_HIP(hipGetLastError());
    std::vector<uint8_t> bits(GG);
    MF_HIP(hipMemcpyAsync(bits.data(), d_bits, GG, hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    MF_HIP(hipFree(d_bits));
    std::vector<int> ptr(G + 1, 0), idx;
    int mx = 0;
    for (int c = 0; c < G; c++) {
        for (int d = 0; d < G; d++)
            if (d != c && bits[(size_t)c * G + d]) idx.push_back(d);
        ptr[c + 1] = (int)idx.size();
        mx = std::max(mx, ptr[c + 1] - ptr[c]);
    }
    if (idx.empty()) idx.push_back(0);
    MF_HIP(hipMalloc((void **)&w.dep_ptr, sizeof(int) * (G + 1)));
    MF_HIP(hipMalloc((void **)&w.dep_idx, sizeof(int) * idx.size()));
    MF_HIP(hipMalloc((void **)&w.flags, sizeof(unsigned) * 16 * (size_t)G));
    MF_HIP(hipMemcpyAsync(w.dep_ptr, ptr.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice, st));
    MF_HIP(hipMemcpyAsync(w.dep_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, st));
    MF_HIP(hipStreamSynchronize(st));
    w.dep_G = G;
    w.dep_trw = trw;
    w.dep_max = mx;
    return MF_OK;
}

// One persistent launch: solves L x = b (x overwritten) on the stream.
// Grid of the persistent solvers: workgroups and tile rows per workgroup.
// grid > 0 is taken as given; by default one workgroup per CU, but no more
// than leave every lane group (B lanes of a 512-thread workgroup) two tile
// rows, since every grid sync polls all workgroups -- except that a system
// the rule puts on 2 .. CUs/2 - 1 workgroups gets one tile row per lane group
// on up to half the CUs.
static void solver_grid(int tilem, int B, int grid, int *G_out, int *trw_out)
{
    const int cus = device_cu_count();
    int G = grid > 0 ? grid : std::min(cus, std::max(1, tilem / (2 * (512 / B))));
    // small systems that still span several workgroups: one tile row per
    // lane group on up to half the CUs (256^2 Laplacian: 3.8 ms on 128
    // workgroups against 4.1 on 64 and 5.0 on 256); one workgroup stays one
    if (grid <= 0 && G > 1 && G < cus / 2) G = std::min(cus / 2, tilem / (512 / B));
    G = std::max(1, std::min(G, tilem));
    const int trw = (tilem + G - 1) / G;
    *G_out = (tilem + trw - 1) / trw;  // no empty workgroups
    *trw_out = trw;
}

// diagnostic phase clock words of the persistent launches (opts.profile)
static unsigned long long *d_phase = nullptr;
void release_phase_buffer()
{
    if (d_phase) (void)hipFree(d_phase);
    d_phase = nullptr;
}

// Device buffers of one persistent launch: the published vectors, the sync
// words and control words, the neighbour lists (dep_max < 0: none, every
// publish is a grid barrier) and the streamed kernels' vectors.
struct WsView {
    double *g64 = nullptr, *h64 = nullptr, *q64 = nullptr;
    float *g32 = nullptr;
    unsigned long long *slots = nullptr;
    int *ctl = nullptr;
    double *out = nullptr;
    const int *dep_ptr = nullptr, *dep_idx = nullptr;
    unsigned *flags = nullptr;
    int dep_max = -1;
    double *vec = nullptr;
    long long vstride = 0;
    bool allow_profile = true;  // the phase clock allocates its buffer on first use
};

// One persistent launch on the grid (G workgroups of trw tile rows) with the
// kernel kc over the buffers w; allocates nothing unless w.allow_profile and
// a phase profile is asked for.
static int launch_persistent(mf_matrix m, const Layout &L, const Layout *L64, int method, const double *d_b,
                             double *d_x, double tol, const mf_opts &o, int G, int trw, const KernelChoice &kc,
                             const WsView &w, bool sync, InnerResult *res, int *grid_out, int *block_out,
                             int *rpt_out)
{
    hipStream_t st = stream();
    SolveArgs A;
    A.L = L.view(m->d_tile_ptr);
    A.blob_bytes = L.blob_bytes;
    A.use64 = L64 && L64 != &L ? 1 : 0;
    A.L64 = A.use64 ? L64->view(m->d_tile_ptr) : A.L;
    if (A.use64 && L64->iw != L.iw) return fail(MF_E_FORMAT, "layouts with different index forms");
    A.tilem = m->tilem;
    A.G = G;
    A.trw = trw;
    A.b = d_b;
    A.x = d_x;
    A.g64 = w.g64;
    A.g32 = w.g32;
    A.h64 = w.h64;
    A.q64 = w.q64;
    A.slots = w.slots;
    A.ctl = w.ctl;
    A.out = w.out;
    A.tol = tol;
    A.rr_factor = o.rr_factor;
    A.maxit = o.maxit;
    // kernels that sum low-precision tiles in FP64 gather the FP64 vector only
    A.has64 = L.has64 || kc.acc64 ? 1 : 0;
    A.haslow = L.haslow && !kc.acc64 ? 1 : 0;
    A.spin_limit = o.spin_limit;
    A.dep_ptr = w.dep_ptr;
    A.dep_idx = w.dep_idx;
    A.flags = w.flags;
    // BiCGSTAB's syncs after publishing p and s carry no reduction: they can
    // be neighbour waits (CG's publish sync also reduces (r, r))
    A.nbr = method == MF_BICGSTAB && o.sync != MF_SYNC_GLOBAL && w.flags && w.dep_max >= 0 && w.dep_max <= 64 ? 1 : 0;
    A.vec = kc.streamed ? w.vec : nullptr;
    A.vstride = w.vstride;
    if (kc.streamed && !w.vec) return fail(MF_E_ARG, "streamed kernel without its vectors");
    if (res) {
        res->sync_deps = A.nbr ? w.dep_max : -1;
        res->streamed = kc.streamed;
    }
    static const bool phase_profile = std::getenv("MF_PHASE_PROFILE") != nullptr;
    const size_t phase_words = 16 + 8 * 1024;
    const bool prof = w.allow_profile && (phase_profile || o.profile);
    if (prof && !d_phase) MF_HIP(hipMalloc((void **)&d_phase, phase_words * sizeof(unsigned long long)));
    A.phase = prof ? d_phase : nullptr;
    if (A.phase) MF_HIP(hipMemsetAsync(d_phase, 0, phase_words * sizeof(unsigned long long), st));
    k_reset_ctl<<<1, 1024, 0, st>>>(w.slots, SLOT_BUFS * SLOT_WORDS * G, w.flags, w.flags ? 16 * G : 0, w.ctl);
    MF_HIP(hipGetLastError());
    void *args[] = {&A};
    // Cooperative launch guarantees that all G workgroups are resident.  Some
    // runtimes refuse it (e.g. while another process holds the device); then
    // the occupancy API must show G resident workgroups and a plain launch is
    // used.  Either way the barrier waits are bounded and the abort word is
    // sticky, so a workgroup that could not start in time finds the abort
    // raised and every workgroup exits.
    static const bool no_coop = std::getenv("MF_NO_COOP") != nullptr;
    hipError_t le = no_coop ? hipErrorNotSupported
                            : hipLaunchCooperativeKernel((const void *)kc.fn, dim3(G), dim3(kc.tpb), args, 0, st);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        int per_cu = 0;
        MF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kc.fn, kc.tpb, 0));
        if (per_cu * device_cu_count() < G)
            return fail(MF_E_UNSUPPORTED, "grid does not fit the device for a persistent launch");
        MF_HIP(hipLaunchKernel((const void *)kc.fn, dim3(G), dim3(kc.tpb), args, 0, st));
    }
    if (grid_out) *grid_out = G;
    if (block_out) *block_out = kc.tpb;
    if (rpt_out) *rpt_out = kc.rpt;
    if (sync && res) {
        int ctl[8];
        double out[4];
        MF_HIP(hipMemcpyAsync(ctl, w.ctl, sizeof(ctl), hipMemcpyDeviceToHost, st));
        MF_HIP(hipMemcpyAsync(out, w.out, sizeof(out), hipMemcpyDeviceToHost, st));
        MF_HIP(hipStreamSynchronize(st));
        res->status = ctl[0] ? MF_E_TIMEOUT : ctl[1];
        res->iterations = ctl[2];
        res->replacements = ctl[3];
        res->bnorm = out[0];
        res->rt2 = out[1];
        res->g2 = out[2];
        if (A.phase) {
            unsigned long long ph[8];
            MF_HIP(hipMemcpy(ph, A.phase, sizeof(ph), hipMemcpyDeviceToHost));
            const double it = std::max(1, res->iterations);
            for (int k = 0; k < 4; k++) res->phase_us[k] = ph[k] / it / 100.0;
            res->phase_us[4] = ph[4] / 100.0;
        }
        if (A.phase && phase_profile) {
            unsigned long long ph[8];
            MF_HIP(hipMemcpy(ph, A.phase, sizeof(ph), hipMemcpyDeviceToHost));
            const double it = std::max(1, res->iterations);
            std::fprintf(stderr,
                         "[mf phase] G=%d TPB=%d RPT=%d it=%d  us/iter: spmv %.2f  reduce1 %.2f  update %.2f  "
                         "reduce2 %.2f  | true-residual passes %.1f us total\n",
                         G, kc.tpb, kc.rpt, res->iterations, ph[0] / it / 100.0, ph[1] / it / 100.0,
                         ph[2] / it / 100.0, ph[3] / it / 100.0, ph[4] / 100.0);
            std::vector<unsigned long long> sv((size_t)8 * G);
            MF_HIP(hipMemcpy(sv.data(), A.phase + 16, sv.size() * 8, hipMemcpyDeviceToHost));
            for (int e = 0; e < 4; e++) {
                const unsigned long long *p = sv.data() + (size_t)e * 2 * G;
                if (!p[0]) continue;
                unsigned long long a0 = ~0ull, a1 = 0, x0 = ~0ull, x1 = 0;
                for (int g = 0; g < G; g++) {
                    a0 = std::min(a0, p[2 * g]);
                    a1 = std::max(a1, p[2 * g]);
                    x0 = std::min(x0, p[2 * g + 1]);
                    x1 = std::max(x1, p[2 * g + 1]);
                }
                std::fprintf(stderr, "[mf phase] barrier %u: arrival spread %.2f us, last arrival -> first exit %.2f us, "
                             "exit spread %.2f us\n", PHASE_EP0 + e, (a1 - a0) / 100.0, ((double)x0 - (double)a1) / 100.0,
                             (x1 - x0) / 100.0);
            }
            // MF_PHASE_PROFILE=2: per workgroup, the time from leaving one
            // stamped sync to arriving at the next (its own work between them)
            if (std::atoi(std::getenv("MF_PHASE_PROFILE")) == 2) {
                for (int e = 1; e < 4; e++) {
                    const unsigned long long *p = sv.data() + (size_t)e * 2 * G, *q = p - 2 * G;
                    if (!p[0] || !q[0]) continue;
                    for (int g = 0; g < G; g++) {
                        const int per = G >> 3, rem = G & 7, xcd = g & 7, k = g >> 3;
                        const int chunk = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + k;
                        std::fprintf(stderr, "[mf phase wg] sync %u wg %d xcd %d chunk %d work %.2f us\n",
                                     PHASE_EP0 + e, g, xcd, chunk, ((double)p[2 * g] - (double)q[2 * g + 1]) / 100.0);
                    }
                }
            }
        }
    }
    return MF_OK;
}

// One persistent launch with the matrix's own workspace (allocated, or
// grown, on first use for the grid and kernel).
static int launch_inner(mf_matrix m, const Layout &L, const Layout *L64, int method, const double *d_b,
                        double *d_x, double tol, const mf_opts &o, bool sync, InnerResult *res, int *grid_out,
                        int *block_out, int *rpt_out)
{
    int G = 1, trw = 1;
    solver_grid(m->tilem, m->B, o.grid, &G, &trw);
    KernelChoice kc = pick(m->B, L.iw, method, (long long)trw * m->B, false, layout_cm(L), G);
    if (!kc.fn) return fail(MF_E_UNSUPPORTED, "no persistent kernel for this tile size and index form");
    // streaming the vectors of a system whose rows fit on chip is switched
    // off: two runs that forced it on small grids ended with a GPU fault
    // whose cause is not known
    if (o.vectors == MF_VECTORS_HBM && !kc.streamed)
        return fail(MF_E_UNSUPPORTED, "vectors in HBM only for systems beyond register/LDS residency");
    if (o.vectors == MF_VECTORS_ONCHIP && kc.streamed)
        return fail(MF_E_UNSUPPORTED, "vectors on chip: the system exceeds register/LDS residency at this grid");
    int rc = ensure_ws(m, G);
    if (rc) return rc;
    if (kc.streamed) {
        rc = ensure_vec(m, stream_vectors(method));
        if (rc) return rc;
    }
    rc = ensure_deps(m, G, trw);
    if (rc) return rc;
    const Workspace &ws = m->ws;
    WsView w;
    w.g64 = ws.g64;
    w.h64 = ws.h64;
    w.q64 = ws.q64;
    w.g32 = ws.g32;
    w.slots = ws.slots;
    w.ctl = ws.ctl;
    w.out = ws.out;
    w.dep_ptr = ws.dep_ptr;
    w.dep_idx = ws.dep_idx;
    w.flags = ws.flags;
    w.dep_max = ws.dep_max;
    w.vec = ws.vec;
    w.vstride = (long long)ws.ncol;
    return launch_persistent(m, L, L64, method, d_b, d_x, tol, o, G, trw, kc, w, sync, res, grid_out, block_out,
                             rpt_out);
}

// r = b - A x over layout L; returns ||r|| (FP64) via host sum of partials.
static int residual(mf_matrix m, const Layout &L, const double *d_b, const double *d_x, double *d_r,
                    double *norm)
{
    hipStream_t st = stream();
    const int B = m->B;
    const int nb = (m->tilem * B + 255) / 256;
    DevL V = L.view(m->d_tile_ptr);
#define MF_RES(BB, W) k_residual<BB, W><<<nb, 256, 0, st>>>(V, m->tilem, d_b, d_x, d_r, m->ws.red)
#define MF_RES_B(BB) if (L.iw == 1) MF_RES(BB, 1); else if (L.iw == 2) MF_RES(BB, 2); else MF_RES(BB, 4)
    if (B == 16) { MF_RES_B(16); }
    else if (B == 8) { MF_RES_B(8); }
    else { MF_RES_B(32); }
#undef MF_RES_B
#undef MF_RES
    MF_HIP(hipGetLastError());
    std::vector<double> part(nb);
    MF_HIP(hipMemcpyAsync(part.data(), m->ws.red, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : part) s += v;
    *norm = std::sqrt(s);
    return MF_OK;
}

static int solve_direct(mf_matrix m, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep)
{
    hipStream_t st = stream();
    std::memset(rep, 0, sizeof(*rep));
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    if (o.index_width < 0 || o.index_width == 3 || o.index_width > 4)
        return fail(MF_E_ARG, "index width must be 0, 1, 2 or 4");
    if (o.vectors < MF_VECTORS_AUTO || o.vectors > MF_VECTORS_ONCHIP) return fail(MF_E_ARG, "vectors");
    double t0 = now_ms();
    int rc;
    const int want = o.index_width ? o.index_width : m->want_iw;
    if (!m->L64.built || m->L64.want_iw != want) {
        rc = build_layout(m, m->L64, 0.0, 0, want);
        if (rc) return rc;
    }
    const bool mixed = o.precision == MF_MIXED;
    if (mixed && (!m->Lmix.built || m->Lmix.tau != o.tau || m->Lmix.allow_mask != o.allow_mask ||
                  m->Lmix.want_iw != want)) {
        rc = build_layout(m, m->Lmix, o.tau, o.allow_mask, want);
        if (rc) return rc;
    }
    rep->ms_setup = now_ms() - t0;
    const Layout &L = mixed ? m->Lmix : m->L64;
    // lossless layout: the stored operator is A itself, so the kernel's own
    // true-residual test is final; a lossy layout either runs its passes on
    // the FP64 layout (ir_max_outer = 0) or is refined in FP64 around it
    const bool exact = !mixed || o.tau == 0.0;
    rc = ensure_ws(m, std::max(1, o.grid > 0 ? o.grid : device_cu_count()));
    if (rc) return rc;
    {  // the grid's neighbour lists, outside the timed region
        int G0 = 1, trw0 = 1;
        solver_grid(m->tilem, m->B, o.grid, &G0, &trw0);
        rc = ensure_deps(m, G0, trw0);
        if (rc) return rc;
    }
    Workspace &w = m->ws;
    const size_t n = (size_t)m->tilem * m->B;
    hipEvent_t e0, e1;
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    MF_HIP(hipEventRecord(e0, st));
    InnerResult ir;
    int G = 0, TPB = 0, RPT = 0;
    int status = MF_OK, iters = 0, repl = 0, outer = 0;
    double relres = -1.0;
    if (exact || o.ir_max_outer <= 0) {
        // lossless: the kernel's own passes are final; lossy without outer
        // refinement: the passes run on the FP64 layout (reliable updates
        // against A itself) inside the one launch
        rc = launch_inner(m, L, exact ? nullptr : &m->L64, o.method, d_b, d_x, o.tol, o, true, &ir, &G, &TPB, &RPT);
        if (rc) return rc;
        status = ir.status;
        iters = ir.iterations;
        repl = ir.replacements;
        outer = 1;
        if (ir.bnorm > 0 && ir.rt2 >= 0) relres = std::sqrt(ir.rt2) / ir.bnorm;
    } else {
        // x = 0; r = b; repeat: solve L d = r to a relative tolerance, x += d,
        // r = b - A x in FP64 (FP64 layout)
        MF_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
        MF_HIP(hipMemcpyAsync(w.r_outer, d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        double bnorm = 0.0, rnorm = 0.0;
        rc = residual(m, m->L64, d_b, d_x, w.r_outer, &bnorm);
        if (rc) return rc;
        rnorm = bnorm;
        status = MF_E_MAXIT;
        for (outer = 0; outer < o.ir_max_outer && rnorm > o.tol * bnorm; ++outer) {
            const double inner_tol = std::max(o.tol * bnorm / rnorm, o.ir_inner_tol);
            rc = launch_inner(m, L, nullptr, o.method, w.r_outer, w.d, inner_tol, o, true, &ir, &G, &TPB, &RPT);
            if (rc) return rc;
            iters += ir.iterations;
            repl += ir.replacements;
            if (ir.status == MF_E_TIMEOUT || ir.status == MF_E_NONFINITE || ir.status == MF_E_DIVERGED) {
                status = ir.status;
                break;
            }
            k_axpy<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(d_x, w.d, (long long)n);
            rc = residual(m, m->L64, d_b, d_x, w.r_outer, &rnorm);
            if (rc) return rc;
        }
        if (rnorm <= o.tol * bnorm) status = MF_OK;
        relres = bnorm > 0 ? rnorm / bnorm : 0.0;
    }
    // lossy tiles that diverged, broke down, produced non-finite sums or ran
    // out of iterations: the same solve on the FP64 layout
    int fallback = 0;
    if (!exact && (status == MF_E_DIVERGED || status == MF_E_BREAKDOWN || status == MF_E_NONFINITE ||
                   status == MF_E_MAXIT)) {
        rc = launch_inner(m, m->L64, nullptr, o.method, d_b, d_x, o.tol, o, true, &ir, &G, &TPB, &RPT);
        if (rc) return rc;
        status = ir.status;
        iters += ir.iterations;
        repl += ir.replacements;
        fallback = 1;
    }
    MF_HIP(hipEventRecord(e1, st));
    MF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    MF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // independent FP64 check of the returned x (outside the timed region)
    double bn = 0.0, rn = 0.0;
    MF_HIP(hipMemsetAsync(w.tmp, 0, sizeof(double) * n, st));
    rc = residual(m, m->L64, d_b, w.tmp, w.d, &bn);
    if (rc) return rc;
    rc = residual(m, m->L64, d_b, d_x, w.d, &rn);
    if (rc) return rc;
    rep->relres_true = bn > 0 ? rn / bn : rn;
    (void)relres;
    rep->status = status;
    rep->iterations = iters;
    rep->outer = outer;
    rep->replacements = repl;
    rep->ms_solve = ms;
    rep->grid = G;
    rep->block = TPB;
    rep->rows_per_thread = RPT;
    for (int k = 0; k < 5; k++) rep->phase_us[k] = ir.phase_us[k];
    rep->sync_deps = ir.sync_deps;
    rep->streamed = ir.streamed;
    rep->fallback = fallback;
    return status == MF_E_TIMEOUT ? fail(MF_E_TIMEOUT, "grid barrier timeout") : MF_OK;
}

// ------------------------------------------------ allocation-free solves
// For callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid it): the buffers of one persistent
// launch -- sync words, control words, the published vector (FP64 and FP32)
// and the true-residual pass's two publishing slots -- live in one static
// device array per GPU, held until the process ends.  It is sized for
// NA_MAX_ROWS padded rows and NA_MAX_G workgroups:
//   slots  SLOT_BUFS x SLOT_WORDS x NA_MAX_G words of 8 bytes (196,608 B)
//   ctl    8 ints, out 4 doubles (256 B each)
//   g64, h64, q64  8 B per row; g32  4 B per row
// = 58,917,376 + 197,120 bytes (about 56.4 MiB).
// Every call on a GPU uses that GPU's array and leaves its contents; the
// sync and control words are cleared by a fill kernel at the start of each
// call, and every entry of the vector buffers is written in a launch before
// it is read.  Calls that overlap on one GPU race on the array.
constexpr size_t NA_MAX_ROWS = size_t(1) << 21;
constexpr int NA_MAX_G = 1024;
constexpr size_t NA_OFF_CTL = (size_t)SLOT_BUFS * SLOT_WORDS * NA_MAX_G * 8;
constexpr size_t NA_OFF_OUT = NA_OFF_CTL + 256;
constexpr size_t NA_OFF_G64 = NA_OFF_OUT + 256;
constexpr size_t NA_OFF_H64 = NA_OFF_G64 + 8 * NA_MAX_ROWS;
constexpr size_t NA_OFF_Q64 = NA_OFF_H64 + 8 * NA_MAX_ROWS;
constexpr size_t NA_OFF_G32 = NA_OFF_Q64 + 8 * NA_MAX_ROWS;
constexpr size_t NA_BYTES = NA_OFF_G32 + 4 * NA_MAX_ROWS;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEV = 64;
static void *g_scratch[NA_MAX_DEV];          // device address of scratch, per GPU
static hipEvent_t g_na_ev[NA_MAX_DEV][2];    // timing events of the allocation-free solves

// Once per device, with that device current (the library's device set-up):
// the device address of that GPU's copy of scratch, and two timing events.
int noalloc_setup(int dev)
{
    if (dev < 0 || dev >= NA_MAX_DEV) return fail(MF_E_NODEV, "device index beyond the allocation-free table");
    if (g_scratch[dev]) return MF_OK;
    void *p = nullptr;
    MF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(MF_E_HIP, "no device address for the static solve buffers");
    for (hipEvent_t &e : g_na_ev[dev]) MF_HIP(hipEventCreate(&e));
    g_scratch[dev] = p;
    return MF_OK;
}

// At library finalisation: the events go, the array stays (the next set-up
// looks its address up again).
void noalloc_release()
{
    for (int d = 0; d < NA_MAX_DEV; d++)
        if (g_scratch[d]) {
            for (hipEvent_t &e : g_na_ev[d]) (void)hipEventDestroy(e);
            g_scratch[d] = nullptr;
        }
}

// The allocation-free solve of m in its own row order (solve_noalloc below
// picks m or its reordered copy); rep gets everything but the timing.
static int noalloc_core(mf_matrix m, int dev, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep)
{
    hipStream_t st = stream();
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    const size_t ncol = (size_t)std::max(m->tilem, m->tilen) * m->B;
    if (ncol > NA_MAX_ROWS)
        return fail(MF_E_ARG, "allocation-free solve: " + std::to_string(ncol) + " padded rows exceed the limit of " +
                                  std::to_string(NA_MAX_ROWS));
    if (o.method != MF_CG && o.method != MF_BICGSTAB) return fail(MF_E_ARG, "method");
    if (o.ir_max_outer > 0) return fail(MF_E_UNSUPPORTED, "allocation-free solve: no outer refinement loop");
    const bool mixed = o.precision == MF_MIXED;
    const int want = o.index_width ? o.index_width : m->want_iw;
    if (!m->L64.built || m->L64.want_iw != want)
        return fail(MF_E_ARG, "allocation-free solve: FP64 layout not built (mf_matrix_layout or a solve first)");
    if (mixed && (!m->Lmix.built || m->Lmix.tau != o.tau || m->Lmix.allow_mask != o.allow_mask ||
                  m->Lmix.want_iw != want))
        return fail(MF_E_ARG, "allocation-free solve: mixed layout for these tau/allow_mask not built "
                              "(mf_matrix_layout first)");
    const Layout &L = mixed ? m->Lmix : m->L64;
    const bool exact = !mixed || o.tau == 0.0;
    int G = 1, trw = 1;
    solver_grid(m->tilem, m->B, o.grid, &G, &trw);
    if (G > NA_MAX_G) return fail(MF_E_ARG, "allocation-free solve: more than 1024 workgroups");
    const KernelChoice kc = pick(m->B, L.iw, o.method, (long long)trw * m->B, false, layout_cm(L), G);
    if (!kc.fn || kc.streamed)
        return fail(MF_E_UNSUPPORTED, "allocation-free solve: system beyond register/LDS residency at this grid");
    unsigned char *base = static_cast<unsigned char *>(g_scratch[dev]);
    WsView w;
    w.slots = reinterpret_cast<unsigned long long *>(base);
    w.ctl = reinterpret_cast<int *>(base + NA_OFF_CTL);
    w.out = reinterpret_cast<double *>(base + NA_OFF_OUT);
    w.g64 = reinterpret_cast<double *>(base + NA_OFF_G64);
    w.h64 = reinterpret_cast<double *>(base + NA_OFF_H64);
    w.q64 = reinterpret_cast<double *>(base + NA_OFF_Q64);
    w.g32 = reinterpret_cast<float *>(base + NA_OFF_G32);
    w.allow_profile = false;  // no neighbour lists (grid barriers), no phase clock
    InnerResult ir;
    int TPB = 0, RPT = 0, Gl = 0;
    int ctl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double out[4] = {0.0, 0.0, 0.0, 0.0};
    MF_HIP(hipEventRecord(g_na_ev[dev][0], st));
    // one launch on Lr (true-residual passes on L64r when given); control
    // words back to the host
    auto run = [&](const Layout &Lr, const Layout *L64r, const KernelChoice &k) -> int {
        int rc = launch_persistent(m, Lr, L64r, o.method, d_b, d_x, o.tol, o, G, trw, k, w, false, &ir, &Gl, &TPB,
                                   &RPT);
        if (rc) return rc;
        MF_HIP(hipMemcpyAsync(ctl, w.ctl, sizeof(ctl), hipMemcpyDeviceToHost, st));
        MF_HIP(hipMemcpyAsync(out, w.out, sizeof(out), hipMemcpyDeviceToHost, st));
        MF_HIP(hipStreamSynchronize(st));
        return MF_OK;
    };
    int rc = run(L, exact ? nullptr : &m->L64, kc);
    if (rc) return rc;
    int status = ctl[0] ? MF_E_TIMEOUT : ctl[1], iters = ctl[2], repl = ctl[3];
    // lossy tiles that failed: the same solve on the FP64 layout, as mf_solve_device
    if (!exact && (status == MF_E_DIVERGED || status == MF_E_BREAKDOWN || status == MF_E_NONFINITE ||
                   status == MF_E_MAXIT)) {
        const KernelChoice k64 = pick(m->B, m->L64.iw, o.method, (long long)trw * m->B, false, layout_cm(m->L64), G);
        if (!k64.fn || k64.streamed) return fail(MF_E_UNSUPPORTED, "allocation-free solve: no FP64 kernel");
        rc = run(m->L64, nullptr, k64);
        if (rc) return rc;
        status = ctl[0] ? MF_E_TIMEOUT : ctl[1];
        iters += ctl[2];
        repl += ctl[3];
        rep->fallback = 1;
    }
    MF_HIP(hipEventRecord(g_na_ev[dev][1], st));
    MF_HIP(hipEventSynchronize(g_na_ev[dev][1]));
    float ms = 0.0f;
    MF_HIP(hipEventElapsedTime(&ms, g_na_ev[dev][0], g_na_ev[dev][1]));
    rep->status = status;
    rep->iterations = iters;
    rep->replacements = repl;
    rep->outer = 1;
    rep->relres_true = out[0] > 0 && out[1] >= 0 ? std::sqrt(out[1]) / out[0] : (out[0] > 0 ? -1.0 : 0.0);
    rep->grid = Gl;
    rep->block = TPB;
    rep->rows_per_thread = RPT;
    rep->sync_deps = -1;
    return rep->status == MF_E_TIMEOUT ? fail(MF_E_TIMEOUT, "grid barrier timeout") : MF_OK;
}

// SpMV y = A x (device pointers); x has tilen*B entries.
int spmv_device(mf_matrix m, const Layout &L, const double *d_x, double *d_y, bool convert)
{
    hipStream_t st = stream();
    const int B = m->B;
    const long long ncol = (long long)m->tilen * B;
    int rc = ensure_ws(m, 1);
    if (rc) return rc;
    if (L.haslow && convert)
        k_to_f32<<<(unsigned)((ncol + 255) / 256), 256, 0, st>>>(d_x, m->ws.g32, ncol);
    const int nr = L.haslow && L.iw < 4 ? 2 : 1;
    const int nb = (int)(((long long)m->tilem * B + SPMV_TPB * nr - 1) / (SPMV_TPB * nr));
    DevL V = L.view(m->d_tile_ptr);
#define MF_SPMV(BB, W, NR, C) k_spmv<BB, W, NR, C><<<nb, SPMV_TPB, 0, st>>>(V, m->tilem, d_x, m->ws.g32, d_y)
#define MF_SPMV_W(BB, NR, C) if (L.iw == 1) MF_SPMV(BB, 1, NR, C); else if (L.iw == 2) MF_SPMV(BB, 2, NR, C); else MF_SPMV(BB, 4, NR, C)
#define MF_SPMV_B(BB, C) if (nr == 2) { MF_SPMV_W(BB, 2, C); } else { MF_SPMV_W(BB, 1, C); }
    if (B == 16) {
        const int cm = layout_cm(L);
        if (cm == CM_F64) { MF_SPMV_B(16, CM_F64); }
        else if (cm == CM_E4M3) { MF_SPMV_B(16, CM_E4M3); }
        else if (cm == CM_LOW) { MF_SPMV_B(16, CM_LOW); }
        else { MF_SPMV_B(16, CM_ALL); }
    }
    else if (B == 8) { MF_SPMV_B(8, CM_ALL); }
    else { MF_SPMV_B(32, CM_ALL); }
#undef MF_SPMV_B
#undef MF_SPMV_W
#undef MF_SPMV
    MF_HIP(hipGetLastError());
    return MF_OK;
}

// ------------------------------------------------------- row reordering
// Rows of very different lengths pad every slab to the tile row's longest
// row, and equal tile-row chunks give the persistent workgroups unequal work.
// The solvers then run on a reordered copy of A: rows sorted by length
// (longest first) form the tile rows, so a tile row's rows have similar
// lengths, and the tile rows are dealt round robin to the chunks of the
// persistent grid, so every workgroup gets a share of long and short rows.
// A row longer than SPLIT_ENTRIES is split over S = 2, 4, 8 or 16 (<= B)
// lanes of its tile row, each holding a consecutive segment of it, so that
// no lane walks more than about SPLIT_ENTRIES entries (random SPD with
// Zipf row lengths: rows of ~2000 entries would otherwise be a chain of
// ~500 dependent quads on one lane); the segments' products are summed
// across the lanes (DevL::split).  The copy has n' = tilem' B rows: every
// row of A once (at the first lane of its segment), plus rows without
// entries (the other lanes of a split row, and the unfilled lanes of a tile
// row), with b = 0 and x = 0 on them throughout.  CG and BiCGSTAB on the
// copy with P b give P x; their iterates equal those on A in exact
// arithmetic.
constexpr int SPLIT_ENTRIES = 64;
constexpr int SPLIT_PAD = 2048;  // zero split entries after tilem' (tile rows a kernel may look past)

__global__ void k_permute_in(const int *__restrict__ perm, int n, const double *__restrict__ src,
                             double *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = perm[i] >= 0 ? src[perm[i]] : 0.0;
}

__global__ void k_permute_out(const int *__restrict__ perm, int n, const double *__restrict__ src,
                              double *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && perm[i] >= 0) dst[perm[i]] = src[i];
}

// lanes per row for a row of len entries (a power of two, at most B)
static int split_lanes(int len, int B)
{
    int S = 1;
    while (S < B && (len + S - 1) / S > SPLIT_ENTRIES) S <<= 1;
    return S;
}

static int build_reordered(mf_matrix m, int grid)
{
    const int n = m->rowA, B = m->B, nnz = m->nnz;
    std::vector<int> rp(n + 1), ci(std::max(nnz, 1));
    std::vector<double> va(std::max(nnz, 1));
    MF_HIP(hipMemcpy(rp.data(), m->d_rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (nnz) {
        MF_HIP(hipMemcpy(ci.data(), m->d_col, sizeof(int) * nnz, hipMemcpyDeviceToHost));
        MF_HIP(hipMemcpy(va.data(), m->d_val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    }
    auto len = [&](int r) { return rp[r + 1] - rp[r]; };
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len(a) > len(b); });
    // tile rows of the copy: consecutive rows (in length order) with the same
    // lanes per row S, B / S rows each
    struct TR { int first, count, S; };
    std::vector<TR> trs;
    for (int i = 0; i < n;) {
        const int S = split_lanes(len(order[i]), B), per = B / S;
        int c = 0;
        while (c < per && i + c < n && split_lanes(len(order[i + c]), B) == S) c++;
        trs.push_back({i, c, S});
        i += c;
    }
    const int tilem2 = std::max(1, (int)trs.size());
    if (trs.empty()) trs.push_back({0, 0, 1});
    const long long n2l = (long long)tilem2 * B;
    if (n2l > INT32_MAX / 2) return fail(MF_E_UNSUPPORTED, "reordered copy too large");
    const int n2 = (int)n2l;
    // tile rows dealt round robin to the chunks of the persistent grid
    int G = 1, trw = 1;
    solver_grid(tilem2, B, grid, &G, &trw);
    std::vector<int> start(G), cap(G), fill(G, 0), pos(tilem2);
    for (int g = 0; g < G; g++) {
        start[g] = g * trw;
        cap[g] = std::max(0, std::min(trw, tilem2 - g * trw));
    }
    for (int j = 0, g = 0; j < tilem2; j++) {
        while (fill[g] >= cap[g]) g = (g + 1) % G;
        pos[j] = start[g] + fill[g]++;
        g = (g + 1) % G;
    }
    // new row of every row of A; source row (or -1) of every row of the copy
    std::vector<int> perm(n2, -1), inv(n);
    std::vector<uint8_t> split(tilem2 + SPLIT_PAD, 0);
    bool any_split = false;
    int split_rows = 0;
    for (int j = 0; j < tilem2; j++) {
        const TR &t = trs[j];
        split[pos[j]] = (uint8_t)t.S;
        any_split = any_split || t.S > 1;
        if (t.S > 1) split_rows += t.count;
        for (int q = 0; q < t.count; q++) {
            const int nw = pos[j] * B + q * t.S;
            perm[nw] = order[t.first + q];
            inv[order[t.first + q]] = nw;
        }
    }
    // CSR of the copy: a split row's entries go in S consecutive segments to
    // its S lanes (columns renumbered)
    std::vector<int> rp2(n2 + 1, 0);
    for (int j = 0; j < tilem2; j++) {
        const TR &t = trs[j];
        for (int q = 0; q < t.count; q++) {
            const int o = order[t.first + q], L = len(o), seg = (L + t.S - 1) / t.S;
            for (int k = 0; k < t.S; k++)
                rp2[pos[j] * B + q * t.S + k + 1] = std::max(0, std::min(seg, L - k * seg));
        }
    }
    for (int i = 0; i < n2; i++) rp2[i + 1] += rp2[i];
    std::vector<int> ci2(std::max(nnz, 1));
    std::vector<double> va2(std::max(nnz, 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int j = 0; j < tilem2; j++) {
        const TR &t = trs[j];
        for (int q = 0; q < t.count; q++) {
            const int o = order[t.first + q];
            // the S segments of row o are consecutive in the copy's CSR
            for (int k = rp[o], d = rp2[pos[j] * B + q * t.S]; k < rp[o + 1]; k++, d++) {
                ci2[d] = inv[ci[k]];
                va2[d] = va[k];
            }
        }
    }
    if (m->child) mf_matrix_destroy(m->child);
    m->child = nullptr;
    mf_matrix c = nullptr;
    int rc = mf_matrix_create(n2, n2, nnz, rp2.data(), ci2.data(), va2.data(), nullptr, B, &c);
    if (rc) return rc;
    c->want_iw = m->want_iw;
    m->child = c;
    m->child_grid = grid;
    m->child_split_rows = split_rows;
    hipStream_t st = stream();
    if (any_split) {
        MF_HIP(hipMalloc((void **)&c->d_split, split.size()));
        MF_HIP(hipMemcpyAsync(c->d_split, split.data(), split.size(), hipMemcpyHostToDevice, st));
    }
    for (void *p : {(void *)m->d_perm, (void *)m->d_pb, (void *)m->d_px})
        if (p) MF_HIP(hipFree(p));
    m->d_perm = nullptr;
    m->d_pb = m->d_px = nullptr;
    MF_HIP(hipMalloc((void **)&m->d_perm, sizeof(int) * (size_t)n2));
    MF_HIP(hipMalloc((void **)&m->d_pb, sizeof(double) * (size_t)n2));
    MF_HIP(hipMalloc((void **)&m->d_px, sizeof(double) * (size_t)n2));
    // on the library stream, ordered before the solves that read them
    MF_HIP(hipMemcpyAsync(m->d_perm, perm.data(), sizeof(int) * (size_t)n2, hipMemcpyHostToDevice, st));
    MF_HIP(hipMemsetAsync(m->d_pb, 0, sizeof(double) * (size_t)n2, st));
    MF_HIP(hipMemsetAsync(m->d_px, 0, sizeof(double) * (size_t)n2, st));
    MF_HIP(hipStreamSynchronize(st));
    return MF_OK;
}

int solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep)
{
    std::memset(rep, 0, sizeof(*rep));
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    if (o.reorder < MF_REORDER_AUTO || o.reorder > MF_REORDER_ON) return fail(MF_E_ARG, "reorder mode");
    double pad = 0.0;
    bool use = false;
    double t0 = now_ms();
    if (o.reorder != MF_REORDER_OFF) {
        const int want = o.index_width ? o.index_width : m->want_iw;
        if (!m->L64.built || m->L64.want_iw != want) {
            int rc = build_layout(m, m->L64, 0.0, 0, want);
            if (rc) return rc;
        }
        // stored entries of the natural FP64 layout against the nonzeros
        pad = (double)m->L64.blob_bytes / (8.0 + m->L64.iw) / std::max(1, m->nnz);
        use = o.reorder == MF_REORDER_ON || pad > 2.0;
    }
    if (!use) {
        int rc = solve_direct(m, d_b, d_x, o, rep);
        rep->pad_factor = pad;
        return rc;
    }
    if (!m->child || m->child_grid != o.grid) {
        int rc = build_reordered(m, o.grid);
        if (rc) return rc;
    }
    const double ms_reorder = now_ms() - t0;
    hipStream_t st = stream();
    const int n = m->child->rowA;  // rows of the copy (every row of A, split segments, padding)
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipEvent_t e[4];
    for (hipEvent_t &ev : e) MF_HIP(hipEventCreate(&ev));
    MF_HIP(hipEventRecord(e[0], st));
    k_permute_in<<<nb, 256, 0, st>>>(m->d_perm, n, d_b, m->d_pb);
    MF_HIP(hipGetLastError());
    MF_HIP(hipEventRecord(e[1], st));
    mf_opts oc = o;
    oc.reorder = MF_REORDER_OFF;
    int rc = solve_direct(m->child, m->d_pb, m->d_px, oc, rep);
    MF_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * (size_t)m->tilem * m->B, st));
    MF_HIP(hipEventRecord(e[2], st));
    k_permute_out<<<nb, 256, 0, st>>>(m->d_perm, n, m->d_px, d_x);
    MF_HIP(hipGetLastError());
    MF_HIP(hipEventRecord(e[3], st));
    MF_HIP(hipEventSynchronize(e[3]));
    float a = 0.0f, b = 0.0f;
    MF_HIP(hipEventElapsedTime(&a, e[0], e[1]));
    MF_HIP(hipEventElapsedTime(&b, e[2], e[3]));
    for (hipEvent_t &ev : e) (void)hipEventDestroy(ev);
    rep->ms_solve += a + b;  // the permutations count as part of the solve
    rep->ms_setup += ms_reorder;
    rep->reordered = 1;
    rep->pad_factor = pad;
    rep->split_rows = m->child_split_rows;
    return rc;
}

__global__ void k_zero(double *__restrict__ x, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// mf_solve_device_noalloc: the matrix itself, or (opts.reorder as in
// solve_device) its row-reordered copy, which an earlier mf_solve_device
// must have built -- the copy's buffers exist then, and only kernels run here
// (fills are kernels, so the call can sit in a captured graph).
int solve_noalloc(mf_matrix m, int dev, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep)
{
    hipStream_t st = stream();
    std::memset(rep, 0, sizeof(*rep));
    if (dev < 0 || dev >= NA_MAX_DEV || !g_scratch[dev])
        return fail(MF_E_NODEV, "allocation-free solve: device not set up (mf_init)");
    if (m->rowA != m->colA) return fail(MF_E_ARG, "solver needs a square matrix");
    if (o.reorder < MF_REORDER_AUTO || o.reorder > MF_REORDER_ON) return fail(MF_E_ARG, "reorder mode");
    const size_t ncol = (size_t)std::max(m->tilem, m->tilen) * m->B;
    if (ncol > NA_MAX_ROWS)
        return fail(MF_E_ARG, "allocation-free solve: " + std::to_string(ncol) + " padded rows exceed the limit of " +
                                  std::to_string(NA_MAX_ROWS));
    const int want = o.index_width ? o.index_width : m->want_iw;
    if (!m->L64.built || m->L64.want_iw != want)
        return fail(MF_E_ARG, "allocation-free solve: FP64 layout not built (mf_matrix_layout or a solve first)");
    double pad = 0.0;
    bool use = false;
    if (o.reorder != MF_REORDER_OFF) {
        pad = (double)m->L64.blob_bytes / (8.0 + m->L64.iw) / std::max(1, m->nnz);
        use = o.reorder == MF_REORDER_ON || pad > 2.0;
    }
    if (use && (!m->child || m->child_grid != o.grid))
        return fail(MF_E_ARG, "allocation-free solve: the row-reordered copy is not built (an earlier "
                              "mf_solve_device with these options builds it)");
    mf_opts oc = o;
    oc.reorder = MF_REORDER_OFF;
    MF_HIP(hipEventRecord(g_na_ev[dev][0], st));
    int rc;
    if (!use) {
        rc = noalloc_core(m, dev, d_b, d_x, oc, rep);
    } else {
        const int n2 = m->child->rowA;
        const unsigned nb = (unsigned)((n2 + 255) / 256);
        k_permute_in<<<nb, 256, 0, st>>>(m->d_perm, n2, d_b, m->d_pb);
        MF_HIP(hipGetLastError());
        rc = noalloc_core(m->child, dev, m->d_pb, m->d_px, oc, rep);
        if (rc && rc != MF_E_TIMEOUT) return rc;
        const long long nx = (long long)m->tilem * m->B;
        k_zero<<<(unsigned)((nx + 255) / 256), 256, 0, st>>>(d_x, nx);
        k_permute_out<<<nb, 256, 0, st>>>(m->d_perm, n2, m->d_px, d_x);
        MF_HIP(hipGetLastError());
        rep->reordered = 1;
        rep->split_rows = m->child_split_rows;
    }
    if (rc && rc != MF_E_TIMEOUT) return rc;
    MF_HIP(hipEventRecord(g_na_ev[dev][1], st));
    MF_HIP(hipEventSynchronize(g_na_ev[dev][1]));
    float ms = 0.0f;
    MF_HIP(hipEventElapsedTime(&ms, g_na_ev[dev][0], g_na_ev[dev][1]));
    rep->ms_solve = ms;
    rep->pad_factor = pad;
    return rc;
}

}  // namespace mf
