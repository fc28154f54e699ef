// Batched persistent CG: many independent SPD systems of uneven size, each
// with K right-hand sides, solved by one launch captured in a hipGraph.
//
// Every system keeps its own compute layouts (built on the GPU exactly as for
// a single matrix); the batch concatenates them into one block-diagonal
// layout (tile rows, tile columns and slab offsets rebased), once in FP64 and
// once with per-tile precision classes.  Vectors live in HBM, the K columns
// of a row interleaved ([row][K]), so one slab read and one gather index feed
// K FMAs: the matrix is streamed once per iteration for all right-hand sides.
// System s runs on a group of W_s workgroups (W_s sized to its work); the
// group's reductions are block-local when W_s = 1 and otherwise use a
// tagged-word sync restricted to the group (GroupSync).
// Each right-hand side has its own CG scalars; a column that has converged
// is frozen (alpha = beta = 0) while the others continue, and the system
// finishes when an FP64 true-residual pass over the FP64 layout shows every
// column at the tolerance.  With lossy tiles the same passes (reliable
// updates against the exact operator) carry the solve to 1e-10.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mf_matrix.h"

namespace mf {

constexpr int BK_MAXV = 24;  // values per group allreduce: 3 K with K <= 8

struct BatchArgs {
    DevL L;    // inner layout (mixed classes, or FP64)
    DevL L64;  // FP64 layout: true residuals
    int G, nrhs;
    int nlines;                           // sync lines per parity buffer (G, or the on-chip tasks' members)
    const int *wg_task;                   // [G + 1]: tasks wg_task[g] .. wg_task[g + 1] - 1 of workgroup g
    const int *task;                      // [5 t]: system, tile rows tr0, tr1, group's first workgroup, group size;
                                          // on-chip kernel [6 t]: system, tr0, tr1, first sync line, gang size,
                                          // index in the gang
    const double *b;                      // [rows][K]
    double *x, *r, *p, *s;                // [rows][K]
    float *r32;                           // FP32 copy of r for low-precision tiles
    unsigned long long *slots;            // sync words [2][G][2 * BK_MAXV]
    int *ctl;                             // [0] abort word
    int *sys_out;                         // [4 s + 0] status, +1 iterations, +2 replacements, +3 time (us)
    double *sys_relres;                   // max over columns of the true relative residual
    double tol, rr_factor;
    int maxit, haslow;
    long long spin_limit;
    unsigned long long *prof;  // optional per workgroup: [4 g] spmv, update, sync, pass (100 MHz ticks)
};

__device__ __forceinline__ void st_ag(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ag(float *p, float v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Group allreduce over the W workgroups of one system (a group of one only
// reduces across its waves).  Each FP64 partial goes out as two 64-bit words
// whose high halves carry the sync epoch and whose low halves carry one half
// of the value each, so a word showing the current epoch holds the current
// half whatever order the words arrive in; every workgroup of the group polls
// the group's words until all show the epoch, lane q reading member q's words
// (members q + 64, ... too), so one poll costs one memory latency whatever the
// group size, and sums them in a fixed order (per lane, then a butterfly over
// the wave: identical bits everywhere).  Two buffers alternate by epoch
// parity: a buffer is rewritten only after every member has passed the next
// sync, which it can do only after reading the buffer.
// The epoch restarts at 0 for every task, and the sync words are cleared once
// per launch, so a group's lines (gw0 .. gw0 + gn - 1) must not carry words of
// an earlier task of the same launch: batch_assign gives a workgroup's
// multi-workgroup task the first place in its list and no workgroup two such
// tasks, and batch_assign_oc gives every gang lines of its own.
template <int TPB>
struct GroupSync {
    double *s_red;  // [TPB/64][BK_MAXV]
    double *s_bc;   // [BK_MAXV]
    int *s_flag;
    int gi, gn, gw0;  // index in group, group size, first workgroup of the group
    unsigned ep = 0;

    // value(k), k < N: this thread's partial of sum k; the sums are left in
    // s_bc (read them there before the next allreduce)
    template <int N, typename F>
    __device__ bool allreduce_f(const BatchArgs &A, F value, bool publish)
    {
        static_assert(N <= BK_MAXV, "group reducer width");
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const double s = wave_sum(value(k));
            if (lane == 0) s_red[wv * BK_MAXV + k] = s;
        }
        if (publish && gn > 1) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        ++ep;
        if (tid < 64) {
            double part = 0.0;
            if (lane < N)
                for (int w = 0; w < TPB / 64; w++) part += s_red[w * BK_MAXV + lane];
            int ab = 0;
            if (gn == 1) {
                if (lane < N) s_bc[lane] = part;
            } else {
                unsigned long long *W = A.slots + (size_t)(ep & 1u) * A.nlines * 2 * BK_MAXV;
                const unsigned long long tag = (unsigned long long)ep << 32;
                if (lane < N) {
                    const unsigned long long bits = (unsigned long long)__double_as_longlong(part);
                    unsigned long long *mine = W + (size_t)(gw0 + gi) * 2 * BK_MAXV + 2 * lane;
                    __hip_atomic_store(mine, tag | (bits & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(mine + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                long long spins = 0;
                double acc[N];
                for (;;) {
                    // the abort word loads with the members' words (one latency a round)
                    const int abw = ld_agent(A.ctl);
                    bool ok = true;
#pragma unroll
                    for (int k = 0; k < N; k++) acc[k] = 0.0;
                    for (int q = lane; q < gn; q += 64) {
                        const unsigned long long *w = W + (size_t)(gw0 + q) * 2 * BK_MAXV;
#pragma unroll
                        for (int k = 0; k < N; k++) {
                            const unsigned long long lo =
                                __hip_atomic_load(w + 2 * k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            const unsigned long long hi =
                                __hip_atomic_load(w + 2 * k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            ok = ok && (lo >> 32) == ep && (hi >> 32) == ep;
                            acc[k] += __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
                        }
                    }
                    if (__all(ok)) break;
                    if (__builtin_amdgcn_readfirstlane(abw)) { ab = 1; break; }
                    if (++spins > A.spin_limit) {
                        if (lane == 0) __hip_atomic_store(A.ctl, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ab = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const double t = wave_sum(acc[k]);
                    if (lane == 0) s_bc[k] = t;
                }
                if (!ab && publish) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            if (lane == 0) *s_flag = ab;
        }
        __syncthreads();
        return !*s_flag;
    }

    template <int N>
    __device__ bool allreduce(const BatchArgs &A, const double (&v)[N], bool publish)
    {
        return allreduce_f<N>(A, [&](int k) { return v[k]; }, publish);
    }
};

// Entries k0 .. k0 + KL - 1 of row col of a [row][K] vector (KL = 1, 2 or 8;
// k0 is a multiple of KL, so the vector loads are aligned).
template <int K, int KL>
__device__ __forceinline__ void ldKL(const double *__restrict__ g, uint32_t col, int k0, double (&o)[KL])
{
    const double *p = g + (size_t)col * K + k0;
    if constexpr (KL == 1) {
        o[0] = *p;
    } else {
#pragma unroll
        for (int j = 0; j < KL / 2; j++) {
            const double2 d = reinterpret_cast<const double2 *>(p)[j];
            o[2 * j] = d.x;
            o[2 * j + 1] = d.y;
        }
    }
}
template <int K, int KL>
__device__ __forceinline__ void ldKL(const float *__restrict__ g, uint32_t col, int k0, float (&o)[KL])
{
    const float *p = g + (size_t)col * K + k0;
    if constexpr (KL == 1) {
        o[0] = *p;
    } else {
#pragma unroll
        for (int j = 0; j < KL / 2; j++) {
            const float2 d = reinterpret_cast<const float2 *>(p)[j];
            o[2 * j] = d.x;
            o[2 * j + 1] = d.y;
        }
    }
}

// One class slab of this lane's row applied to its KL columns k0 .. k0 + KL - 1.
template <int B, int IW, int CLS, int K, int KL, typename AccT, typename GT>
__device__ __forceinline__ void slabK(const DevL &L, int2 h, int tcr, int t0, int r, int lanebase, int k0,
                                      const GT *__restrict__ g, AccT (&acc)[KL])
{
    const int nq = (h.y + 3) >> 2;
    const uint8_t *vals = L.blob + (size_t)(unsigned)h.x * 8;
    const uint8_t *idx = vals + (size_t)nq * B * 4 * cls_width(CLS);
    for (int j = 0; j < nq; j++) {
        Quad<CLS> v;
        IdxQuad<IW> id;
        v.load(vals, (size_t)j * B + r);
        id.load(idx, (size_t)j * B + r);
        GT xg[4][KL];
        static_for<4>([&](auto e) {
            ldKL<K, KL>(g, gather_col<B, IW>(id.template get<e>(), tcr, t0, L, lanebase), k0, xg[e]);
        });
        static_for<4>([&](auto e) {
            const AccT a = (AccT)v.template get<e>();
#pragma unroll
            for (int k = 0; k < KL; k++) acc[k] = fma_t(a, (AccT)xg[e][k], acc[k]);
        });
    }
}

// index base of tile row tr (as load_tcr)
template <int IW, int B>
__device__ __forceinline__ int row_tcr(const DevL &L, int tr, int t0, int nt, int r)
{
    return IW == 1 ? (r < nt ? L.stc[t0 + r] * B : 0) : IW == 2 ? tr * B : 0;
}

// y = A_row g for the lane's KL columns: FP64 tiles in FP64 over g64, the
// other classes in FP32 over g32, as the single-vector SpMV.
template <int B, int IW, int K, int KL>
__device__ __forceinline__ void row_spmvK(const DevL &L, int tr, int r, int lanebase, int k0,
                                          const double *__restrict__ g64, const float *__restrict__ g32,
                                          double (&y)[KL])
{
    const int4 *h = reinterpret_cast<const int4 *>(L.trh + (size_t)tr * 4);
    const int4 h01 = h[0], h23 = h[1];
    const int t0 = L.tile_ptr[tr], nt = L.tile_ptr[tr + 1] - t0;
    const int tcr = row_tcr<IW, B>(L, tr, t0, nt, r);
    double a64[KL];
    float a32[KL];
#pragma unroll
    for (int k = 0; k < KL; k++) {
        a64[k] = 0.0;
        a32[k] = 0.0f;
    }
    if (h01.y) slabK<B, IW, 0, K, KL>(L, make_int2(h01.x, h01.y), tcr, t0, r, lanebase, k0, g64, a64);
    if (h01.w) slabK<B, IW, 1, K, KL>(L, make_int2(h01.z, h01.w), tcr, t0, r, lanebase, k0, g32, a32);
    if (h23.y) slabK<B, IW, 2, K, KL>(L, make_int2(h23.x, h23.y), tcr, t0, r, lanebase, k0, g32, a32);
    if (h23.w) slabK<B, IW, 3, K, KL>(L, make_int2(h23.z, h23.w), tcr, t0, r, lanebase, k0, g32, a32);
#pragma unroll
    for (int k = 0; k < KL; k++) y[k] = a64[k] + (double)a32[k];
}

// A x and A p on the FP64 layout (class 0 only) for the lane's KL columns.
template <int B, int IW, int K, int KL>
__device__ __forceinline__ void row_spmv64_xp(const DevL &L, int tr, int r, int lanebase, int k0,
                                              const double *__restrict__ gx, const double *__restrict__ gp,
                                              double (&ax)[KL], double (&ap)[KL])
{
    const int2 h = reinterpret_cast<const int2 *>(L.trh)[(size_t)tr * 4];
    const int t0 = L.tile_ptr[tr], nt = L.tile_ptr[tr + 1] - t0;
    const int tcr = row_tcr<IW, B>(L, tr, t0, nt, r);
#pragma unroll
    for (int k = 0; k < KL; k++) ax[k] = ap[k] = 0.0;
    const int nq = (h.y + 3) >> 2;
    const uint8_t *vals = L.blob + (size_t)(unsigned)h.x * 8;
    const uint8_t *idx = vals + (size_t)nq * B * 4 * 8;
    for (int j = 0; j < nq; j++) {
        Quad<0> v;
        IdxQuad<IW> id;
        v.load(vals, (size_t)j * B + r);
        id.load(idx, (size_t)j * B + r);
        static_for<4>([&](auto e) {
            const uint32_t c = gather_col<B, IW>(id.template get<e>(), tcr, t0, L, lanebase);
            double vx[KL], vp[KL];
            ldKL<K, KL>(gx, c, k0, vx);
            ldKL<K, KL>(gp, c, k0, vp);
            const double a = v.template get<e>();
#pragma unroll
            for (int k = 0; k < KL; k++) {
                ax[k] = __builtin_fma(a, vx[k], ax[k]);
                ap[k] = __builtin_fma(a, vp[k], ap[k]);
            }
        });
    }
}

// Per-column CG scalars of the system at hand, in LDS: gamma, previous
// gamma, previous alpha, reference ||r||^2 of the last replacement, ||b||^2,
// (r, s) and (p, s) of the last replacement, alpha, beta; mode bits: 0
// restart, 1 replaced, 2 fresh, 3 after replacement, 4 frozen; breakdown
// flags: bit 0 in a fresh step, bit 1 otherwise.
template <int K, int TPB>
struct BatchShared {
    double red[(TPB / 64) * BK_MAXV];
    double bc[BK_MAXV];
    int flag;
    double g[K], gp[K], ap[K], ref[K], b2[K], rs[K], ps[K], al[K], be[K];
    double tprev[K];  // true (r, r) of each column at the last true-residual pass
    int mode[K];
    int bd;
    int use64;        // 1: the system's SpMVs read the FP64 layout
};

// CG on one system (tile rows tr0 .. tr1 - 1 of it in this workgroup);
// returns false when the group sync gave up (abort word raised).
template <int B, int IW, int K, int TPB>
__device__ bool cg_system(const BatchArgs &A, BatchShared<K, TPB> &sh, int sys, int tr0, int tr1, int gw0, int gn)
{
    // K = 8: four lanes per row, two columns each (the lanes of a row read the
    // same slab quads, which one wave fetches once); K = 1: one lane per row
    // SpMV phase: one lane per row, all K columns (fewer memory instructions
    // per FMA than splitting the columns, which measured slower).  The FP64
    // true-residual pass, which needs A x and A p, splits the K = 8 columns
    // over four lanes of a row (two each) to keep its registers below the
    // SpMV's.
    constexpr int GPB = TPB / B;                      // tile rows per sweep
    constexpr int CGP = K == 8 ? 4 : 1, KLP = K / CGP;  // pass: lanes per row, columns per lane
    constexpr int GPBP = TPB / (B * CGP);
    double *sh_g = sh.g, *sh_gp = sh.gp, *sh_ap = sh.ap, *sh_ref = sh.ref, *sh_b2 = sh.b2, *sh_rs = sh.rs,
           *sh_ps = sh.ps, *sh_al = sh.al, *sh_be = sh.be;
    int *sh_mode = sh.mode;
    int &sh_bd = sh.bd;
    const int g = blockIdx.x, tid = threadIdx.x;
    const int grp = tid / B, lane = tid % B, lanebase = (tid & 63) & ~(B - 1);
    const int grpp = tid / (B * CGP), k0p = ((tid / B) % CGP) * KLP;  // pass mapping
    GroupSync<TPB> S{sh.red, sh.bc, &sh.flag, g - gw0, gn, gw0};
    const int KA = A.nrhs;
    const double tol2 = A.tol * A.tol, rr2 = A.rr_factor * A.rr_factor;
    const unsigned long long t_start = wall_clock64();
    auto finish = [&](int status, int iter, int repl, double relres) {
        if (g == gw0 && tid == 0) {
            A.sys_out[4 * sys] = status;
            A.sys_out[4 * sys + 1] = iter;
            // bit 30: the system left its lossy tiles for the FP64 layout
            A.sys_out[4 * sys + 2] = repl | (sh.use64 && A.haslow ? 1 << 30 : 0);
            // time of the system on its first workgroup (us; the clock runs at 100 MHz)
            A.sys_out[4 * sys + 3] = (int)min((wall_clock64() - t_start) / 100ull, 0x7fffffffull);
            A.sys_relres[sys] = relres;
        }
    };

    double v[K];
    {
        const int km = tid % K;
        double acc = 0.0;
        for (size_t e = (size_t)tr0 * B * K + tid; e < (size_t)tr1 * B * K; e += TPB) {
            const double bv = A.b[e];
            st_ag(A.x + e, 0.0);
            st_ag(A.p + e, 0.0);
            A.s[e] = 0.0;
            st_ag(A.r + e, bv);
            if (A.haslow) st_ag(A.r32 + e, (float)bv);
            acc += bv * bv;
        }
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = k == km ? acc : 0.0;
    }
    if (!S.allreduce(A, v, true)) return false;
    if (tid < K) {
        sh_b2[tid] = sh_g[tid] = sh_ref[tid] = sh.tprev[tid] = sh.bc[tid];
        sh_gp[tid] = sh_ap[tid] = 1.0;
        sh_mode[tid] = 1;
    }
    if (tid == 0) sh.use64 = A.haslow ? 0 : 1;
    __syncthreads();
    int iter = 0, repl = 0;
    bool force = false;
    unsigned long long *pf = A.prof && tid == 0 ? A.prof + 4 * (size_t)g : nullptr;
    unsigned long long tk = pf ? wall_clock64() : 0;
    auto lap = [&](int k) {
        if (pf) {
            const unsigned long long now = wall_clock64();
            pf[k] += now - tk;
            tk = now;
        }
    };
    for (;;) {
        bool conv = true, due = A.rr_factor > 0.0, finite = true;
        for (int k = 0; k < KA; k++) {
            const double gk = sh_g[k];
            finite = finite && isfinite(gk);
            conv = conv && gk <= tol2 * sh_b2[k];
            due = due && gk <= rr2 * sh_ref[k];
        }
        if (!finite) { finish(MF_E_NONFINITE, iter, repl, -1.0); return true; }
        if (conv || due || force) {
            // r = b - A x and s = A p on the FP64 layout; (r,r), (r,s), (p,s)
            if (!S.allreduce(A, v, true)) return false;  // x and p of every row visible
            double t0[KLP], t1[KLP], t2[KLP];
#pragma unroll
            for (int j = 0; j < KLP; j++) t0[j] = t1[j] = t2[j] = 0.0;
            for (int tr = tr0 + grpp; tr < tr1; tr += GPBP) {
                const size_t row = (size_t)tr * B + lane;
                double ax[KLP], ap[KLP];
                row_spmv64_xp<B, IW, K, KLP>(A.L64, tr, lane, lanebase, k0p, A.x, A.p, ax, ap);
#pragma unroll
                for (int j = 0; j < KLP; j++) {
                    const size_t e = row * K + k0p + j;
                    const double rv = A.b[e] - ax[j], pv = A.p[e];
                    st_ag(A.r + e, rv);
                    if (A.haslow) st_ag(A.r32 + e, (float)rv);
                    A.s[e] = ap[j];
                    t0[j] += rv * rv;
                    t1[j] += rv * ap[j];
                    t2[j] += pv * ap[j];
                }
            }
            // sum k: column k % K of (r, r), (r, s), (p, s) for k / K = 0, 1, 2
            auto tv = [&](int k) {
                const int c = k % K, j = c - k0p, q = k / K;
                double v = 0.0;
#pragma unroll
                for (int i = 0; i < KLP; i++)
                    if (j == i) v = q == 0 ? t0[i] : q == 1 ? t1[i] : t2[i];
                return v;
            };
            if (!S.template allreduce_f<3 * K>(A, tv, true)) return false;
            lap(3);
            ++repl;
            bool done = true;
            double rel = 0.0;
            for (int k = 0; k < KA; k++) {
                done = done && sh.bc[k] <= tol2 * sh_b2[k];
                if (sh_b2[k] > 0.0) rel = fmax(rel, sqrt(sh.bc[k] / sh_b2[k]));
            }
            if (!isfinite(rel)) { finish(MF_E_NONFINITE, iter, repl, rel); return true; }
            if (done) { finish(MF_OK, iter, repl, rel); return true; }
            // lossy tiles at their accuracy floor: a pass that follows a drop
            // of the recurrence by rr_factor but shows a column's true residual
            // down by less than 2x (or a breakdown) moves the system to the FP64
            // layout for the rest of its solve (x, r and the direction are kept)
            if (tid == 0) {
                bool stall = false;
                for (int k = 0; k < KA; k++) {
                    stall = stall || (sh.bc[k] > tol2 * sh_b2[k] && sh.bc[k] > 0.25 * sh.tprev[k]);
                    sh.tprev[k] = sh.bc[k];
                }
                if (stall) sh.use64 = 1;
            }
            if (tid < K) {
                const bool restart = force || sh.bc[tid] > 4.0 * sh_g[tid];
                sh_mode[tid] = (restart ? 1 : 0) | 2;
                sh_rs[tid] = sh.bc[K + tid];
                sh_ps[tid] = sh.bc[2 * K + tid];
                sh_g[tid] = sh_ref[tid] = sh.bc[tid];
            }
            force = false;
            __syncthreads();
            continue;
        }
        if (iter >= A.maxit) {
            double rel = 0.0;
            for (int k = 0; k < KA; k++)
                if (sh_b2[k] > 0.0) rel = fmax(rel, sqrt(sh_g[k] / sh_b2[k]));
            finish(MF_E_MAXIT, iter, repl, rel);
            return true;
        }
        if (tid < K) {
            const int m = sh_mode[tid];
            const bool fresh = iter == 0 || (m & 1);
            const bool frozen = tid >= KA || sh_g[tid] <= tol2 * sh_b2[tid];
            sh_be[tid] = (fresh || frozen) ? 0.0 : sh_g[tid] / sh_gp[tid];
            sh_mode[tid] = (fresh ? 4 : 0) | ((!fresh && (m & 2)) ? 8 : 0) | (frozen ? 16 : 0);
        }
        if (tid == 0) sh_bd = 0;
        __syncthreads();
        // w = A r; s = w + beta s; (w, r)
        double d[K];
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = 0.0;
        for (int tr = tr0 + grp; tr < tr1; tr += GPB) {
            const size_t row = (size_t)tr * B + lane;
            double w[K];
            row_spmvK<B, IW, K, K>(sh.use64 ? A.L64 : A.L, tr, lane, lanebase, 0, A.r, A.r32, w);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const size_t e = row * K + k;
                A.s[e] = w[k] + sh_be[k] * A.s[e];
                d[k] += w[k] * A.r[e];
            }
        }
        lap(0);
        if (!S.allreduce(A, d, false)) return false;
        lap(2);
        if (tid < K) {
            const int m = sh_mode[tid];
            const double be = sh_be[tid], gk = sh_g[tid];
            const double dk = sh.bc[tid];
            const double denom = (m & 4)   ? dk
                                 : (m & 8) ? dk + be * (2.0 * sh_rs[tid] + be * sh_ps[tid])
                                           : dk - be * gk / sh_ap[tid];
            if (m & 16) {
                sh_al[tid] = 0.0;
            } else if (!(denom > 0.0) || !isfinite(denom)) {
                sh_al[tid] = 0.0;
                atomicOr(&sh_bd, (m & 4) ? 1 : 2);
            } else {
                sh_al[tid] = gk / denom;
            }
        }
        __syncthreads();
        if (sh_bd) {
            // rounding broke (p, A p) > 0: recompute the residual and restart;
            // a breakdown in a fresh step is final
            if (sh_bd & 1) { finish(MF_E_BREAKDOWN, iter, repl, -1.0); return true; }
            force = true;
            continue;
        }
        // p, x, r updates stream the workgroup's elements [row][K] with unit
        // stride; every thread keeps one column (TPB is a multiple of K)
        double gg[K];
        {
            const int km = tid % K;
            const double al = sh_al[km], be = sh_be[km];
            const size_t e1 = (size_t)tr1 * B * K;
            double acc = 0.0;
            constexpr int U = 4;
            for (size_t e0 = (size_t)tr0 * B * K + tid; e0 < e1; e0 += (size_t)U * TPB) {
                double rv0[U], pv0[U], xv0[U], sv0[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const size_t e = min(e0 + (size_t)u * TPB, e1 - 1);
                    rv0[u] = A.r[e];
                    pv0[u] = A.p[e];
                    xv0[u] = A.x[e];
                    sv0[u] = A.s[e];
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const size_t e = e0 + (size_t)u * TPB;
                    if (e < e1) {
                        const double pv = rv0[u] + be * pv0[u];
                        const double rv = rv0[u] - al * sv0[u];
                        st_ag(A.p + e, pv);
                        st_ag(A.x + e, xv0[u] + al * pv);
                        st_ag(A.r + e, rv);
                        if (A.haslow) st_ag(A.r32 + e, (float)rv);
                        acc += rv * rv;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) gg[k] = k == km ? acc : 0.0;
        }
        lap(1);
        if (!S.allreduce(A, gg, true)) return false;
        lap(2);
        if (tid < K) {
            if (!(sh_mode[tid] & 16)) {
                sh_gp[tid] = sh_g[tid];
                sh_ap[tid] = sh_al[tid];
            }
            sh_g[tid] = sh.bc[tid];
            sh_mode[tid] = 0;
        }
        __syncthreads();
        ++iter;
    }
}

template <int B, int IW, int K, int TPB>
__global__ __launch_bounds__(TPB) void k_cg_batched(BatchArgs A)
{
    __shared__ BatchShared<K, TPB> sh;
    for (int t = A.wg_task[blockIdx.x]; t < A.wg_task[blockIdx.x + 1]; t++) {
        const int *q = A.task + 5 * (size_t)t;
        if (!cg_system<B, IW, K, TPB>(A, sh, q[0], q[1], q[2], q[3], q[4])) return;
        __syncthreads();  // LDS scalars are reused by the next system
    }
}

// ------------------------------------------------ on-chip batched solver
// Each system runs on a gang of workgroups holding at most OC_RPT * TPB rows
// each, with the Krylov vectors of its rows on chip for the whole solve: x
// and p of the thread's rows (8 columns each) in registers, r and s in LDS;
// only r (and its FP32 copy for low-precision tiles) goes to HBM, published
// for the gang's gathers, so an iteration moves the matrix and one vector
// instead of four vectors read and written.  Hestenes-Stiefel CG per column
// with s = A p by recurrence and (p, A p) summed directly, two gang syncs per
// iteration, FP64 true-residual passes on the FP64 layout (x and p published
// for them), converged columns frozen, lossy systems moved to the FP64
// layout when their true residual stalls -- the same algorithm as cg_system.
// A workgroup runs its tasks in the order of their scheduled start; a gang's
// members share one start, so no cycle of waits can form.  Each task has its
// own sync lines (first line, gang size), so the epochs of one task never
// meet words of another.
constexpr int OC_TPB = 512, OC_RPT = 1, OC_K = 8;

template <int TPB, int RPT>
struct OcShared {
    double x[RPT * OC_K * TPB], r[RPT * OC_K * TPB], s[RPT * OC_K * TPB], p[RPT * OC_K * TPB];  // [i][k][thread]
    double red[(TPB / 64) * BK_MAXV];
    double bc[BK_MAXV];
    int flag;
    double g[OC_K], gp[OC_K], ref[OC_K], b2[OC_K], al[OC_K], be[OC_K], tprev[OC_K];
    int mode[OC_K];
    int bd, use64;
};

template <int B, int IW, int TPB, int RPT>
__device__ bool cg_system_oc(const BatchArgs &A, OcShared<TPB, RPT> &sh, int sys, int tr0, int tr1, int line0,
                             int gn, int gi)
{
    constexpr int K = OC_K, GPB = TPB / B;
    const int tid = threadIdx.x, grp = tid / B, lane = tid % B, lanebase = (tid & 63) & ~(B - 1);
    GroupSync<TPB> S{sh.red, sh.bc, &sh.flag, gi, gn, line0};
    const int KA = A.nrhs;
    const double tol2 = A.tol * A.tol, rr2 = A.rr_factor * A.rr_factor;
    const unsigned long long t_start = wall_clock64();
    auto tile_row = [&](int i) { return tr0 + i * GPB + grp; };
    auto row_of = [&](int i) { return (size_t)tile_row(i) * B + lane; };
    auto R = [&](int i, int k) -> double & { return sh.r[(i * K + k) * TPB + tid]; };
    auto Sv = [&](int i, int k) -> double & { return sh.s[(i * K + k) * TPB + tid]; };
    auto P = [&](int i, int k) -> double & { return sh.p[(i * K + k) * TPB + tid]; };
    auto X = [&](int i, int k) -> double & { return sh.x[(i * K + k) * TPB + tid]; };
    auto publish_r = [&](int i) {
        const size_t e = row_of(i) * K;
#pragma unroll
        for (int k = 0; k < K; k++) {
            st_ag(A.r + e + k, R(i, k));
            if (A.haslow) st_ag(A.r32 + e + k, (float)R(i, k));
        }
    };
    auto finish = [&](int status, int iter, int repl, double relres) {
#pragma unroll
        for (int i = 0; i < RPT; i++)
            if (tile_row(i) < tr1) {
                const size_t e = row_of(i) * K;
#pragma unroll
                for (int k = 0; k < K; k++) A.x[e + k] = X(i, k);
            }
        if (gi == 0 && tid == 0) {
            A.sys_out[4 * sys] = status;
            A.sys_out[4 * sys + 1] = iter;
            A.sys_out[4 * sys + 2] = repl | (sh.use64 && A.haslow ? 1 << 30 : 0);
            A.sys_out[4 * sys + 3] = (int)min((wall_clock64() - t_start) / 100ull, 0x7fffffffull);
            A.sys_relres[sys] = relres;
        }
    };

    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = 0.0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            X(i, k) = R(i, k) = Sv(i, k) = P(i, k) = 0.0;
        }
        if (tile_row(i) < tr1) {
            const size_t e = row_of(i) * K;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const double bv = A.b[e + k];
                R(i, k) = bv;
                v[k] += bv * bv;
            }
            publish_r(i);
        }
    }
    if (!S.allreduce(A, v, true)) return false;
    if (tid < K) {
        sh.b2[tid] = sh.g[tid] = sh.ref[tid] = sh.tprev[tid] = sh.bc[tid];
        sh.gp[tid] = 1.0;
        sh.mode[tid] = 1;
    }
    if (tid == 0) sh.use64 = A.haslow ? 0 : 1;
    __syncthreads();
    int iter = 0, repl = 0;
    bool force = false;
    for (;;) {
        bool conv = true, due = A.rr_factor > 0.0, finite = true;
        for (int k = 0; k < KA; k++) {
            const double gk = sh.g[k];
            finite = finite && isfinite(gk);
            conv = conv && gk <= tol2 * sh.b2[k];
            due = due && gk <= rr2 * sh.ref[k];
        }
        if (!finite) { finish(MF_E_NONFINITE, iter, repl, -1.0); return true; }
        if (conv || due || force) {
            // r = b - A x and s = A p over the FP64 layout, from x and p
            // published for the gang
#pragma unroll
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    const size_t e = row_of(i) * K;
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        st_ag(A.x + e + k, X(i, k));
                        st_ag(A.p + e + k, P(i, k));
                    }
                }
            if (!S.allreduce(A, v, true)) return false;
            double t0[K];
#pragma unroll
            for (int k = 0; k < K; k++) t0[k] = 0.0;
#pragma unroll
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    const size_t e = row_of(i) * K;
                    // two halves of the columns (fewer live registers)
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        double ax[K / 2], ap[K / 2];
                        row_spmv64_xp<B, IW, K, K / 2>(A.L64, tile_row(i), lane, lanebase, h * (K / 2), A.x, A.p,
                                                       ax, ap);
#pragma unroll
                        for (int j = 0; j < K / 2; j++) {
                            const int k = h * (K / 2) + j;
                            const double rv = A.b[e + k] - ax[j];
                            R(i, k) = rv;
                            Sv(i, k) = ap[j];
                            t0[k] += rv * rv;
                        }
                    }
                    publish_r(i);
                }
            if (!S.allreduce(A, t0, true)) return false;
            ++repl;
            bool done = true;
            double rel = 0.0;
            for (int k = 0; k < KA; k++) {
                done = done && sh.bc[k] <= tol2 * sh.b2[k];
                if (sh.b2[k] > 0.0) rel = fmax(rel, sqrt(sh.bc[k] / sh.b2[k]));
            }
            if (!isfinite(rel)) { finish(MF_E_NONFINITE, iter, repl, rel); return true; }
            if (done) { finish(MF_OK, iter, repl, rel); return true; }
            if (tid == 0) {
                bool stall = false;
                for (int k = 0; k < KA; k++) {
                    stall = stall || (sh.bc[k] > tol2 * sh.b2[k] && sh.bc[k] > 0.25 * sh.tprev[k]);
                    sh.tprev[k] = sh.bc[k];
                }
                if (stall) sh.use64 = 1;
            }
            if (tid < K) {
                const bool restart = force || sh.bc[tid] > 4.0 * sh.g[tid];
                sh.mode[tid] = restart ? 1 : 0;
                sh.g[tid] = sh.ref[tid] = sh.bc[tid];
            }
            force = false;
            __syncthreads();
            continue;
        }
        if (iter >= A.maxit) {
            double rel = 0.0;
            for (int k = 0; k < KA; k++)
                if (sh.b2[k] > 0.0) rel = fmax(rel, sqrt(sh.g[k] / sh.b2[k]));
            finish(MF_E_MAXIT, iter, repl, rel);
            return true;
        }
        if (tid < K) {
            const int m = sh.mode[tid];
            const bool fresh = iter == 0 || (m & 1);
            const bool frozen = tid >= KA || sh.g[tid] <= tol2 * sh.b2[tid];
            sh.be[tid] = (fresh || frozen) ? 0.0 : sh.g[tid] / sh.gp[tid];
            sh.mode[tid] = (fresh ? 4 : 0) | (frozen ? 16 : 0);
        }
        if (tid == 0) sh.bd = 0;
        __syncthreads();
        // w = A r; s = w + beta s; p = r + beta p; (p, s)
        double d[K];
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = 0.0;
        const bool u64 = sh.use64;
#pragma unroll
        for (int i = 0; i < RPT; i++)
            if (tile_row(i) < tr1) {
                double w[K];
                row_spmvK<B, IW, K, K>(u64 ? A.L64 : A.L, tile_row(i), lane, lanebase, 0, A.r, A.r32, w);
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double bk = sh.be[k];
                    const double sv = w[k] + bk * Sv(i, k);
                    Sv(i, k) = sv;
                    const double pv = R(i, k) + bk * P(i, k);
                    P(i, k) = pv;
                    d[k] += pv * sv;
                }
            }
        if (!S.allreduce(A, d, false)) return false;
        if (tid < K) {
            const int m = sh.mode[tid];
            const double denom = sh.bc[tid];
            if (m & 16) {
                sh.al[tid] = 0.0;
            } else if (!(denom > 0.0) || !isfinite(denom)) {
                sh.al[tid] = 0.0;
                atomicOr(&sh.bd, (m & 4) ? 1 : 2);
            } else {
                sh.al[tid] = sh.g[tid] / denom;
            }
        }
        __syncthreads();
        if (sh.bd) {
            if (sh.bd & 1) { finish(MF_E_BREAKDOWN, iter, repl, -1.0); return true; }
            force = true;
            continue;
        }
        double gg[K];
#pragma unroll
        for (int k = 0; k < K; k++) gg[k] = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++)
            if (tile_row(i) < tr1) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double ak = sh.al[k];
                    X(i, k) = X(i, k) + ak * P(i, k);
                    const double rv = R(i, k) - ak * Sv(i, k);
                    R(i, k) = rv;
                    gg[k] += rv * rv;
                }
                publish_r(i);
            }
        if (!S.allreduce(A, gg, true)) return false;
        if (tid < K) {
            if (!(sh.mode[tid] & 16)) sh.gp[tid] = sh.g[tid];
            sh.g[tid] = sh.bc[tid];
            sh.mode[tid] = 0;
        }
        __syncthreads();
        ++iter;
    }
}

template <int B, int IW, int TPB, int RPT>
__global__ __launch_bounds__(TPB, 1) void k_cg_batched_oc(BatchArgs A)
{
    __shared__ OcShared<TPB, RPT> sh;
    for (int t = A.wg_task[blockIdx.x]; t < A.wg_task[blockIdx.x + 1]; t++) {
        const int *q = A.task + 6 * (size_t)t;
        if (!cg_system_oc<B, IW, TPB, RPT>(A, sh, q[0], q[1], q[2], q[3], q[4], q[5])) return;
        __syncthreads();  // LDS is reused by the next system
    }
}

__global__ void k_batch_reset(unsigned long long *slots, long long nslots, int *ctl, int *sys_out, int nsys)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nslots) slots[i] = 0ull;
    if (i < 8) ctl[i] = 0;
    if (i < 4ll * nsys) sys_out[i] = (i % 4 == 0) ? MF_E_TIMEOUT : 0;
}

__global__ void k_add_int(const int *__restrict__ src, int n, int add, int *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
}

// Four-byte slab indices hold the matrix column itself: shift every index of
// one system's slabs (blob at hand = that system's copy) by its first row.
template <int B>
__global__ void k_rebase_cols(const int2 *__restrict__ trh, int tilem, uint8_t *__restrict__ blob, unsigned add)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * tilem) return;
    const int2 h = trh[i];
    const int c = i & 3, nq = (h.y + 3) >> 2;
    if (!nq) return;
    uint32_t *idx = reinterpret_cast<uint32_t *>(blob + (size_t)(unsigned)h.x * 8 + (size_t)nq * B * 4 * cls_width(c));
    for (int e = 0; e < nq * B * 4; e++) idx[e] += add;
}

__global__ void k_rebase_trh(const int2 *__restrict__ src, int n, unsigned add8, int2 *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int2 h = src[i];
        dst[i] = make_int2((int)((unsigned)h.x + add8), h.y);
    }
}

}  // namespace mf

using namespace mf;

// One concatenated (block-diagonal) compute layout of the batch.
struct BatchLayout {
    int2 *trh = nullptr;
    int *stc = nullptr;
    uint8_t *blob = nullptr;
    size_t blob_bytes = 0;
    bool has64 = false, haslow = false;
    long long nnz_code[5] = {0, 0, 0, 0, 0};
    void release()
    {
        if (trh) (void)hipFree(trh);
        if (stc) (void)hipFree(stc);
        if (blob) (void)hipFree(blob);
        *this = BatchLayout();
    }
};

struct mf_batch_s {
    int nsys = 0, nrhs = 0, K = 0, B = 16, TPB = 256;
    bool mixed = false;
    int iw = 1;  // slab index width shared by every system
    mf_opts o;
    std::vector<int> n, tilem, trbase, nnz;
    std::vector<std::vector<int>> h_rp, h_ci;  // host copies of the systems (rebuilds)
    std::vector<std::vector<double>> h_va;
    std::vector<int> weights;                  // iteration weights of the assignment
    std::vector<double> wtime;                 // measured work per system (us on one workgroup), or empty
    std::string tuned;                         // what the tuner chose
    std::vector<long long> tbase;
    std::vector<double> tau;
    int tilem_total = 0;
    long long tilenum_total = 0;
    size_t rows = 0;
    int *d_tile_ptr = nullptr;
    BatchLayout L64, Lmix;
    int G = 0;
    std::vector<int> sys_nwg;
    bool oc = false;       // on-chip kernel (vectors in registers/LDS, gangs of workgroups)
    int fell_last = 0;     // systems of the last solve that left lossy tiles
    int nlines = 0;        // sync lines per parity buffer
    int *d_wg = nullptr;  // wg_task [G + 1], then task records [5][tasks]
    double *d_vec = nullptr;  // b, x, r, p, s: 5 x rows x K
    float *d_r32 = nullptr;
    unsigned long long *d_slots = nullptr;
    int *d_ctl = nullptr, *d_sys_out = nullptr;
    double *d_relres = nullptr;
    double *h_stage = nullptr;  // pinned [rows][K]
    unsigned long long *d_prof = nullptr;  // optional phase counters
    const void *fn = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    BatchArgs args{};
    double ms_setup = 0.0;
    double bytes_alg = 0.0;  // per iteration: matrix once + K (16 n) vector bytes
};

static void batch_release_data(mf_batch_s *bt);

static void batch_release(mf_batch_s *bt)
{
    if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    batch_release_data(bt);
    if (bt->d_wg) (void)hipFree(bt->d_wg);
    if (bt->d_slots) (void)hipFree(bt->d_slots);
    if (bt->d_prof) (void)hipFree(bt->d_prof);
}

typedef const void *KernelFn;

template <int B, int IW>
static KernelFn batched_fn(int K)
{
    if (K == 1) return (KernelFn)k_cg_batched<B, IW, 1, 256>;
    return (KernelFn)k_cg_batched<B, IW, 8, 256>;
}

template <int B> static KernelFn batched_fn_iw(int iw, int K)
{
    return iw == 1 ? batched_fn<B, 1>(K) : iw == 2 ? batched_fn<B, 2>(K) : batched_fn<B, 4>(K);
}

template <int B> static KernelFn batched_oc_fn(int iw)
{
    return iw == 1   ? (KernelFn)k_cg_batched_oc<B, 1, OC_TPB, OC_RPT>
           : iw == 2 ? (KernelFn)k_cg_batched_oc<B, 2, OC_TPB, OC_RPT>
                     : (KernelFn)k_cg_batched_oc<B, 4, OC_TPB, OC_RPT>;
}

static KernelFn pick_batched_oc(int B, int iw)
{
    switch (B) {
    case 8: return batched_oc_fn<8>(iw);
    case 16: return batched_oc_fn<16>(iw);
    default: return batched_oc_fn<32>(iw);
    }
}

static KernelFn pick_batched(int B, int iw, int K)
{
    switch (B) {
    case 8: return batched_fn_iw<8>(iw, K);
    case 16: return batched_fn_iw<16>(iw, K);
    default: return batched_fn_iw<32>(iw, K);
    }
}

// Concatenates one layout per system into a block-diagonal layout.
static int concat_layouts(mf_batch_s *bt, const std::vector<mf_matrix> &mats, bool mix, BatchLayout &out)
{
    hipStream_t st = stream();
    std::vector<size_t> base(bt->nsys + 1, 0);
    for (int s = 0; s < bt->nsys; s++) {
        const Layout &L = mix ? mats[s]->Lmix : mats[s]->L64;
        base[s + 1] = base[s] + (L.blob_bytes + 255) / 256 * 256;
        out.has64 = out.has64 || L.has64;
        out.haslow = out.haslow || L.haslow;
        for (int c = 0; c < 5; c++) out.nnz_code[c] += L.nnz_per_code[c];
    }
    out.blob_bytes = base[bt->nsys];
    if (out.blob_bytes / 8 > 0xFFFFFFFFull) return fail(MF_E_UNSUPPORTED, "batched layout larger than 32 GB");
    MF_HIP(hipMalloc((void **)&out.trh, sizeof(int2) * 4 * ((size_t)bt->tilem_total + 1)));
    MF_HIP(hipMalloc((void **)&out.stc, sizeof(int) * ((size_t)bt->tilenum_total + 1)));
    MF_HIP(hipMalloc((void **)&out.blob, std::max<size_t>(out.blob_bytes, 64)));
    MF_HIP(hipMemsetAsync(out.trh, 0, sizeof(int2) * 4 * ((size_t)bt->tilem_total + 1), st));
    MF_HIP(hipMemsetAsync(out.stc, 0, sizeof(int) * ((size_t)bt->tilenum_total + 1), st));
    for (int s = 0; s < bt->nsys; s++) {
        const Layout &L = mix ? mats[s]->Lmix : mats[s]->L64;
        const mf_matrix m = mats[s];
        const int n4 = 4 * m->tilem;
        k_rebase_trh<<<(n4 + 255) / 256, 256, 0, st>>>(L.trh, n4, (unsigned)(base[s] / 8), out.trh + (size_t)4 * bt->trbase[s]);
        if (m->tilenum)
            k_add_int<<<(m->tilenum + 255) / 256, 256, 0, st>>>(L.stc, m->tilenum, bt->trbase[s],
                                                               out.stc + bt->tbase[s]);
        if (L.blob_bytes) {
            MF_HIP(hipMemcpyAsync(out.blob + base[s], L.blob, L.blob_bytes, hipMemcpyDeviceToDevice, st));
            // four-byte indices are matrix columns: shift them by the system's
            // first row (the one- and two-byte forms are tile-row relative)
            if (L.iw == 4 && bt->trbase[s]) {
                const unsigned add = (unsigned)bt->trbase[s] * bt->B;
                const unsigned nb = (unsigned)((n4 + 255) / 256);
                if (bt->B == 8) k_rebase_cols<8><<<nb, 256, 0, st>>>(L.trh, m->tilem, out.blob + base[s], add);
                else if (bt->B == 16) k_rebase_cols<16><<<nb, 256, 0, st>>>(L.trh, m->tilem, out.blob + base[s], add);
                else k_rebase_cols<32><<<nb, 256, 0, st>>>(L.trh, m->tilem, out.blob + base[s], add);
            }
        }
    }
    MF_HIP(hipGetLastError());
    return MF_OK;
}

// Work assignment.  Work of system s = iterations x (bytes per iteration +
// a fixed per-iteration cost); a system split over w > 1 workgroups also pays
// its group syncs every iteration.  For a bound T on the work of one
// workgroup, every system above T gets the fewest workgroups that bring its
// share (plus the sync cost) under T, and the others are packed largest
// first onto the least loaded workgroup (after the split system on a split
// system's workgroup); bisection finds the smallest T that fits g_target
// workgroups, and that assignment is used.
// iters: per-system iteration counts of an earlier solve (or estimates).
static int batch_assign_oc(mf_batch_s *bt, const std::vector<int> &iters);

static int batch_assign(mf_batch_s *bt, int g_target, const std::vector<int> &iters)
{
    if (bt->oc) return batch_assign_oc(bt, iters);
    // work in microseconds of one workgroup: measured (wtime) when a solve
    // has timed every system, else the byte model at about 2.5 KB per us plus
    // ~3 us per iteration of scalar work and block reductions; a split
    // system's three group syncs cost ~9 us per iteration
    constexpr double RATE = 2.5e3, C_ITER = 3.0, C_SYNC = 9.0;
    const int n_sys = bt->nsys;
    const bool measured = (int)bt->wtime.size() == n_sys;
    std::vector<double> work(n_sys), its(n_sys);
    double total = 0.0;
    for (int s = 0; s < n_sys; s++) {
        its[s] = iters.empty() ? 1.0 : std::max(1, iters[s]);
        work[s] = measured ? std::max(bt->wtime[s], 1.0)
                           : its[s] * ((12.0 * bt->nnz[s] + 48.0 * bt->K * bt->tilem[s] * bt->B) / RATE + C_ITER);
        total += work[s];
    }
    g_target = std::max(1, g_target);
    const int gpb = bt->TPB / bt->B;
    std::vector<int> order(n_sys);
    for (int s = 0; s < n_sys; s++) order[s] = s;
    std::sort(order.begin(), order.end(), [&](int a, int c) { return work[a] > work[c] || (work[a] == work[c] && a < c); });
    // workgroups of system s under bound T (1 when it fits one workgroup)
    auto nwg = [&](int s, double T) {
        if (work[s] <= T) return 1;
        const double room = T - its[s] * C_SYNC;
        const int cap = std::max(1, bt->tilem[s] / gpb);
        if (room <= 0.0) return cap;
        return std::min(cap, std::max(2, (int)std::ceil(work[s] / room)));
    };
    // assignment under T: per-system workgroups, then the other systems
    // packed largest first onto the least loaded workgroup, the workgroups of
    // split systems included (there a packed system runs after the split one,
    // so the group is in step during the split system's syncs); returns the
    // largest workgroup load and the number of workgroups used
    std::vector<int> w_of(n_sys), pack_of(n_sys);
    auto plan = [&](double T) {
        int gbig = 0, nsmall = 0;
        for (int s = 0; s < n_sys; s++) {
            w_of[s] = nwg(s, T);
            if (w_of[s] > 1) gbig += w_of[s];
            else nsmall++;
        }
        if (gbig > g_target) return std::make_pair(1e300, gbig);
        const int G = std::min(g_target, gbig + nsmall);
        std::vector<double> load(G, 0.0);
        int g = 0;
        for (int s : order)
            if (w_of[s] > 1)
                for (int q = 0; q < w_of[s]; q++, g++) load[g] = work[s] / w_of[s] + its[s] * C_SYNC;
        double worst = 0.0;
        for (int q = 0; q < G; q++) worst = std::max(worst, load[q]);
        for (int s : order) {
            if (w_of[s] > 1) continue;
            int best = 0;
            for (int q = 1; q < G; q++)
                if (load[q] < load[best]) best = q;
            load[best] += work[s];
            pack_of[s] = best;
            worst = std::max(worst, load[best]);
        }
        return std::make_pair(G ? worst : 1e300, G);
    };
    double lo = total / g_target, hi = std::max(total, lo) * 2.0;
    for (int it = 0; it < 40; it++) {
        const double mid = 0.5 * (lo + hi);
        if (plan(mid).first <= mid) hi = mid;
        else lo = mid;
    }
    const int G = plan(hi).second;
    bt->sys_nwg.assign(w_of.begin(), w_of.end());
    std::vector<std::vector<int>> tasks(G);  // per workgroup: task records (5 ints)
    int g = 0;
    for (int s : order) {
        const int w = w_of[s];
        if (w == 1) continue;
        for (int q = 0; q < w; q++, g++)
            tasks[g] = {s, bt->trbase[s] + (int)((long long)bt->tilem[s] * q / w),
                        bt->trbase[s] + (int)((long long)bt->tilem[s] * (q + 1) / w), g - q, w};
    }
    for (int s : order) {
        if (w_of[s] > 1) continue;
        const int wg = pack_of[s];
        tasks[wg].insert(tasks[wg].end(), {s, bt->trbase[s], bt->trbase[s] + bt->tilem[s], wg, 1});
    }
    // GroupSync restarts its epoch per task on lines cleared once per launch:
    // a multi-workgroup task must be its workgroups' first and only one
    for (int q = 0; q < G; q++)
        for (size_t t = 5; t < tasks[q].size(); t += 5)
            if (tasks[q][t + 4] > 1) return fail(MF_E_UNSUPPORTED, "batch plan: a split system after another task");
    std::vector<int> h(G + 1, 0);
    for (int q = 0; q < G; q++) h[q + 1] = h[q] + (int)tasks[q].size() / 5;
    for (int q = 0; q < G; q++) h.insert(h.end(), tasks[q].begin(), tasks[q].end());
    if (bt->d_wg) MF_HIP(hipFree(bt->d_wg));
    if (bt->d_slots) MF_HIP(hipFree(bt->d_slots));
    bt->d_wg = nullptr;
    bt->d_slots = nullptr;
    MF_HIP(hipMalloc((void **)&bt->d_wg, sizeof(int) * h.size()));
    MF_HIP(hipMemcpy(bt->d_wg, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    MF_HIP(hipMalloc((void **)&bt->d_slots, sizeof(unsigned long long) * 2 * 2 * BK_MAXV * (size_t)G));
    bt->G = G;
    bt->nlines = G;
    bt->TPB = 256;
    return MF_OK;
}

// On-chip assignment: system s needs W_s = ceil(tile rows / (OC_RPT x
// OC_TPB / B)) workgroups (its vectors then fit their registers and LDS),
// all running it together for t_s.  List scheduling on the resident grid (one
// workgroup per CU): systems largest gang first (then longest), each on the
// W_s workgroups free earliest, starting when the last of them is free.  A
// workgroup's tasks are in start order, and every member of a gang has the
// same start, so the earliest unfinished task can always run (no cycle of
// waits).  t_s: measured on an earlier solve (wtime), else iterations x
// (bytes of the matrix and the published vector per workgroup at ~10 GB/s
// per CU, plus ~8 us of syncs and scalar work).
static int batch_assign_oc(mf_batch_s *bt, const std::vector<int> &iters)
{
    const int n_sys = bt->nsys, G = device_cu_count();
    const int trw = OC_RPT * (OC_TPB / bt->B);
    const bool measured = (int)bt->wtime.size() == n_sys;
    std::vector<int> W(n_sys);
    std::vector<double> t(n_sys);
    for (int s = 0; s < n_sys; s++) {
        W[s] = (bt->tilem[s] + trw - 1) / trw;
        if (W[s] > G) return fail(MF_E_UNSUPPORTED, "system larger than the device holds on chip");
        const double its = iters.empty() ? 1.0 : std::max(1, iters[s]);
        const double bytes = (12.0 * bt->nnz[s] + 2.0 * 64.0 * bt->tilem[s] * bt->B) / W[s];
        t[s] = measured ? std::max(bt->wtime[s], 1.0) : its * (bytes / 1.0e4 + 8.0);
    }
    std::vector<int> order(n_sys);
    for (int s = 0; s < n_sys; s++) order[s] = s;
    std::sort(order.begin(), order.end(), [&](int a, int c) {
        return W[a] != W[c] ? W[a] > W[c] : (t[a] != t[c] ? t[a] > t[c] : a < c);
    });
    std::vector<double> avail(G, 0.0);
    std::vector<std::vector<int>> tasks(G);  // per workgroup: 6-int records
    std::vector<int> idx(G);
    int line = 0;
    for (int s : order) {
        for (int g = 0; g < G; g++) idx[g] = g;
        std::partial_sort(idx.begin(), idx.begin() + W[s], idx.end(),
                          [&](int a, int c) { return avail[a] < avail[c] || (avail[a] == avail[c] && a < c); });
        double start = 0.0;
        for (int q = 0; q < W[s]; q++) start = std::max(start, avail[idx[q]]);
        for (int q = 0; q < W[s]; q++) {
            const int g = idx[q];
            const int a = bt->trbase[s] + (int)((long long)bt->tilem[s] * q / W[s]);
            const int b = bt->trbase[s] + (int)((long long)bt->tilem[s] * (q + 1) / W[s]);
            tasks[g].insert(tasks[g].end(), {s, a, b, line, W[s], q});
            avail[g] = start + t[s];
        }
        line += W[s];
    }
    bt->sys_nwg.assign(W.begin(), W.end());
    std::vector<int> h(G + 1, 0);
    for (int q = 0; q < G; q++) h[q + 1] = h[q] + (int)tasks[q].size() / 6;
    for (int q = 0; q < G; q++) h.insert(h.end(), tasks[q].begin(), tasks[q].end());
    if (bt->d_wg) MF_HIP(hipFree(bt->d_wg));
    if (bt->d_slots) MF_HIP(hipFree(bt->d_slots));
    bt->d_wg = nullptr;
    bt->d_slots = nullptr;
    MF_HIP(hipMalloc((void **)&bt->d_wg, sizeof(int) * h.size()));
    MF_HIP(hipMemcpy(bt->d_wg, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    MF_HIP(hipMalloc((void **)&bt->d_slots, sizeof(unsigned long long) * 2 * 2 * BK_MAXV * (size_t)line));
    bt->G = G;
    bt->nlines = line;
    bt->TPB = OC_TPB;
    return MF_OK;
}

// Builds the launch arguments and captures reset + persistent launch in a
// graph (the reset is a kernel: no memset nodes).
static int batch_capture(mf_batch_s *bt)
{
    hipStream_t st = stream();
    BatchArgs &A = bt->args;
    const BatchLayout &Li = bt->mixed ? bt->Lmix : bt->L64;
    A.L = DevL{bt->d_tile_ptr, Li.trh, Li.stc, Li.blob};
    A.L64 = DevL{bt->d_tile_ptr, bt->L64.trh, bt->L64.stc, bt->L64.blob};
    A.G = bt->G;
    A.nlines = bt->nlines;
    A.nrhs = bt->nrhs;
    A.wg_task = bt->d_wg;
    A.task = bt->d_wg + bt->G + 1;
    const size_t vn = bt->rows * bt->K;
    A.b = bt->d_vec;
    A.x = bt->d_vec + vn;
    A.r = bt->d_vec + 2 * vn;
    A.p = bt->d_vec + 3 * vn;
    A.s = bt->d_vec + 4 * vn;
    A.r32 = bt->d_r32;
    A.slots = bt->d_slots;
    A.ctl = bt->d_ctl;
    A.sys_out = bt->d_sys_out;
    A.sys_relres = bt->d_relres;
    A.tol = bt->o.tol;
    // an inner layout without low-precision tiles is the FP64 operator with
    // FP64 arithmetic: replacements only at convergence, as in FP64 mode
    A.rr_factor = Li.haslow ? bt->o.rr_factor : 0.0;
    A.maxit = bt->o.maxit;
    A.haslow = Li.haslow ? 1 : 0;
    A.spin_limit = bt->o.spin_limit;
    A.prof = nullptr;
    if (std::getenv("MF_PHASE_PROFILE")) {
        if (bt->d_prof) (void)hipFree(bt->d_prof);
        bt->d_prof = nullptr;
        MF_HIP(hipMalloc((void **)&bt->d_prof, sizeof(unsigned long long) * 4 * bt->G));
        MF_HIP(hipMemset(bt->d_prof, 0, sizeof(unsigned long long) * 4 * bt->G));
        A.prof = bt->d_prof;
    }
    bt->fn = bt->oc ? pick_batched_oc(bt->B, bt->iw) : pick_batched(bt->B, bt->iw, bt->K);
    int per_cu = 0;
    MF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bt->fn, bt->TPB, 0));
    if ((long long)per_cu * device_cu_count() < bt->G)
        return fail(MF_E_UNSUPPORTED, "batched grid does not fit the device");
    if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    bt->gexec = nullptr;
    bt->graph = nullptr;
    const long long nslots = 2ll * 2 * BK_MAXV * bt->nlines;
    const long long nr = std::max<long long>(nslots, 4ll * bt->nsys);
    MF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    k_batch_reset<<<(unsigned)((nr + 255) / 256), 256, 0, st>>>(bt->d_slots, nslots, bt->d_ctl, bt->d_sys_out, bt->nsys);
    void *argv[] = {&bt->args};
    hipError_t le = hipLaunchKernel(bt->fn, dim3(bt->G), dim3(bt->TPB), argv, 0, st);
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (le != hipSuccess) return hip_fail(le, "batched launch (capture)");
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
    bt->graph = graph;
    MF_HIP(hipGraphInstantiate(&bt->gexec, graph, nullptr, nullptr, 0));
    return MF_OK;
}

static void batch_release_data(mf_batch_s *bt)
{
    bt->L64.release();
    bt->Lmix.release();
    void *ps[] = {bt->d_tile_ptr, bt->d_vec, bt->d_r32, bt->d_ctl, bt->d_sys_out, bt->d_relres};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (bt->h_stage) (void)hipHostFree(bt->h_stage);
    bt->d_tile_ptr = nullptr;
    bt->d_vec = nullptr;
    bt->d_r32 = nullptr;
    bt->d_ctl = bt->d_sys_out = nullptr;
    bt->d_relres = nullptr;
    bt->h_stage = nullptr;
}

// (Re)builds layouts and vectors for the current tile size and thresholds.
static int batch_build(mf_batch_s *bt)
{
    hipStream_t st = stream();
    batch_release_data(bt);
    std::vector<mf_matrix> mats(bt->nsys, nullptr);
    auto cleanup = [&]() {
        for (mf_matrix m : mats)
            if (m) mf_matrix_destroy(m);
    };
    int rc = MF_OK;
    // one index width for all systems: the request resolved against the
    // forms that every system fits
    int fits = 3;
    for (int s = 0; s < bt->nsys && !rc; s++) {
        rc = mf_matrix_create(bt->n[s], bt->n[s], bt->h_rp[s][bt->n[s]], bt->h_rp[s].data(), bt->h_ci[s].data(),
                              bt->h_va[s].data(), nullptr, bt->B, &mats[s]);
        int f = 0;
        if (!rc) rc = index_fits(mats[s], &f);
        fits &= f;
    }
    const int iw = resolve_index_width(bt->o.index_width, fits);
    for (int s = 0; s < bt->nsys && !rc; s++) {
        rc = build_layout(mats[s], mats[s]->L64, 0.0, 0, iw, false);
        if (!rc && bt->mixed) rc = build_layout(mats[s], mats[s]->Lmix, bt->tau[s], bt->o.allow_mask, iw, false);
        if (!rc && (mats[s]->L64.iw != iw || (bt->mixed && mats[s]->Lmix.iw != iw)))
            rc = fail(MF_E_FORMAT, "slab index width differs between systems");
    }
    if (rc) { cleanup(); return rc; }
    bt->iw = iw;
    bt->tilem.resize(bt->nsys);
    bt->trbase.assign(bt->nsys + 1, 0);
    bt->tbase.assign(bt->nsys + 1, 0);
    bt->nnz.resize(bt->nsys);
    for (int s = 0; s < bt->nsys; s++) {
        bt->tilem[s] = mats[s]->tilem;
        bt->nnz[s] = mats[s]->nnz;
        bt->trbase[s + 1] = bt->trbase[s] + mats[s]->tilem;
        bt->tbase[s + 1] = bt->tbase[s] + mats[s]->tilenum;
    }
    bt->tilem_total = bt->trbase[bt->nsys];
    bt->tilenum_total = bt->tbase[bt->nsys];
    if ((long long)bt->tilem_total * bt->B * bt->K * 8 >= (1ll << 32) / 2) {
        cleanup();
        return fail(MF_E_UNSUPPORTED, "batched vectors of 2 GB or more");
    }
    bt->rows = (size_t)bt->tilem_total * bt->B;
    rc = hipMalloc((void **)&bt->d_tile_ptr, sizeof(int) * ((size_t)bt->tilem_total + 1)) == hipSuccess ? MF_OK
                                                                                                  : fail(MF_E_NOMEM, "tile_ptr");
    for (int s = 0; s < bt->nsys && !rc; s++)
        k_add_int<<<(mats[s]->tilem + 1 + 255) / 256, 256, 0, st>>>(mats[s]->d_tile_ptr, mats[s]->tilem + 1,
                                                                     (int)bt->tbase[s], bt->d_tile_ptr + bt->trbase[s]);
    if (!rc) rc = concat_layouts(bt, mats, false, bt->L64);
    if (!rc && bt->mixed) rc = concat_layouts(bt, mats, true, bt->Lmix);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(MF_E_HIP, "batched layout copy");
    cleanup();
    if (rc) return rc;
    // algorithmic bytes per iteration: the matrix once (tile formula, each
    // system as in the single-matrix case) plus 16 n bytes per column
    const BatchLayout &Li = bt->mixed ? bt->Lmix : bt->L64;
    const double wv[5] = {8, 0, 1, 2, 4};
    double vt = 0.0;
    for (int c = 0; c < 5; c++) vt += Li.nnz_code[c] * (wv[c] + 1);
    bt->bytes_alg = vt + (double)bt->tilenum_total * (8 + bt->B) + 4.0 * (bt->tilem_total + bt->nsys) +
                    16.0 * bt->K * (double)bt->rows;
    const size_t vn = bt->rows * bt->K;
    MF_HIP(hipMalloc((void **)&bt->d_vec, sizeof(double) * 5 * vn));
    MF_HIP(hipMalloc((void **)&bt->d_r32, sizeof(float) * vn));
    MF_HIP(hipMemset(bt->d_vec, 0, sizeof(double) * 5 * vn));
    MF_HIP(hipMemset(bt->d_r32, 0, sizeof(float) * vn));
    MF_HIP(hipMalloc((void **)&bt->d_ctl, sizeof(int) * 8));
    MF_HIP(hipMalloc((void **)&bt->d_sys_out, sizeof(int) * 4 * bt->nsys));
    MF_HIP(hipMalloc((void **)&bt->d_relres, sizeof(double) * bt->nsys));
    MF_HIP(hipHostMalloc((void **)&bt->h_stage, sizeof(double) * vn, hipHostMallocDefault));
    return MF_OK;
}

// Workgroups of the batched kernel that can be resident at once.
static int batch_capacity(mf_batch_s *bt)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pick_batched(bt->B, bt->iw, bt->K), bt->TPB, 0) !=
        hipSuccess)
        per_cu = 1;
    return std::max(1, per_cu) * device_cu_count();
}

// grid target: opts.grid when set, else the resident capacity (never more)
static int batch_target(mf_batch_s *bt)
{
    const int cap = batch_capacity(bt);
    return bt->o.grid > 0 ? std::min(bt->o.grid, cap) : cap;
}

static int batch_run(mf_batch_s *bt, float *ms)
{
    hipStream_t st = stream();
    hipEvent_t e0, e1;
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    MF_HIP(hipEventRecord(e0, st));
    MF_HIP(hipGraphLaunch(bt->gexec, st));
    MF_HIP(hipEventRecord(e1, st));
    MF_HIP(hipEventSynchronize(e1));
    MF_HIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MF_OK;
}

extern "C" {

int mf_batch_create(int nsys, const int *n, const int *const *rowptr, const int *const *colidx,
                    const double *const *val, int nrhs, const mf_opts *opts, mf_batch *out)
{
    clear_error();
    *out = nullptr;
    if (nsys <= 0 || !n || !rowptr || !colidx || !val) return fail(MF_E_ARG, "empty batch");
    if (nrhs < 1 || nrhs > 8) return fail(MF_E_ARG, "nrhs must be 1..8");
    if (opts && (opts->index_width < 0 || opts->index_width == 3 || opts->index_width > 4))
        return fail(MF_E_ARG, "index width must be 0, 1, 2 or 4");
    if (opts && (opts->vectors < MF_VECTORS_AUTO || opts->vectors > MF_VECTORS_ONCHIP))
        return fail(MF_E_ARG, "vectors");
    int rc = ensure_device();
    if (rc) return rc;
    mf_batch_s *bt = new mf_batch_s();
    if (opts) bt->o = *opts;
    else mf_opts_default(&bt->o);
    if (bt->o.method != MF_CG) {
        delete bt;
        return fail(MF_E_UNSUPPORTED, "the batched solver runs CG");
    }
    const double t0 = now_ms();
    bt->nsys = nsys;
    bt->nrhs = nrhs;
    bt->K = nrhs == 1 ? 1 : 8;
    bt->mixed = bt->o.precision == MF_MIXED;
    bt->n.assign(n, n + nsys);
    bt->tau.assign(nsys, bt->o.tau);
    bt->B = bt->o.block == 8 || bt->o.block == 32 ? bt->o.block : 16;
    bt->h_rp.resize(nsys);
    bt->h_ci.resize(nsys);
    bt->h_va.resize(nsys);
    bt->weights.resize(nsys);
    for (int s = 0; s < nsys && !rc; s++) {
        if (n[s] <= 0 || !rowptr[s]) rc = fail(MF_E_ARG, "empty system");
        else {
            const int nz = rowptr[s][n[s]];
            bt->h_rp[s].assign(rowptr[s], rowptr[s] + n[s] + 1);
            bt->h_ci[s].assign(colidx[s], colidx[s] + nz);
            bt->h_va[s].assign(val[s], val[s] + nz);
            // CG iterations grow like the square root of the condition number,
            // about n^(1/d) for d-dimensional stencils: sqrt(n) as a first
            // guess until a solve has measured them
            bt->weights[s] = (int)std::lround(std::sqrt((double)n[s]));
        }
    }
    if (!rc) rc = batch_build(bt);
    if (!rc) {
        // the on-chip kernel when asked for: the 8-column form (2 to 8
        // right-hand sides), every system's gang within the resident grid
        if (bt->o.vectors == MF_VECTORS_ONCHIP) {
            const int trw = OC_RPT * (OC_TPB / bt->B);
            bool fit = bt->K == OC_K;
            for (int s = 0; s < nsys && fit; s++) fit = (bt->tilem[s] + trw - 1) / trw <= device_cu_count();
            if (!fit) rc = fail(MF_E_UNSUPPORTED, "on-chip batch: needs 2 to 8 right-hand sides and systems that fit the grid");
            bt->oc = fit;
        }
    }
    if (!rc) rc = batch_assign(bt, batch_target(bt), bt->weights);
    if (!rc) rc = batch_capture(bt);
    if (rc) {
        batch_release(bt);
        delete bt;
        return rc;
    }
    bt->ms_setup = now_ms() - t0;
    live_add(bt, 1);
    *out = bt;
    return MF_OK;
}

void mf_batch_destroy(mf_batch bt)
{
    if (!bt || !live_remove(bt, 1)) return;
    batch_release(bt);
    delete bt;
}

}  // extern "C"

// Right-hand sides into the device layout ([row][K], padded rows zero).
static int batch_stage(mf_batch_s *bt, const double *const *b)
{
    hipStream_t st = stream();
    const int K = bt->K;
    const size_t vn = bt->rows * K;
    std::memset(bt->h_stage, 0, sizeof(double) * vn);
    for (int s = 0; s < bt->nsys; s++) {
        const size_t r0 = (size_t)bt->trbase[s] * bt->B;
        for (int k = 0; k < bt->nrhs; k++)
            for (int i = 0; i < bt->n[s]; i++) bt->h_stage[(r0 + i) * K + k] = b[s][(size_t)k * bt->n[s] + i];
    }
    MF_HIP(hipMemcpyAsync(bt->d_vec, bt->h_stage, sizeof(double) * vn, hipMemcpyHostToDevice, st));
    MF_HIP(hipStreamSynchronize(st));
    return MF_OK;
}

struct BatchOutcome {
    int status = MF_OK, itmax = 0, repl = 0, fell = 0;
    double worst = 0.0;
    std::vector<int> iters;
    std::vector<double> relres;
    std::vector<double> us;  // per system: time on its first workgroup
};

// Work of every system on one workgroup, from a solve's per-system times: a
// system split over w workgroups took (work / w + iterations x sync) there.
static void batch_measure(mf_batch_s *bt, const BatchOutcome &oc)
{
    constexpr double C_SYNC = 9.0;
    bt->wtime.assign(bt->nsys, 0.0);
    for (int s = 0; s < bt->nsys; s++) {
        const int w = s < (int)bt->sys_nwg.size() ? bt->sys_nwg[s] : 1;
        const double t = oc.us[s];
        // on chip: a gang's time as it ran
        bt->wtime[s] = bt->oc ? t : w > 1 ? std::max(t - oc.iters[s] * C_SYNC, 0.5 * t) * w : t;
    }
}

static int batch_outcome(mf_batch_s *bt, BatchOutcome &oc)
{
    std::vector<int> so(4 * (size_t)bt->nsys);
    oc.relres.assign(bt->nsys, 0.0);
    oc.iters.assign(bt->nsys, 0);
    oc.us.assign(bt->nsys, 0.0);
    int ab = 0;
    MF_HIP(hipMemcpy(so.data(), bt->d_sys_out, sizeof(int) * so.size(), hipMemcpyDeviceToHost));
    MF_HIP(hipMemcpy(oc.relres.data(), bt->d_relres, sizeof(double) * bt->nsys, hipMemcpyDeviceToHost));
    MF_HIP(hipMemcpy(&ab, bt->d_ctl, sizeof(int), hipMemcpyDeviceToHost));
    oc.status = ab ? MF_E_TIMEOUT : MF_OK;
    oc.itmax = oc.repl = oc.fell = 0;
    oc.worst = 0.0;
    for (int s = 0; s < bt->nsys; s++) {
        oc.iters[s] = so[4 * s + 1];
        oc.us[s] = so[4 * s + 3];
        oc.itmax = std::max(oc.itmax, so[4 * s + 1]);
        oc.repl += so[4 * s + 2] & 0x3fffffff;
        oc.fell += so[4 * s + 2] >> 30;
        oc.worst = std::max(oc.worst, oc.relres[s]);
        if (oc.status == MF_OK && so[4 * s] != MF_OK) oc.status = so[4 * s];
    }
    return MF_OK;
}

static void batch_report(mf_batch_s *bt, const BatchOutcome &oc, float ms, mf_report *rep)
{
    if (!rep) return;
    std::memset(rep, 0, sizeof(*rep));
    rep->status = oc.status;
    rep->iterations = oc.itmax;
    rep->outer = 1;
    rep->replacements = oc.repl;
    rep->fallback = oc.fell;  // systems that moved to the FP64 layout
    rep->relres_true = oc.worst;
    rep->ms_solve = ms;
    rep->ms_setup = bt->ms_setup;
    rep->grid = bt->G;
    rep->block = bt->TPB;
    rep->rows_per_thread = 0;
}

extern "C" {

int mf_batch_solve(mf_batch bt, const double *const *b, double *const *x, int *iters, double *relres, mf_report *rep)
{
    clear_error();
    if (!bt || !b || !x) return fail(MF_E_ARG, "null argument");
    int rc = batch_stage(bt, b);
    float ms = 0.0f;
    if (!rc) rc = batch_run(bt, &ms);
    BatchOutcome oc;
    if (!rc) rc = batch_outcome(bt, oc);
    if (rc) return rc;
    const int K = bt->K;
    const size_t vn = bt->rows * K;
    MF_HIP(hipMemcpy(bt->h_stage, bt->args.x, sizeof(double) * vn, hipMemcpyDeviceToHost));
    for (int s = 0; s < bt->nsys; s++) {
        const size_t r0 = (size_t)bt->trbase[s] * bt->B;
        for (int k = 0; k < bt->nrhs; k++)
            for (int i = 0; i < bt->n[s]; i++) x[s][(size_t)k * bt->n[s] + i] = bt->h_stage[(r0 + i) * K + k];
        if (iters) iters[s] = oc.iters[s];
        if (relres) relres[s] = oc.relres[s];
    }
    batch_report(bt, oc, ms, rep);
    bt->fell_last = oc.fell;
    if (bt->args.prof) {
        std::vector<unsigned long long> pf(4 * (size_t)bt->G);
        MF_HIP(hipMemcpy(pf.data(), bt->args.prof, sizeof(unsigned long long) * pf.size(), hipMemcpyDeviceToHost));
        double sum[4] = {0, 0, 0, 0}, tmax = 0;
        int gmax = 0;
        for (int g = 0; g < bt->G; g++) {
            double t = 0;
            for (int k = 0; k < 4; k++) {
                sum[k] += pf[4 * g + k] / 100.0;
                t += pf[4 * g + k] / 100.0;
            }
            if (t > tmax) { tmax = t; gmax = g; }
        }
        std::fprintf(stderr, "[mf batch phase] mean us per workgroup: spmv %.0f update %.0f sync %.0f pass %.0f | "
                     "busiest workgroup %d: spmv %.0f update %.0f sync %.0f pass %.0f\n",
                     sum[0] / bt->G, sum[1] / bt->G, sum[2] / bt->G, sum[3] / bt->G, gmax, pf[4 * gmax] / 100.0,
                     pf[4 * gmax + 1] / 100.0, pf[4 * gmax + 2] / 100.0, pf[4 * gmax + 3] / 100.0);
        // MF_PHASE_PROFILE=2: every workgroup's time and tasks (system, rows,
        // iterations, group size, measured time of the system)
        if (std::atoi(std::getenv("MF_PHASE_PROFILE")) == 2) {
            std::vector<int> h(bt->G + 1);
            MF_HIP(hipMemcpy(h.data(), bt->d_wg, sizeof(int) * (bt->G + 1), hipMemcpyDeviceToHost));
            const int TS = bt->oc ? 6 : 5;  // ints per task record
            std::vector<int> tk(TS * (size_t)h[bt->G]);
            MF_HIP(hipMemcpy(tk.data(), bt->d_wg + bt->G + 1, sizeof(int) * tk.size(), hipMemcpyDeviceToHost));
            for (int g = 0; g < bt->G; g++) {
                double t = 0;
                for (int k = 0; k < 4; k++) t += pf[4 * g + k] / 100.0;
                std::fprintf(stderr, "[mf batch wg] %d total %.0f spmv %.0f update %.0f sync %.0f |", g, t,
                             pf[4 * g] / 100.0, pf[4 * g + 1] / 100.0, pf[4 * g + 2] / 100.0);
                for (int q = h[g]; q < h[g + 1]; q++) {
                    const int *r = tk.data() + TS * (size_t)q;
                    std::fprintf(stderr, " s%d n=%d it=%d w=%d us=%.0f;", r[0], bt->n[r[0]], oc.iters[r[0]], r[4],
                                 oc.us[r[0]]);
                }
                std::fprintf(stderr, "\n");
            }
        }
        MF_HIP(hipMemset(bt->args.prof, 0, sizeof(unsigned long long) * pf.size()));
    }
    return oc.status == MF_E_TIMEOUT ? fail(MF_E_TIMEOUT, "group sync timeout") : MF_OK;
}

// Autotuner.  Trial solves of the caller's right-hand sides, each timed with
// device events.  Every configuration is solved twice: on the current
// assignment, and again after the workgroups were re-assigned from the time
// every system took in the first solve (the kernel times each system on its
// first workgroup), so systems that need many iterations or run slowly get
// more workgroups and the packed workgroups finish together.  Configurations:
//  1. the current one;
//  2. tile sizes 8, 16 and 32 (layouts rebuilt on the GPU);
//  3. when the caller asked for mixed precision: tile thresholds tau = 2^-24
//     (FP32-rounded tiles) and 2^-11 (FP16-rounded), kept only if every
//     column still reaches the tolerance.
// The fastest converged configuration stays captured in the batch's graph,
// assigned from its own measured times.
int mf_batch_tune(mf_batch bt, const double *const *b, mf_report *rep)
{
    clear_error();
    if (!bt || !b) return fail(MF_E_ARG, "null argument");
    struct Cfg {
        int B;
        bool mixed;
        double tau;
    };
    auto apply = [&](const Cfg &c, bool rebuild) -> int {
        bt->B = c.B;
        bt->mixed = c.mixed;
        bt->tau.assign(bt->nsys, c.tau);
        int rc = rebuild ? batch_build(bt) : MF_OK;
        if (!rc) rc = batch_assign(bt, batch_target(bt), bt->weights);
        if (!rc) rc = batch_capture(bt);
        return rc;
    };
    auto trial = [&](float *ms, BatchOutcome &oc) -> int {
        int rc = batch_stage(bt, b);
        if (!rc) rc = batch_run(bt, ms);
        if (!rc) rc = batch_outcome(bt, oc);
        return rc;
    };
    // a configuration is tried twice: once on the current assignment, then
    // re-assigned from the per-system times that solve measured
    auto measured_trial = [&](float *ms, BatchOutcome &oc) -> int {
        int rc = trial(ms, oc);
        if (rc || oc.status != MF_OK) return rc;
        bt->weights = oc.iters;
        batch_measure(bt, oc);
        rc = batch_assign(bt, batch_target(bt), bt->weights);
        if (!rc) rc = batch_capture(bt);
        if (!rc) rc = trial(ms, oc);
        return rc;
    };
    const bool want_mixed = bt->o.precision == MF_MIXED;
    Cfg cur{bt->B, bt->mixed, bt->tau.empty() ? 0.0 : bt->tau[0]};
    float ms = 0.0f;
    BatchOutcome oc;
    int rc = measured_trial(&ms, oc);
    if (rc) return rc;
    if (oc.status != MF_OK) return fail(oc.status, "tuning trial did not converge");
    Cfg best = cur;
    float best_ms = ms;
    std::vector<double> best_wtime = bt->wtime;
    std::vector<int> best_weights = bt->weights;
    char buf[256];
    std::string log;
    std::snprintf(buf, sizeof buf, "B=%d %s: %.2f ms", cur.B, cur.mixed ? "mixed" : "fp64", ms);
    log += buf;
    auto consider = [&](const Cfg &c, const char *label) -> int {
        int rc = apply(c, true);
        if (!rc) rc = measured_trial(&ms, oc);
        if (rc) return rc;
        std::snprintf(buf, sizeof buf, "; %s: %.2f ms%s", label, ms, oc.status ? " (not converged)" : "");
        log += buf;
        if (oc.status == MF_OK && oc.worst <= bt->o.tol && ms < best_ms) {
            best = c;
            best_ms = ms;
            best_wtime = bt->wtime;
            best_weights = bt->weights;
        }
        return MF_OK;
    };
    for (int B : {8, 16, 32}) {
        if (B == cur.B) continue;
        char label[32];
        std::snprintf(label, sizeof label, "B=%d", B);
        rc = consider({B, cur.mixed, cur.tau}, label);
        if (rc) return rc;
    }
    if (want_mixed) {
        for (double tau : {std::ldexp(1.0, -24), std::ldexp(1.0, -11)}) {
            char label[32];
            std::snprintf(label, sizeof label, "tau=2^%d", (int)std::lround(std::log2(tau)));
            rc = consider({best.B, true, tau}, label);
            if (rc) return rc;
        }
    }
    // the best configuration again, assigned from its own measured times
    bt->wtime = best_wtime;
    bt->weights = best_weights;
    rc = apply(best, true);
    if (!rc) rc = trial(&ms, oc);
    if (rc) return rc;
    std::snprintf(buf, sizeof buf, " -> B=%d %s tau=%g, %d workgroups", best.B, best.mixed ? "mixed" : "fp64",
                  best.tau, bt->G);
    bt->tuned = log + buf;
    batch_report(bt, oc, ms, rep);
    return MF_OK;
}

int mf_batch_info(mf_batch bt, mf_batch_info_t *info)
{
    if (!bt || !info) return fail(MF_E_ARG, "null argument");
    info->nsys = bt->nsys;
    info->nrhs = bt->nrhs;
    info->B = bt->B;
    info->G = bt->G;
    info->rows = (long long)bt->rows;
    info->tilenum = bt->tilenum_total;
    info->bytes_alg = bt->bytes_alg;
    info->ms_setup = bt->ms_setup;
    const BatchLayout &Li = bt->mixed ? bt->Lmix : bt->L64;
    for (int c = 0; c < 5; c++) info->nnz_per_code[c] = Li.nnz_code[c];
    info->mixed = bt->mixed ? 1 : 0;
    info->tau = bt->tau.empty() ? 0.0 : bt->tau[0];
    std::snprintf(info->tuned, sizeof info->tuned, "%s", bt->tuned.c_str());
    info->index_width = bt->iw;
    info->on_chip = bt->oc ? 1 : 0;
    info->fell_back = bt->fell_last;
    return MF_OK;
}

int mf_solve_batched(int nsys, const int *n, const int *const *rowptr, const int *const *colidx,
                     const double *const *val, int nrhs, const double *const *b, double *const *x, const mf_opts *o,
                     int *iters, double *relres, mf_report *rep)
{
    mf_batch bt = nullptr;
    int rc = mf_batch_create(nsys, n, rowptr, colidx, val, nrhs, o, &bt);
    if (rc) return rc;
    rc = mf_batch_solve(bt, b, x, iters, relres, rep);
    mf_batch_destroy(bt);
    return rc;
}

}  // extern "C"
