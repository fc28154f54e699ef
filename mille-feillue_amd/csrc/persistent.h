// Persistent CG / BiCGSTAB kernels (one workgroup per CU, vectors on chip,
// self-validating grid sync words, neighbour waits where a sync carries no
// reduction) and their launch selection, shared by the
// translation units that instantiate them for the tile sizes 16 (solver.hip),
// 8 (solver_b8.hip) and 32 (solver_b32.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mf_matrix.h"

namespace mf {

// --------------------------------------------------------- solver kernels
constexpr int SYNC_K = 5;  // values per grid allreduce, at most
constexpr int LDS_GATHER = 2048;  // rows of a one-workgroup solve gathered from LDS
// A true-residual pass on a lossy layout whose squared norm exceeds the best
// earlier one by this factor (the norm grew 4x while the recurrence claimed
// progress) ends the launch with MF_E_DIVERGED; the host then solves on the
// FP64 layout.
constexpr double DIVERGE_GROWTH = 16.0;

struct SolveArgs {
    DevL L;
    unsigned long long blob_bytes;  // bytes of L's slab blob
    DevL L64;          // FP64 layout for the true-residual passes of a lossy layout
    int use64;         // 1: passes use L64 (headers read from L64), else L
    int tilem, G, trw;
    const double *b;
    double *x;
    double *g64;       // published Krylov vector (FP64)
    float *g32;        // its FP32 copy for low-precision tiles
    double *h64, *q64; // publishing slots of the true-residual pass
    unsigned long long *slots;  // sync words [SLOT_BUFS][G][SLOT_WORDS]
    int *ctl;          // 0 abort, 1 status, 2 iterations, 3 replacements
    double *out;       // 0 ||b||, 1 final ||r_true||^2, 2 final updated ||r||^2
    double tol, rr_factor;
    int maxit, has64, haslow;
    long long spin_limit;
    unsigned long long *phase;  // optional: [0] spmv, [1] reduce after spmv,
                                // [2] vector update, [3] reduce after update,
                                // [4] true-residual passes (100 MHz ticks, block 0)
    // neighbour syncs (nbr = 1): chunk c's SpMV gathers rows of the chunks
    // dep_idx[dep_ptr[c] .. dep_ptr[c + 1]) (at most 64) besides its own;
    // flags[16 c] is chunk c's publish epoch
    const int *dep_ptr, *dep_idx;
    unsigned *flags;
    int nbr;
    // Krylov vectors of the streamed kernels (systems beyond register/LDS
    // residency): vector k of the padded rows at vec + k * vstride
    double *vec;
    long long vstride;
};

// Phase clock of block 0 (diagnostic; enabled when A.phase is set).
struct PhaseClock {
    unsigned long long *acc;
    unsigned long long t;
    __device__ PhaseClock(unsigned long long *p) : acc(blockIdx.x == 0 && threadIdx.x == 0 ? p : nullptr), t(0)
    {
        if (acc) t = wall_clock64();
    }
    __device__ void lap(int k)
    {
        if (acc) {
            const unsigned long long now = wall_clock64();
            acc[k] += now - t;
            t = now;
        }
    }
};

constexpr unsigned PHASE_EP0 = 201;

// Stores of vector entries that other workgroups read after the next grid
// sync: agent-scope stores, complete (visible to every XCD) once the storing
// wave's memory counter has drained.
__device__ __forceinline__ void st_agent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(float *p, float v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sync words: per epoch buffer, each workgroup owns one 64-byte line of
// SLOT_WORDS 64-bit words; an empty word holds SLOT_EMPTY, a signalling-NaN
// pattern that no arithmetic produces (operations return quiet NaNs).
constexpr int SLOT_WORDS = 8, SLOT_BUFS = 3;
constexpr unsigned long long SLOT_EMPTY = 0x7ff0dead5ee00001ull;
static_assert(SYNC_K <= SLOT_WORDS, "sync line");

// Grid synchronisation and deterministic allreduce without read-modify-write
// atomics or cache write-back.  At epoch e every workgroup stores its N FP64
// partials into its line of buffer e mod 3; wave 0 of every workgroup polls
// the G x N words until none is empty, then sums the partials in workgroup
// order (lane-strided sums, then a fixed butterfly), so all workgroups obtain
// identical bits and take identical branches.  After the poll a workgroup
// empties its own line of buffer (e + 2) mod 3: that buffer was last read at
// epoch e - 1, which every workgroup has left once all epoch-e words are
// there, and is next filled at epoch e + 2.  Before its epoch e + 1 stores a
// workgroup waits for its own earlier stores (the emptying among them) to
// complete, so a reader that has seen a workgroup's epoch e + 1 words finds
// its epoch e + 2 line either empty or current.
// publish: vector entries written before the sync (with st_agent) are read
// by other workgroups after it -- each wave drains its stores before the
// workgroup barrier that precedes the sync words, and the readers invalidate
// their caches (agent-scope acquire) after the poll.  The wait is bounded:
// after spin_limit polls the workgroup raises the abort word and every
// workgroup leaves (returns false).
template <int TPB, int K>
struct Reducer {
    static_assert(K <= SYNC_K, "reducer width");
    double *s_red;  // [TPB/64][K]
    double *s_bc;   // [K]
    int *s_flag;
    unsigned ep = 0;

    template <int N>
    __device__ bool allreduce(const SolveArgs &A, double (&v)[N], bool publish)
    {
        static_assert(N >= 1 && N <= K, "reducer width");
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
        for (int k = 0; k < N; k++) {
            double s = wave_sum(v[k]);
            if (lane == 0) s_red[wv * K + k] = s;
        }
        if (publish) __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
        __syncthreads();
        ++ep;
        // diagnostic stamps of syncs PHASE_EP0 .. PHASE_EP0 + 3: arrival and
        // exit time of every workgroup
        const bool stamp = A.phase && tid == 0 && ep - PHASE_EP0 < 4u;
        unsigned long long *st = stamp ? A.phase + 16 + (size_t)(ep - PHASE_EP0) * 2 * A.G + 2 * blockIdx.x : nullptr;
        if (stamp) st[0] = wall_clock64();
        if (A.G == 1) {
            // one workgroup: the reduction stays in LDS; published entries
            // are read back through this CU's L1, which the acquire refreshes
            if (tid < 64) {
                if (lane < N) {
                    double sum = 0.0;
                    for (int w = 0; w < TPB / 64; w++) sum += s_red[w * K + lane];
                    s_bc[lane] = sum;
                }
                if (publish) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            if (stamp) st[1] = wall_clock64();
            __syncthreads();
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = s_bc[k];
            return true;
        }
        if (tid < 64) {
            const unsigned G = (unsigned)A.G;
            unsigned long long *W = A.slots + (size_t)(ep % SLOT_BUFS) * G * SLOT_WORDS;
            __builtin_amdgcn_s_waitcnt(0);  // the emptying of this line has completed
            if (lane < N) {
                double s = 0.0;
                for (int w = 0; w < TPB / 64; w++) s += s_red[w * K + lane];
                __hip_atomic_store(W + (size_t)blockIdx.x * SLOT_WORDS + lane,
                                   (unsigned long long)__double_as_longlong(s), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            double acc[N];
            long long spins = 0;
            int ab = 0;
            for (;;) {
                // one round: the abort word and every slot word of this lane
                // (PQ workgroups per lane at a time, clamped addresses) are
                // loaded before the first wait, so a round costs one memory
                // latency rather than one per workgroup the lane reads
                constexpr int PQ = 4;
                const int abw = ld_agent(A.ctl);
                bool ok = true;
#pragma unroll
                for (int k = 0; k < N; k++) acc[k] = 0.0;
                for (unsigned g0 = lane; g0 < G; g0 += 64 * PQ) {
                    unsigned long long w[PQ][N];
#pragma unroll
                    for (int q = 0; q < PQ; q++) {
                        const unsigned g = min(g0 + 64u * q, G - 1);
#pragma unroll
                        for (int k = 0; k < N; k++)
                            w[q][k] = __hip_atomic_load(W + (size_t)g * SLOT_WORDS + k, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                    }
#pragma unroll
                    for (int q = 0; q < PQ; q++) {
                        if (g0 + 64u * q >= G) continue;
#pragma unroll
                        for (int k = 0; k < N; k++) {
                            ok = ok && w[q][k] != SLOT_EMPTY;
                            acc[k] += __longlong_as_double((long long)w[q][k]);
                        }
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
            if (!ab) {
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const double s = wave_sum(acc[k]);
                    if (lane == 0) s_bc[k] = s;
                }
                if (lane < SLOT_WORDS)
                    __hip_atomic_store(A.slots + (size_t)((ep + 2) % SLOT_BUFS) * G * SLOT_WORDS +
                                           (size_t)blockIdx.x * SLOT_WORDS + lane,
                                       SLOT_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (publish) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            if (lane == 0) *s_flag = ab;
        }
        if (stamp) st[1] = wall_clock64();
        __syncthreads();
        if (*s_flag) return false;
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = s_bc[k];
        return true;
    }

    // Grid barrier after publishing vector entries.
    __device__ bool barrier(const SolveArgs &A)
    {
        double z[1] = {0.0};
        return allreduce(A, z, true);
    }

    // Publish without a grid rendezvous: after this workgroup's vector stores
    // have drained, its chunk's epoch word is raised, and the workgroup waits
    // only until the chunks its SpMV gathers from show the same epoch (then
    // invalidates its caches, as after a publishing allreduce).  No vector is
    // overwritten before the next allreduce, which every workgroup reaches
    // only after its SpMV, so readers never see a later value than the one
    // published here.  Bounded like allreduce.
    unsigned nep = 0;
    __device__ bool neighbours(const SolveArgs &A, int chunk)
    {
        const int tid = threadIdx.x, lane = tid & 63;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
        __syncthreads();
        ++nep;
        if (tid < 64) {
            if (lane == 0)
                __hip_atomic_store(A.flags + (size_t)chunk * 16, nep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int d0 = A.dep_ptr[chunk], nd = A.dep_ptr[chunk + 1] - d0;
            const unsigned *f = lane < nd ? A.flags + (size_t)A.dep_idx[d0 + lane] * 16 : nullptr;
            long long spins = 0;
            int ab = 0;
            for (;;) {
                // the flag and the abort word in one round
                const unsigned fv = f ? ld_agent(f) : nep;
                const int abw = ld_agent(A.ctl);
                const bool ok = fv >= nep;
                if (__all(ok)) break;
                if (__builtin_amdgcn_readfirstlane(abw)) { ab = 1; break; }
                if (++spins > A.spin_limit) {
                    if (lane == 0) __hip_atomic_store(A.ctl, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ab = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ab) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (lane == 0) *s_flag = ab;
        }
        __syncthreads();
        return !*s_flag;
    }

    // The sync after publishing the vector the next SpMV gathers.
    __device__ bool published(const SolveArgs &A, int chunk)
    {
        return A.nbr && A.G > 1 ? neighbours(A, chunk) : barrier(A);
    }
};

// tile rows whose loads are interleaved inside the persistent kernels' SpMV
// (four-byte indices take four times the index registers: two rows)
__host__ __device__ constexpr int spmv_rows(int rpt, int iw)
{
    return rpt < (iw < 4 ? 4 : 2) ? rpt : (iw < 4 ? 4 : 2);
}

__device__ __forceinline__ bool finite_all(double a, double b) { return isfinite(a) && isfinite(b); }

// Hides the value of v from the optimiser.  Applied to the thread's row base
// at the top of every solver iteration so that the 64-bit addresses of the
// RPT rows are recomputed where they are used instead of being hoisted out of
// the loop (which would keep RPT x arrays addresses live and spill).
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Slab bytes a resident kernel keeps in LDS per workgroup (0: none).  The
// LDS left beside the kernels' vectors, headers and gathered-vector copy at
// 512 threads and 1, 2 or 4 rows per thread.
constexpr int lds_slab_bytes(int rpt) { return rpt == 1 ? 112 * 1024 : rpt == 2 ? 96 * 1024 : rpt == 4 ? 64 * 1024 : 0; }

// Stages the headers of tile rows tr0 .. tr0 + TRW - 1 of a chunk (zeros
// from tr1 on) and their first tiles.  When the chunk's slabs fit in LSB
// bytes they are copied to s_slab and the staged offsets are rebased onto
// that copy (returns true).  A chunk's slabs are one byte range of the blob:
// slabs follow tile-row order, and class 0's header of every tile row holds
// the row's start, empty or not, so the range runs from tile row tr0's start
// to tile row tr1's (the blob's end after the last tile row).  Offsets are
// multiples of 16 bytes, which keeps the copy's quads aligned.  Every thread
// calls it; a __syncthreads must follow.
template <int TPB>
__device__ bool stage_chunk(const SolveArgs &A, int tr0, int tr1, int TRW, int4 *s_h, int *s_t0, uint8_t *s_slab,
                            int LSB)
{
    const int tid = threadIdx.x;
    size_t lo = 0;
    bool lds = false;
    if (LSB > 0 && tr0 < tr1) {
        lo = slab_off(A.L.trh[(size_t)tr0 * 4].x);
        const size_t hi = tr1 < A.tilem ? slab_off(A.L.trh[(size_t)tr1 * 4].x) : A.blob_bytes;
        lds = hi - lo <= (size_t)LSB;
        if (lds) {
            const uint4 *src = reinterpret_cast<const uint4 *>(A.L.blob + lo);
            uint4 *dst = reinterpret_cast<uint4 *>(s_slab);
            for (size_t i = tid; i < (hi - lo) / 16; i += TPB) dst[i] = src[i];
        }
    }
    const int d = (int)(lo / 8);  // in the headers' 8-byte units (even: bit 0 is the pattern flag)
    for (int q = tid; q <= TRW; q += TPB) {
        const int tr = tr0 + q;
        if (q < TRW && tr < tr1) {
            const int4 *h = reinterpret_cast<const int4 *>(A.L.trh + (size_t)tr * 4);
            int4 a = h[0], b = h[1];
            if (lds) {
                a.x -= d;
                a.z -= d;
                b.x -= d;
                b.z -= d;
            }
            s_h[2 * q] = a;
            s_h[2 * q + 1] = b;
        } else if (q < TRW) {  // beyond the chunk: no slabs
            s_h[2 * q] = s_h[2 * q + 1] = make_int4(0, 0, 0, 0);
        }
        s_t0[q] = A.L.tile_ptr[min(tr, A.tilem)];
    }
    return lds;
}

// CG (Hestenes-Stiefel with s = A p carried by recurrence): registers hold
// x and p of the thread's rows, LDS holds r and s.  Per iteration: [r
// published, (r, r) reduced] -> convergence test -> w = A r, s = w + beta s,
// p = r + beta p and (p, s) in the same pass -> reduce (p, s) -> x, r updates
// -> publish r, reduce (r, r).  The denominator (p, A p) is summed directly
// rather than recurred (Chronopoulos-Gear's (A r, r) - beta (r, r) / alpha
// form), which holds after a residual replacement that keeps the direction
// and, with lossy tiles, halves the iterations on rough coefficients.
template <int B, int IW, int TPB, int RPT, int CM = CM_ALL>
__global__ __launch_bounds__(TPB, 1) void k_cg(SolveArgs A)
{
    __shared__ double s_r[RPT * TPB];
    __shared__ double s_s[RPT * TPB];
    __shared__ double s_red[(TPB / 64) * 4];
    __shared__ double s_bc[4];
    __shared__ int s_flag;
    Reducer<TPB, 4> R{s_red, s_bc, &s_flag};
    constexpr int GPB = TPB / B, NR = spmv_rows(RPT, IW);
    const int tid = threadIdx.x;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    const int grp = tid / B, lane = tid % B, lanebase = (tid & 63) & ~(B - 1);
    const int tr0 = chunk * A.trw, tr1 = min(A.tilem, tr0 + A.trw);
    double xr[RPT], pr[RPT];
    // headers of the block's tile rows, loaded once: the per-iteration SpMV
    // then issues the tile-column word and the slab loads together
    constexpr int TRW = RPT * GPB;
    __shared__ int4 s_h[2 * TRW];
    __shared__ int s_t0[TRW + 1];
    // small chunks keep their slabs in LDS too
    constexpr int LSB = lds_slab_bytes(RPT);
    __shared__ __align__(16) uint8_t s_slab[LSB > 0 ? LSB : 16];
    const bool lds_s = stage_chunk<TPB>(A, tr0, tr1, TRW, s_h, s_t0, s_slab, LSB);
    // one workgroup with the whole gathered vector in LDS (up to LDS_GATHER
    // rows): the SpMV reads published entries there, with no global stores
    // to drain and no cache round trips
    constexpr int LG = RPT <= 4 ? LDS_GATHER : 1;
    __shared__ double s_g64[LG];
    __shared__ float s_g32[LG];
    const bool lds_g = RPT <= 4 && A.G == 1 && (long long)A.tilem * B <= LDS_GATHER;
    if (lds_g)
        for (int i = tid; i < LG; i += TPB) {
            s_g64[i] = 0.0;
            s_g32[i] = 0.0f;
        }
    __syncthreads();
    int rbase = tr0 * B + tid;  // row of step i is rbase + i * TPB
    auto tile_row = [&](int i) { return tr0 + i * GPB + grp; };
    auto row_of = [&](int i) { return (size_t)(rbase + i * TPB); };
    // the solve, with the gathered vector and the slabs each in LDS or in
    // global memory (up to four inlined copies, each with its address spaces
    // known to the compiler)
    auto solve = [&](double *g64, float *g32, auto lds_tag, auto slab_tag) {
        constexpr bool lds = decltype(lds_tag)::value;
        // the layout as the SpMV and the passes read it: slabs in LDS or HBM
        DevL Lw = A.L;
        if constexpr (decltype(slab_tag)::value) Lw.blob = s_slab;
        auto publish = [&](size_t row, double v) {
            if (lds) {
                if (A.has64) g64[row] = v;
                if (A.haslow) g32[row] = (float)v;
            } else {
                if (A.has64) st_agent(g64 + row, v);
                if (A.haslow) st_agent(g32 + row, (float)v);
            }
        };
        auto finish = [&](int status, int iter, int repl, double rt2, double g2) {
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) A.x[row_of(i)] = xr[i];
            if (blockIdx.x == 0 && tid == 0) {
                A.ctl[1] = status;
                A.ctl[2] = iter;
                A.ctl[3] = repl;
                A.out[1] = rt2;
                A.out[2] = g2;
            }
        };

        double v1[1] = {0.0};
    #pragma clang loop unroll(full)
        for (int i = 0; i < RPT; i++) {
            xr[i] = pr[i] = 0.0;
            s_r[i * TPB + tid] = 0.0;
            s_s[i * TPB + tid] = 0.0;
            if (tile_row(i) < tr1) {
                const size_t row = row_of(i);
                const double bv = A.b[row];
                s_r[i * TPB + tid] = bv;
                publish(row, bv);
                v1[0] += bv * bv;
            }
        }
        if (!R.allreduce(A, v1, true)) return;
        const double bnorm2 = v1[0];
        if (blockIdx.x == 0 && tid == 0) A.out[0] = sqrt(bnorm2);
        const double tol2 = A.tol * A.tol * bnorm2;
        if (bnorm2 == 0.0) { finish(MF_OK, 0, 0, 0.0, 0.0); return; }
        double gamma = bnorm2, ref2 = bnorm2, gprev = 1.0, tbest = bnorm2;
        bool restart = false;   // next iteration starts a new direction (beta = 0)
        bool force = false;     // a breakdown asks for a true-residual pass + restart
        int iter = 0, repl = 0;
        PhaseClock clk(A.phase);
        for (;;) {
            rbase = opaque(rbase);
            if (!isfinite(gamma)) { finish(MF_E_NONFINITE, iter, repl, -1.0, gamma); return; }
            const bool conv = gamma <= tol2;
            const bool due = force || (A.rr_factor > 0.0 && gamma <= A.rr_factor * A.rr_factor * ref2);
            if (conv || due) {
                // true residual b - A x and s = A p in one FP64-accumulated pass
    #pragma clang loop unroll(full)
                for (int i = 0; i < RPT; i++)
                    if (tile_row(i) < tr1) {
                        st_agent(A.h64 + row_of(i), xr[i]);
                        st_agent(A.q64 + row_of(i), pr[i]);
                    }
                if (!R.barrier(A)) return;
                double t1[1] = {0.0};
    #pragma unroll 1
                for (int i = 0; i < RPT; i++) {
                    const int tr = tile_row(i);
                    if (tr < tr1) {
                        const double *g[2] = {A.h64, A.q64};
                        double acc[2] = {0.0, 0.0};
                        const int q = i * GPB + grp;
                        if (A.use64)
                            tile_row_spmv_f64_c<B, IW, 2>(A.L64, row_ctx<B, IW>(A.L64, tr, lane), lane, lanebase, g, acc);
                        else
                            tile_row_spmv_f64_c<B, IW, 2>(
                                Lw, row_ctx_from<B, IW>(Lw, tr, s_t0[q], s_t0[q + 1] - s_t0[q], s_h[2 * q], s_h[2 * q + 1], lane),
                                lane, lanebase, g, acc);
                        const size_t row = (size_t)tr * B + lane;
                        const double rv = A.b[row] - acc[0];
                        s_r[i * TPB + tid] = rv;
                        s_s[i * TPB + tid] = acc[1];
                        publish(row, rv);
                        t1[0] += rv * rv;
                    }
                }
                if (!R.allreduce(A, t1, true)) return;
                clk.lap(4);
                ++repl;
                if (!isfinite(t1[0])) { finish(MF_E_NONFINITE, iter, repl, t1[0], gamma); return; }
                if (t1[0] <= tol2) { finish(MF_OK, iter, repl, t1[0], gamma); return; }
                // a lossy layout whose rounded operator lost definiteness (or whose
                // recurrences ran away) shows a true residual above the best one
                if (A.use64 && t1[0] > DIVERGE_GROWTH * tbest) { finish(MF_E_DIVERGED, iter, repl, t1[0], gamma); return; }
                tbest = fmin(tbest, t1[0]);
                // reliable update: keep the direction p (s = A p was refreshed with
                // it); when the recurrence had drifted by more than a factor 2 in
                // norm, restart from the true residual
                restart = force || t1[0] > 4.0 * gamma;
                force = false;
                gamma = ref2 = t1[0];
                continue;
            }
            if (iter >= A.maxit) { finish(MF_E_MAXIT, iter, repl, -1.0, gamma); return; }
            const bool fresh = iter == 0 || restart;
            restart = false;
            const double beta = fresh ? 0.0 : gamma / gprev;
            // w = A r; s = w + beta s
    #pragma unroll 1
            for (int i0 = 0; i0 < RPT; i0 += NR) {
                double w[NR];
                // four-byte indices: four quads of a row in flight up to 8 rows per
                // thread (at 16 the registers spill; BiCGSTAB's allow four only
                // for FP64)
                tile_rows_spmv<B, IW, NR, GPB, RPT <= 8 ? 4 : 2, RPT <= 8 ? 4 : 1, CM>(
                    Lw, s_h, s_t0, i0 * GPB + grp, tile_row(i0), lane, lanebase, g64, g32, w);
    #pragma unroll
                for (int k = 0; k < NR; k++) {
                    const int i = i0 + k;
                    if (tile_row(i) < tr1) s_s[i * TPB + tid] = w[k] + beta * s_s[i * TPB + tid];
                }
            }
            // p = r + beta p; (p, s)
            double red[1] = {0.0};
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    pr[i] = s_r[i * TPB + tid] + beta * pr[i];
                    red[0] += pr[i] * s_s[i * TPB + tid];
                }
            clk.lap(0);
            if (!R.allreduce(A, red, false)) return;
            clk.lap(1);
            const double denom = red[0];
            if (!isfinite(denom)) { finish(MF_E_NONFINITE, iter, repl, -1.0, gamma); return; }
            if (!(denom > 0.0)) {
                // (p, A p) <= 0 from rounding: recompute the residual and restart;
                // a breakdown right after a restart is final
                if (fresh) { finish(MF_E_BREAKDOWN, iter, repl, -1.0, gamma); return; }
                force = true;
                continue;
            }
            const double alpha = gamma / denom;
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) xr[i] = xr[i] + alpha * pr[i];
            double g1[1] = {0.0};
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    const double rv = s_r[i * TPB + tid] - alpha * s_s[i * TPB + tid];
                    s_r[i * TPB + tid] = rv;
                    publish(row_of(i), rv);
                    g1[0] += rv * rv;
                }
            gprev = gamma;
            ++iter;
            clk.lap(2);
            if (!R.allreduce(A, g1, true)) return;
            clk.lap(3);
            gamma = g1[0];
        }
    };
    if constexpr (LSB > 0) {
        if (lds_g) {
            if (lds_s)
                solve(s_g64, s_g32, std::true_type{}, std::true_type{});
            else
                solve(s_g64, s_g32, std::true_type{}, std::false_type{});
            return;
        }
        if (lds_s) {
            solve(A.g64, A.g32, std::false_type{}, std::true_type{});
            return;
        }
    }
    solve(A.g64, A.g32, std::false_type{}, std::false_type{});
}

// rho = (r^, r) and (r, r) of r = s - omega t from the reduced dots
// d = {(t, s), (t, t), (r^, s), (r^, t), (s, s)}; (r, r) is clamped at 0
// against rounding.
__device__ __forceinline__ void next_rho_g2(const double (&d)[5], double omega, double &rho, double &g2)
{
    rho = d[2] - omega * d[3];
    g2 = fmax(0.0, d[4] - omega * (2.0 * d[0] - omega * d[1]));
}

// BiCGSTAB: registers hold x, r, r^, p, s; LDS holds v = A p and t = A s.
// Four syncs per iteration: the reduction after t = A s carries (r^, s),
// (r^, t) and (s, s) besides (t, s) and (t, t), so that rho and (r, r) of the
// new iterate follow as (r^, s) - omega (r^, t) and (s, s) - 2 omega (t, s) +
// omega^2 (t, t) -- the values of (r^, r) and (r, r) in exact arithmetic --
// without a reduction of their own; s and p are published with neighbour
// waits.
template <int B, int IW, int TPB, int RPT, int CM = CM_ALL>
__global__ __launch_bounds__(TPB, 1) void k_bicgstab(SolveArgs A)
{
    __shared__ double s_v[RPT * TPB];
    __shared__ double s_t[RPT * TPB];
    __shared__ double s_red[(TPB / 64) * 5];
    __shared__ double s_bc[5];
    __shared__ int s_flag;
    Reducer<TPB, 5> R{s_red, s_bc, &s_flag};
    constexpr int GPB = TPB / B, NR = spmv_rows(RPT, IW);
    const int tid = threadIdx.x;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    const int grp = tid / B, lane = tid % B, lanebase = (tid & 63) & ~(B - 1);
    const int tr0 = chunk * A.trw, tr1 = min(A.tilem, tr0 + A.trw);
    double xr[RPT], rr[RPT], rh[RPT], pr[RPT], sr[RPT];
    // headers of the block's tile rows, loaded once: the per-iteration SpMV
    // then issues the tile-column word and the slab loads together
    constexpr int TRW = RPT * GPB;
    __shared__ int4 s_h[2 * TRW];
    __shared__ int s_t0[TRW + 1];
    // small chunks keep their slabs in LDS too
    constexpr int LSB = lds_slab_bytes(RPT);
    __shared__ __align__(16) uint8_t s_slab[LSB > 0 ? LSB : 16];
    const bool lds_s = stage_chunk<TPB>(A, tr0, tr1, TRW, s_h, s_t0, s_slab, LSB);
    // one workgroup with the whole gathered vector in LDS (up to LDS_GATHER
    // rows): the SpMV reads published entries there, with no global stores
    // to drain and no cache round trips
    constexpr int LG = RPT <= 4 ? LDS_GATHER : 1;
    __shared__ double s_g64[LG];
    __shared__ float s_g32[LG];
    const bool lds_g = RPT <= 4 && A.G == 1 && (long long)A.tilem * B <= LDS_GATHER;
    if (lds_g)
        for (int i = tid; i < LG; i += TPB) {
            s_g64[i] = 0.0;
            s_g32[i] = 0.0f;
        }
    __syncthreads();
    int rbase = tr0 * B + tid;  // row of step i is rbase + i * TPB
    auto tile_row = [&](int i) { return tr0 + i * GPB + grp; };
    auto row_of = [&](int i) { return (size_t)(rbase + i * TPB); };
    // the solve, with the gathered vector and the slabs each in LDS or in
    // global memory (up to four inlined copies, each with its address spaces
    // known to the compiler)
    auto solve = [&](double *g64, float *g32, auto lds_tag, auto slab_tag) {
        constexpr bool lds = decltype(lds_tag)::value;
        // the layout as the SpMV and the passes read it: slabs in LDS or HBM
        DevL Lw = A.L;
        if constexpr (decltype(slab_tag)::value) Lw.blob = s_slab;
        auto publish = [&](size_t row, double v) {
            if (lds) {
                if (A.has64) g64[row] = v;
                if (A.haslow) g32[row] = (float)v;
            } else {
                if (A.has64) st_agent(g64 + row, v);
                if (A.haslow) st_agent(g32 + row, (float)v);
            }
        };
        auto finish = [&](int status, int iter, int repl, double rt2, double g2) {
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) A.x[row_of(i)] = xr[i];
            if (blockIdx.x == 0 && tid == 0) {
                A.ctl[1] = status;
                A.ctl[2] = iter;
                A.ctl[3] = repl;
                A.out[1] = rt2;
                A.out[2] = g2;
            }
        };
        auto spmv_into = [&](double *dst) {
    #pragma unroll 1
            for (int i0 = 0; i0 < RPT; i0 += NR) {
                double w[NR];
                tile_rows_spmv<B, IW, NR, GPB, 2, 4, CM>(Lw, s_h, s_t0, i0 * GPB + grp, tile_row(i0), lane, lanebase,
                                                         g64, g32, w);
    #pragma unroll
                for (int k = 0; k < NR; k++)
                    if (tile_row(i0 + k) < tr1) dst[(i0 + k) * TPB + tid] = w[k];
            }
        };

        double v1[1] = {0.0};
    #pragma clang loop unroll(full)
        for (int i = 0; i < RPT; i++) {
            xr[i] = rr[i] = rh[i] = pr[i] = sr[i] = 0.0;
            s_v[i * TPB + tid] = 0.0;
            if (tile_row(i) < tr1) {
                const size_t row = row_of(i);
                rr[i] = rh[i] = pr[i] = A.b[row];
                publish(row, pr[i]);
                v1[0] += rr[i] * rr[i];
            }
        }
        if (!R.allreduce(A, v1, true)) return;
        const double bnorm2 = v1[0];
        if (blockIdx.x == 0 && tid == 0) A.out[0] = sqrt(bnorm2);
        const double tol2 = A.tol * A.tol * bnorm2;
        if (bnorm2 == 0.0) { finish(MF_OK, 0, 0, 0.0, 0.0); return; }
        double rho = bnorm2, ref2 = bnorm2, tbest = bnorm2;
        bool restart = false;
        int iter = 0, repl = 0;
        PhaseClock clk(A.phase);
        for (;;) {
            rbase = opaque(rbase);
            spmv_into(s_v);  // v = A p
            clk.lap(0);
            double a1[1] = {0.0};
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) a1[0] += rh[i] * s_v[i * TPB + tid];
            if (!R.allreduce(A, a1, false)) return;
            clk.lap(1);
            if (!(a1[0] != 0.0) || !isfinite(a1[0])) { finish(MF_E_BREAKDOWN, iter, repl, -1.0, -1.0); return; }
            const double alpha = rho / a1[0];
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    sr[i] = rr[i] - alpha * s_v[i * TPB + tid];
                    publish(row_of(i), sr[i]);
                }
            clk.lap(2);
            if (!R.published(A, chunk)) return;
            clk.lap(3);
            spmv_into(s_t);  // t = A s
            clk.lap(0);
            double a2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    const double t = s_t[i * TPB + tid];
                    a2[0] += t * sr[i];
                    a2[1] += t * t;
                    a2[2] += rh[i] * sr[i];
                    a2[3] += rh[i] * t;
                    a2[4] += sr[i] * sr[i];
                }
            if (!R.allreduce(A, a2, false)) return;
            clk.lap(1);
            const double omega = a2[1] > 0.0 ? a2[0] / a2[1] : 0.0;
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    xr[i] = xr[i] + alpha * pr[i] + omega * sr[i];
                    rr[i] = sr[i] - omega * s_t[i * TPB + tid];
                }
            double rho_new, g2;
            next_rho_g2(a2, omega, rho_new, g2);
            ++iter;
            if (!finite_all(rho_new, g2)) { finish(MF_E_NONFINITE, iter, repl, -1.0, g2); return; }
            const bool conv = g2 <= tol2;
            const bool due = A.rr_factor > 0.0 && g2 <= A.rr_factor * A.rr_factor * ref2;
            if (conv || due) {
    #pragma clang loop unroll(full)
                for (int i = 0; i < RPT; i++)
                    if (tile_row(i) < tr1) st_agent(A.h64 + row_of(i), xr[i]);
                if (!R.barrier(A)) return;
    #pragma unroll 1
                for (int i = 0; i < RPT; i++) {
                    const int tr = tile_row(i);
                    if (tr < tr1) {
                        const double *g[1] = {A.h64};
                        double acc[1] = {0.0};
                        const int q = i * GPB + grp;
                        if (A.use64)
                            tile_row_spmv_f64_c<B, IW, 1>(A.L64, row_ctx<B, IW>(A.L64, tr, lane), lane, lanebase, g, acc);
                        else
                            tile_row_spmv_f64_c<B, IW, 1>(
                                Lw, row_ctx_from<B, IW>(Lw, tr, s_t0[q], s_t0[q + 1] - s_t0[q], s_h[2 * q], s_h[2 * q + 1], lane),
                                lane, lanebase, g, acc);
                        s_t[i * TPB + tid] = A.b[(size_t)tr * B + lane] - acc[0];
                    }
                }
                double t2[2] = {0.0, 0.0};
    #pragma clang loop unroll(full)
                for (int i = 0; i < RPT; i++)
                    if (tile_row(i) < tr1) {
                        rr[i] = s_t[i * TPB + tid];
                        t2[0] += rr[i] * rr[i];
                        t2[1] += rr[i] * rh[i];
                    }
                if (!R.allreduce(A, t2, false)) return;
                clk.lap(4);
                ++repl;
                if (!isfinite(t2[0])) { finish(MF_E_NONFINITE, iter, repl, t2[0], g2); return; }
                if (t2[0] <= tol2) { finish(MF_OK, iter, repl, t2[0], g2); return; }
                if (A.use64 && t2[0] > DIVERGE_GROWTH * tbest) { finish(MF_E_DIVERGED, iter, repl, t2[0], g2); return; }
                tbest = fmin(tbest, t2[0]);
                restart = t2[0] > 4.0 * g2;  // recurrence drifted: new direction p = r
                ref2 = t2[0];
                rho_new = t2[1];
            }
            if (iter >= A.maxit) { finish(MF_E_MAXIT, iter, repl, -1.0, g2); return; }
            if (omega == 0.0 || rho == 0.0) { finish(MF_E_BREAKDOWN, iter, repl, -1.0, g2); return; }
            const double beta = restart ? 0.0 : (rho_new / rho) * (alpha / omega);
            restart = false;
    #pragma clang loop unroll(full)
            for (int i = 0; i < RPT; i++)
                if (tile_row(i) < tr1) {
                    pr[i] = rr[i] + beta * (pr[i] - omega * s_v[i * TPB + tid]);
                    publish(row_of(i), pr[i]);
                }
            rho = rho_new;
            clk.lap(2);
            if (!R.published(A, chunk)) return;
            clk.lap(3);
        }
    };
    if constexpr (LSB > 0) {
        if (lds_g) {
            if (lds_s)
                solve(s_g64, s_g32, std::true_type{}, std::true_type{});
            else
                solve(s_g64, s_g32, std::true_type{}, std::false_type{});
            return;
        }
        if (lds_s) {
            solve(A.g64, A.g32, std::false_type{}, std::true_type{});
            return;
        }
    }
    solve(A.g64, A.g32, std::false_type{}, std::false_type{});
}

// ------------------------------------------------- streamed (HBM) kernels
// Systems whose rows per workgroup exceed what registers and LDS hold run the
// same iterations with the Krylov vectors of a workgroup's rows in HBM.  Each
// thread owns the rows tr0 * B + j * TPB + tid of its chunk (j ascending), the
// same rows in the same order as a resident kernel on the same grid, so the
// partial sums -- and with them every iterate -- are bit-identical to it.
// Only the owning thread touches a row of these vectors; the gathered vector
// is published as in the resident kernels.  The tile-row headers of the
// chunk sit in LDS, a window of ST_HDR_BYTES at a time (one window for
// chunks up to ~2700 tile rows, i.e. beyond 10M rows at B = 16 on 256
// workgroups).
constexpr int ST_HDR_BYTES = 96 * 1024;

template <int B, int IW, int TPB>
struct StreamGeom {
    static constexpr int GPB = TPB / B;
    static constexpr int NR = IW < 4 ? 4 : 2;             // tile rows per lane group per SpMV step
    static constexpr int STEP = NR * GPB;                 // tile rows per SpMV step
    static constexpr int HW = (ST_HDR_BYTES / 36) / STEP * STEP;  // tile rows of headers in LDS
    static_assert(HW >= STEP, "header window");
};

// Stages the headers of tile rows [w0, w0 + HW) of chunk [.., tr1) in LDS
// (zero headers beyond the chunk); every thread of the workgroup calls it.
template <int B, int IW, int TPB>
__device__ __forceinline__ void stage_headers(const SolveArgs &A, int4 *s_h, int *s_t0, int w0, int tr1)
{
    constexpr int HW = StreamGeom<B, IW, TPB>::HW;
    __syncthreads();  // the previous window's readers are done
    for (int q = threadIdx.x; q <= HW; q += TPB) {
        const int tr = w0 + q;
        if (q < HW && tr < tr1) {
            const int4 *h = reinterpret_cast<const int4 *>(A.L.trh + (size_t)tr * 4);
            s_h[2 * q] = h[0];
            s_h[2 * q + 1] = h[1];
        } else if (q < HW) {
            s_h[2 * q] = s_h[2 * q + 1] = make_int4(0, 0, 0, 0);
        }
        s_t0[q] = A.L.tile_ptr[min(tr, A.tilem)];
    }
    __syncthreads();
}

// One SpMV pass of a streamed kernel over the chunk [tr0, tr1): y of the
// thread's rows in ascending order, handed to body(row, y).  hw0 tracks the
// window held in LDS (a chunk within one window is staged once per launch).
template <int B, int IW, int TPB, typename F>
__device__ __forceinline__ void stream_spmv(const SolveArgs &A, int4 *s_h, int *s_t0, int &hw0, int tr0, int tr1,
                                            const double *g64, const float *g32, F &&body)
{
    using Geo = StreamGeom<B, IW, TPB>;
    constexpr int GPB = Geo::GPB, NR = Geo::NR, STEP = Geo::STEP, HW = Geo::HW;
    const int tid = threadIdx.x, grp = tid / B, lane = tid % B, lanebase = (tid & 63) & ~(B - 1);
    for (int w0 = tr0; w0 < tr1; w0 += HW) {
        if (hw0 != w0) {
            stage_headers<B, IW, TPB>(A, s_h, s_t0, w0, tr1);
            hw0 = w0;
        }
        const int w1 = min(tr1, w0 + HW);
#pragma unroll 1
        for (int s0 = w0; s0 < w1; s0 += STEP) {
            double w[NR];
            tile_rows_spmv<B, IW, NR, GPB, 4, 4>(A.L, s_h, s_t0, s0 - w0 + grp, s0 + grp, lane, lanebase, g64, g32, w);
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const int tr = s0 + k * GPB + grp;
                if (tr < tr1) body((size_t)tr * B + lane, w[k]);
            }
        }
    }
}

// FP64-accumulated product of NV published vectors over the chunk's tile
// rows (true-residual passes), handed to body(row, acc) in the thread's row
// order; headers are read from the layout (use64: the FP64 layout).
template <int B, int IW, int TPB, int NV, typename F>
__device__ __forceinline__ void stream_spmv_f64(const SolveArgs &A, int tr0, int tr1, const double *const (&g)[NV],
                                                F &&body)
{
    const int tid = threadIdx.x, grp = tid / B, lane = tid % B, lanebase = (tid & 63) & ~(B - 1);
#pragma unroll 1
    for (int tr = tr0 + grp; tr < tr1; tr += TPB / B) {
        double acc[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] = 0.0;
        if (A.use64)
            tile_row_spmv_f64_c<B, IW, NV>(A.L64, row_ctx<B, IW>(A.L64, tr, lane), lane, lanebase, g, acc);
        else
            tile_row_spmv_f64_c<B, IW, NV>(A.L, row_ctx<B, IW>(A.L, tr, lane), lane, lanebase, g, acc);
        body((size_t)tr * B + lane, acc);
    }
}

// CG with the vectors in HBM: the iteration of k_cg (same recurrences, same
// syncs, same order of every partial sum).  Vectors: r, p, s at vec, x in
// the caller's buffer.
template <int B, int IW, int TPB>
__global__ __launch_bounds__(TPB, 1) void k_cg_st(SolveArgs A)
{
    using Geo = StreamGeom<B, IW, TPB>;
    __shared__ int4 s_h[2 * Geo::HW];
    __shared__ int s_t0[Geo::HW + 1];
    __shared__ double s_red[(TPB / 64) * 4];
    __shared__ double s_bc[4];
    __shared__ int s_flag;
    Reducer<TPB, 4> R{s_red, s_bc, &s_flag};
    const int tid = threadIdx.x;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    const int tr0 = chunk * A.trw, tr1 = min(A.tilem, tr0 + A.trw);
    const size_t row0 = (size_t)tr0 * B + tid, row1 = (size_t)tr1 * B;
    double *__restrict__ vr = A.vec, *__restrict__ vp = A.vec + A.vstride, *__restrict__ vs = A.vec + 2 * A.vstride;
    double *__restrict__ x = A.x;
    int hw0 = -1;
    auto publish = [&](size_t row, double v) {
        if (A.has64) st_agent(A.g64 + row, v);
        if (A.haslow) st_agent(A.g32 + row, (float)v);
    };
    auto finish = [&](int status, int iter, int repl, double rt2, double g2) {
        if (blockIdx.x == 0 && tid == 0) {
            A.ctl[1] = status;
            A.ctl[2] = iter;
            A.ctl[3] = repl;
            A.out[1] = rt2;
            A.out[2] = g2;
        }
    };
    double v1[1] = {0.0};
    for (size_t row = row0; row < row1; row += TPB) {
        const double bv = A.b[row];
        x[row] = 0.0;
        vp[row] = 0.0;
        vs[row] = 0.0;
        vr[row] = bv;
        publish(row, bv);
        v1[0] += bv * bv;
    }
    if (!R.allreduce(A, v1, true)) return;
    const double bnorm2 = v1[0];
    if (blockIdx.x == 0 && tid == 0) A.out[0] = sqrt(bnorm2);
    const double tol2 = A.tol * A.tol * bnorm2;
    if (bnorm2 == 0.0) { finish(MF_OK, 0, 0, 0.0, 0.0); return; }
    double gamma = bnorm2, ref2 = bnorm2, gprev = 1.0, tbest = bnorm2;
    bool restart = false, force = false;
    int iter = 0, repl = 0;
    PhaseClock clk(A.phase);
    for (;;) {
        if (!isfinite(gamma)) { finish(MF_E_NONFINITE, iter, repl, -1.0, gamma); return; }
        const bool conv = gamma <= tol2;
        const bool due = force || (A.rr_factor > 0.0 && gamma <= A.rr_factor * A.rr_factor * ref2);
        if (conv || due) {
            // true residual b - A x and s = A p in one FP64-accumulated pass
            for (size_t row = row0; row < row1; row += TPB) {
                st_agent(A.h64 + row, x[row]);
                st_agent(A.q64 + row, vp[row]);
            }
            if (!R.barrier(A)) return;
            double t1[1] = {0.0};
            const double *g[2] = {A.h64, A.q64};
            stream_spmv_f64<B, IW, TPB, 2>(A, tr0, tr1, g, [&](size_t row, const double(&acc)[2]) {
                const double rv = A.b[row] - acc[0];
                vr[row] = rv;
                vs[row] = acc[1];
                publish(row, rv);
                t1[0] += rv * rv;
            });
            if (!R.allreduce(A, t1, true)) return;
            clk.lap(4);
            ++repl;
            if (!isfinite(t1[0])) { finish(MF_E_NONFINITE, iter, repl, t1[0], gamma); return; }
            if (t1[0] <= tol2) { finish(MF_OK, iter, repl, t1[0], gamma); return; }
            // a lossy layout whose rounded operator lost definiteness (or whose
            // recurrences ran away) shows a true residual above the best one
            if (A.use64 && t1[0] > DIVERGE_GROWTH * tbest) { finish(MF_E_DIVERGED, iter, repl, t1[0], gamma); return; }
            tbest = fmin(tbest, t1[0]);
            restart = force || t1[0] > 4.0 * gamma;
            force = false;
            gamma = ref2 = t1[0];
            continue;
        }
        if (iter >= A.maxit) { finish(MF_E_MAXIT, iter, repl, -1.0, gamma); return; }
        const bool fresh = iter == 0 || restart;
        restart = false;
        const double beta = fresh ? 0.0 : gamma / gprev;
        // w = A r; s = w + beta s; p = r + beta p; (p, s)
        double red[1] = {0.0};
        stream_spmv<B, IW, TPB>(A, s_h, s_t0, hw0, tr0, tr1, A.g64, A.g32, [&](size_t row, double w) {
            const double s = w + beta * vs[row];
            vs[row] = s;
            const double p = vr[row] + beta * vp[row];
            vp[row] = p;
            red[0] += p * s;
        });
        clk.lap(0);
        if (!R.allreduce(A, red, false)) return;
        clk.lap(1);
        const double denom = red[0];
        if (!isfinite(denom)) { finish(MF_E_NONFINITE, iter, repl, -1.0, gamma); return; }
        if (!(denom > 0.0)) {
            if (fresh) { finish(MF_E_BREAKDOWN, iter, repl, -1.0, gamma); return; }
            force = true;
            continue;
        }
        const double alpha = gamma / denom;
        double g1[1] = {0.0};
        for (size_t row = row0; row < row1; row += TPB) {
            const double p = vp[row];
            x[row] = x[row] + alpha * p;
            const double rv = vr[row] - alpha * vs[row];
            vr[row] = rv;
            publish(row, rv);
            g1[0] += rv * rv;
        }
        gprev = gamma;
        ++iter;
        clk.lap(2);
        if (!R.allreduce(A, g1, true)) return;
        clk.lap(3);
        gamma = g1[0];
    }
}

// BiCGSTAB with the vectors in HBM: the iteration of k_bicgstab.  Vectors:
// r, r^, p, s, v = A p, t = A s at vec, x in the caller's buffer.
template <int B, int IW, int TPB>
__global__ __launch_bounds__(TPB, 1) void k_bicgstab_st(SolveArgs A)
{
    using Geo = StreamGeom<B, IW, TPB>;
    __shared__ int4 s_h[2 * Geo::HW];
    __shared__ int s_t0[Geo::HW + 1];
    __shared__ double s_red[(TPB / 64) * 5];
    __shared__ double s_bc[5];
    __shared__ int s_flag;
    Reducer<TPB, 5> R{s_red, s_bc, &s_flag};
    const int tid = threadIdx.x;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    const int tr0 = chunk * A.trw, tr1 = min(A.tilem, tr0 + A.trw);
    const size_t row0 = (size_t)tr0 * B + tid, row1 = (size_t)tr1 * B;
    const long long S = A.vstride;
    double *__restrict__ vr = A.vec, *__restrict__ vrh = A.vec + S, *__restrict__ vp = A.vec + 2 * S;
    double *__restrict__ vs = A.vec + 3 * S, *__restrict__ vv = A.vec + 4 * S, *__restrict__ vt = A.vec + 5 * S;
    double *__restrict__ x = A.x;
    int hw0 = -1;
    auto publish = [&](size_t row, double v) {
        if (A.has64) st_agent(A.g64 + row, v);
        if (A.haslow) st_agent(A.g32 + row, (float)v);
    };
    auto finish = [&](int status, int iter, int repl, double rt2, double g2) {
        if (blockIdx.x == 0 && tid == 0) {
            A.ctl[1] = status;
            A.ctl[2] = iter;
            A.ctl[3] = repl;
            A.out[1] = rt2;
            A.out[2] = g2;
        }
    };
    double v1[1] = {0.0};
    for (size_t row = row0; row < row1; row += TPB) {
        const double bv = A.b[row];
        x[row] = 0.0;
        vr[row] = vrh[row] = vp[row] = bv;
        vv[row] = 0.0;
        publish(row, bv);
        v1[0] += bv * bv;
    }
    if (!R.allreduce(A, v1, true)) return;
    const double bnorm2 = v1[0];
    if (blockIdx.x == 0 && tid == 0) A.out[0] = sqrt(bnorm2);
    const double tol2 = A.tol * A.tol * bnorm2;
    if (bnorm2 == 0.0) { finish(MF_OK, 0, 0, 0.0, 0.0); return; }
    double rho = bnorm2, ref2 = bnorm2, tbest = bnorm2;
    bool restart = false;
    int iter = 0, repl = 0;
    PhaseClock clk(A.phase);
    for (;;) {
        // v = A p; (r^, v)
        double a1[1] = {0.0};
        stream_spmv<B, IW, TPB>(A, s_h, s_t0, hw0, tr0, tr1, A.g64, A.g32, [&](size_t row, double w) {
            vv[row] = w;
            a1[0] += vrh[row] * w;
        });
        clk.lap(0);
        if (!R.allreduce(A, a1, false)) return;
        clk.lap(1);
        if (!(a1[0] != 0.0) || !isfinite(a1[0])) { finish(MF_E_BREAKDOWN, iter, repl, -1.0, -1.0); return; }
        const double alpha = rho / a1[0];
        for (size_t row = row0; row < row1; row += TPB) {
            const double s = vr[row] - alpha * vv[row];
            vs[row] = s;
            publish(row, s);
        }
        clk.lap(2);
        if (!R.published(A, chunk)) return;
        clk.lap(3);
        // t = A s; (t, s), (t, t), (r^, s), (r^, t), (s, s)
        double a2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        stream_spmv<B, IW, TPB>(A, s_h, s_t0, hw0, tr0, tr1, A.g64, A.g32, [&](size_t row, double t) {
            vt[row] = t;
            const double s = vs[row], rh = vrh[row];
            a2[0] += t * s;
            a2[1] += t * t;
            a2[2] += rh * s;
            a2[3] += rh * t;
            a2[4] += s * s;
        });
        clk.lap(0);
        if (!R.allreduce(A, a2, false)) return;
        clk.lap(1);
        const double omega = a2[1] > 0.0 ? a2[0] / a2[1] : 0.0;
        for (size_t row = row0; row < row1; row += TPB) {
            const double s = vs[row];
            x[row] = x[row] + alpha * vp[row] + omega * s;
            vr[row] = s - omega * vt[row];
        }
        double rho_new, g2;
        next_rho_g2(a2, omega, rho_new, g2);
        ++iter;
        if (!finite_all(rho_new, g2)) { finish(MF_E_NONFINITE, iter, repl, -1.0, g2); return; }
        const bool conv = g2 <= tol2;
        const bool due = A.rr_factor > 0.0 && g2 <= A.rr_factor * A.rr_factor * ref2;
        if (conv || due) {
            for (size_t row = row0; row < row1; row += TPB) st_agent(A.h64 + row, x[row]);
            if (!R.barrier(A)) return;
            double t2[2] = {0.0, 0.0};
            const double *g[1] = {A.h64};
            stream_spmv_f64<B, IW, TPB, 1>(A, tr0, tr1, g, [&](size_t row, const double(&acc)[1]) {
                const double r = A.b[row] - acc[0];
                vr[row] = r;
                t2[0] += r * r;
                t2[1] += r * vrh[row];
            });
            if (!R.allreduce(A, t2, false)) return;
            clk.lap(4);
            ++repl;
            if (!isfinite(t2[0])) { finish(MF_E_NONFINITE, iter, repl, t2[0], g2); return; }
            if (t2[0] <= tol2) { finish(MF_OK, iter, repl, t2[0], g2); return; }
            if (A.use64 && t2[0] > DIVERGE_GROWTH * tbest) { finish(MF_E_DIVERGED, iter, repl, t2[0], g2); return; }
            tbest = fmin(tbest, t2[0]);
            restart = t2[0] > 4.0 * g2;
            ref2 = t2[0];
            rho_new = t2[1];
        }
        if (iter >= A.maxit) { finish(MF_E_MAXIT, iter, repl, -1.0, g2); return; }
        if (omega == 0.0 || rho == 0.0) { finish(MF_E_BREAKDOWN, iter, repl, -1.0, g2); return; }
        const double beta = restart ? 0.0 : (rho_new / rho) * (alpha / omega);
        restart = false;
        for (size_t row = row0; row < row1; row += TPB) {
            const double p = vr[row] + beta * (vp[row] - omega * vv[row]);
            vp[row] = p;
            publish(row, p);
        }
        rho = rho_new;
        clk.lap(2);
        if (!R.published(A, chunk)) return;
        clk.lap(3);
    }
}

// ------------------------------------------------------------ dispatch
using KernelFn = void (*)(SolveArgs);

struct KernelChoice {
    KernelFn fn = nullptr;
    int tpb = 0, rpt = 0;
    int streamed = 0;  // 1: vectors in HBM (k_cg_st / k_bicgstab_st), rpt = 0
    int acc64 = 0;     // 1: built with CM_ACC64 (publishes the FP64 vector only)
};

constexpr int ST_TPB = 512;
// vectors of the streamed kernels: CG r, p, s; BiCGSTAB r, r^, p, s, v, t
__host__ __device__ constexpr int stream_vectors(int method) { return method == MF_CG ? 3 : 6; }

// 512 threads (8 waves per CU, 256 VGPRs per wave) hold the register-resident
// vectors of up to 16 rows per thread without spilling.  Beyond that (or when
// asked, stream = true) the streamed kernels keep the vectors in HBM.
template <int B, int IW>
KernelChoice pick_kernel(int method, long long rows_per_wg, bool stream)
{
    KernelChoice k;
    const int q = (int)((rows_per_wg + 511) / 512);
    if (stream) {
    } else if (method == MF_CG) {
        if (q <= 1) k = {k_cg<B, IW, 512, 1>, 512, 1};
        else if (q <= 2) k = {k_cg<B, IW, 512, 2>, 512, 2};
        else if (q <= 4) k = {k_cg<B, IW, 512, 4>, 512, 4};
        else if (q <= 8) k = {k_cg<B, IW, 512, 8>, 512, 8};
        // 16 rows per thread: r, s and the headers of 1024 tile rows of 8
        // rows exceed the 160 KB of LDS, so 8-row tiles stop at 8
        else if constexpr (B >= 16) {
            if (q <= 16) k = {k_cg<B, IW, 512, 16>, 512, 16};
        }
    } else {
        if (q <= 1) k = {k_bicgstab<B, IW, 512, 1>, 512, 1};
        else if (q <= 2) k = {k_bicgstab<B, IW, 512, 2>, 512, 2};
        else if (q <= 4) k = {k_bicgstab<B, IW, 512, 4>, 512, 4};
        else if (q <= 8) k = {k_bicgstab<B, IW, 512, 8>, 512, 8};
    }
    if (!k.fn)
        k = {method == MF_CG ? k_cg_st<B, IW, ST_TPB> : k_bicgstab_st<B, IW, ST_TPB>, ST_TPB, 0, 1};
    return k;
}

// Resident kernels built for a layout's class mask (CM_F64: FP64 tiles only;
// CM_LOW: no FP64 tiles) at 4 or more rows per thread, where the SpMV's live
// registers bind (16-row tiles, 128^3 Poisson CG at 16 rows per thread: 256
// VGPRs and 14 spilled with every class, 207 or 252 and none with one of
// these masks).  An empty choice leaves the pick to pick_kernel.
template <int B, int IW, int CM>
KernelChoice pick_kernel_cm(int method, long long rows_per_wg)
{
    KernelChoice k;
    const int q = (int)((rows_per_wg + 511) / 512);
    if (q <= 2) return k;
    if (method == MF_CG) {
        if (q <= 4) k = {k_cg<B, IW, 512, 4, CM>, 512, 4};
        else if (q <= 8) k = {k_cg<B, IW, 512, 8, CM>, 512, 8};
        else if constexpr (B >= 16) {
            if (q <= 16) k = {k_cg<B, IW, 512, 16, CM>, 512, 16};
        }
    } else {
        if (q <= 4) k = {k_bicgstab<B, IW, 512, 4, CM>, 512, 4};
        else if (q <= 8) k = {k_bicgstab<B, IW, 512, 8, CM>, 512, 8};
    }
    return k;
}

template <int B, int CM> KernelChoice pick_tile_cm(int iw, int method, long long rows_per_wg)
{
    return iw == 1   ? pick_kernel_cm<B, 1, CM>(method, rows_per_wg)
           : iw == 2 ? pick_kernel_cm<B, 2, CM>(method, rows_per_wg)
                     : pick_kernel_cm<B, 4, CM>(method, rows_per_wg);
}

// Sync-bound systems on layouts with low-precision tiles -- one row per
// thread, or one workgroup (up to 4 rows per thread) -- get kernels whose
// low-precision tiles accumulate in FP64 (CM_ACC64, every class compiled,
// with the FP64 tiles' blocking: no spills).  From 2 rows per thread on many
// workgroups the FP64 gathers cost more than the iterations they save.
template <int B, int IW>
KernelChoice pick_kernel_acc64(int method, long long rows_per_wg, int G)
{
    constexpr int CMA = CM_ALL | CM_ACC64;
    KernelChoice k;
    const int q = (int)((rows_per_wg + 511) / 512);
    if (q > 1 && G != 1) return k;
    if (method == MF_CG) {
        if (q <= 1) k = {k_cg<B, IW, 512, 1, CMA>, 512, 1};
        else if (q <= 2) k = {k_cg<B, IW, 512, 2, CMA>, 512, 2};
        else if (q <= 4) k = {k_cg<B, IW, 512, 4, CMA>, 512, 4};
    } else {
        if (q <= 1) k = {k_bicgstab<B, IW, 512, 1, CMA>, 512, 1};
        else if (q <= 2) k = {k_bicgstab<B, IW, 512, 2, CMA>, 512, 2};
        else if (q <= 4) k = {k_bicgstab<B, IW, 512, 4, CMA>, 512, 4};
    }
    k.acc64 = k.fn ? 1 : 0;
    return k;
}

template <int B> KernelChoice pick_tile_acc64(int iw, int method, long long rows_per_wg, int G)
{
    return iw == 1   ? pick_kernel_acc64<B, 1>(method, rows_per_wg, G)
           : iw == 2 ? pick_kernel_acc64<B, 2>(method, rows_per_wg, G)
                     : pick_kernel_acc64<B, 4>(method, rows_per_wg, G);
}

// tile sizes and class masks instantiated in their own translation units
KernelChoice pick_b8(int iw, int method, long long rows_per_wg, bool stream);   // solver_b8.hip
KernelChoice pick_b32(int iw, int method, long long rows_per_wg, bool stream);  // solver_b32.hip
KernelChoice pick_b16_f64(int iw, int method, long long rows_per_wg);           // solver_f64.hip
KernelChoice pick_b16_low(int iw, int method, long long rows_per_wg);           // solver_low.hip
KernelChoice pick_b16_e4m3(int iw, int method, long long rows_per_wg);          // solver_e4m3.hip
KernelChoice pick_b16_acc64(int iw, int method, long long rows_per_wg, int G);  // solver_acc64.hip

template <int B> KernelChoice pick_tile(int iw, int method, long long rows_per_wg, bool stream)
{
    return iw == 1   ? pick_kernel<B, 1>(method, rows_per_wg, stream)
           : iw == 2 ? pick_kernel<B, 2>(method, rows_per_wg, stream)
                     : pick_kernel<B, 4>(method, rows_per_wg, stream);
}

}  // namespace mf
