// Device building blocks for gfx950 (wave64): value decoders, the per-row tile
// SpMV used by every kernel, the grid barrier and deterministic reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mf {

// ----------------------------------------------------------------- layout
// Compute layout ("tile-row slabs") of one matrix on the device, built from
// the reference tile arrays by tiles.hip.  Within a tile row the tiles are
// ordered by precision class (reference order inside a class); slot s is a
// tile's position in that order and stc[tile_ptr[tr] + s] its tile column.
// For each class c present in the tile row, row r's nonzeros in the class's
// tiles (slot order, then element order) form a slab of L_c slices: slice k
// holds the k-th nonzero of every row, B consecutive values followed (after
// all L_c value slices) by B consecutive indices (slot << log2(B) | column in
// tile; one byte when every tile row has at most B tiles, else two).  Short
// rows are padded with value 0 / index 0.  trh[tr * 4 + c] = {slab offset / 8,
// L_c}.  Class c: 0 FP64, 1 FP32, 2 FP16, 3 FP8 E4M3 (OCP).
struct DevL {
    const int *tile_ptr;
    const int2 *trh;
    const int *stc;
    const uint8_t *blob;
};

__host__ __device__ constexpr int cls_width(int cls) { return 8 >> cls; }

// OCP E4M3 -> float (no infinities; NaN is never stored).
__device__ __forceinline__ float e4m3_to_f32(uint32_t c)
{
    const uint32_t e = (c >> 3) & 15u, m = c & 7u;
    const float sub = (float)m * 0.001953125f;  // m * 2^-9
    const float nrm = __uint_as_float(((e + 120u) << 23) | (m << 20));
    const float mag = e ? nrm : sub;
    return __uint_as_float(__float_as_uint(mag) | ((c & 0x80u) << 24));
}

__device__ __forceinline__ float f16_to_f32(uint16_t h)
{
    _Float16 v;
    __builtin_memcpy(&v, &h, 2);
    return (float)v;
}

template <int CLS> struct ClsT;
template <> struct ClsT<0> {
    using T = double;
    __device__ static double dec(double v) { return v; }
};
template <> struct ClsT<1> {
    using T = float;
    __device__ static float dec(float v) { return v; }
};
template <> struct ClsT<2> {
    using T = uint16_t;
    __device__ static float dec(uint16_t v) { return f16_to_f32(v); }
};
template <> struct ClsT<3> {
    using T = uint8_t;
    __device__ static float dec(uint8_t v) { return e4m3_to_f32(v); }
};

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <int B> __host__ __device__ constexpr int log2_of() { return B == 8 ? 3 : (B == 16 ? 4 : 5); }

// Per-tile-row context of one lane group (B lanes, lane r = row r).
struct RowCtx {
    int t0;      // first tile of the tile row
    int tc_reg;  // tile column of slot r (one-byte index form)
    int4 h01, h23;  // slab headers of classes 0,1 and 2,3
};

template <int B, bool IDX8>
__device__ __forceinline__ RowCtx row_ctx(const DevL &L, int tr, int r)
{
    RowCtx c;
    c.t0 = L.tile_ptr[tr];
    const int nt = L.tile_ptr[tr + 1] - c.t0;
    c.tc_reg = (IDX8 && r < nt) ? L.stc[c.t0 + r] : 0;
    const int4 *h = reinterpret_cast<const int4 *>(L.trh + (size_t)tr * 4);
    c.h01 = h[0];
    c.h23 = h[1];
    return c;
}

// Gather-vector index of one slab entry.
template <int B, bool IDX8>
__device__ __forceinline__ int gather_index(uint32_t id, const RowCtx &c, const DevL &L, int lanebase)
{
    const int slot = (int)(id >> log2_of<B>()), col = (int)(id & (B - 1));
    const int tc = IDX8 ? __shfl(c.tc_reg, lanebase + slot, 64) : L.stc[c.t0 + slot];
    return tc * B + col;
}

template <bool IDX8>
__device__ __forceinline__ uint32_t load_index(const uint8_t *idx, size_t e)
{
    return IDX8 ? (uint32_t)idx[e] : (uint32_t)reinterpret_cast<const uint16_t *>(idx)[e];
}

// One class slab, U slices per step: all value and index loads of the step
// are issued before the gathers, and all gathers before the FMAs.  AccT is
// double for class 0 (gather g64) and float otherwise (gather g32).
template <int B, bool IDX8, int CLS, int U, typename AccT, typename GT>
__device__ __forceinline__ void slab_spmv(const DevL &L, const RowCtx &c, int2 h, int r, int lanebase,
                                          const GT *__restrict__ g, AccT &acc)
{
    using VT = typename ClsT<CLS>::T;
    const int Lc = h.y;
    const uint8_t *base = L.blob + (size_t)(unsigned)h.x * 8;
    const VT *vals = reinterpret_cast<const VT *>(base);
    const uint8_t *idx = base + (size_t)Lc * B * sizeof(VT);
    for (int k0 = 0; k0 < Lc; k0 += U) {
        VT v[U];
        uint32_t id[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = min(k0 + u, Lc - 1);
            v[u] = vals[(size_t)k * B + r];
            id[u] = load_index<IDX8>(idx, (size_t)k * B + r);
        }
        GT xg[U];
#pragma unroll
        for (int u = 0; u < U; u++) xg[u] = g[gather_index<B, IDX8>(id[u], c, L, lanebase)];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k0 + u < Lc) acc = fma_t((AccT)ClsT<CLS>::dec(v[u]), (AccT)xg[u], acc);
    }
}

// Context of tile row tr from headers already at hand (t0, nt, slab
// headers); only the tile-column word of the lane's slot is loaded.
template <bool IDX8>
__device__ __forceinline__ RowCtx row_ctx_from(const DevL &L, int t0, int nt, int4 h01, int4 h23, int r)
{
    RowCtx c;
    c.t0 = t0;
    c.tc_reg = (IDX8 && r < nt) ? L.stc[t0 + r] : 0;
    c.h01 = h01;
    c.h23 = h23;
    return c;
}

// y_r of tile row tr for lane r: FP64 tiles accumulate in FP64 over g64, then
// FP32, FP16 and FP8 tiles (in that order) accumulate in FP32 over g32.
template <int B, bool IDX8, int U = 8>
__device__ __forceinline__ double tile_row_spmv_c(const DevL &L, const RowCtx &c, int r, int lanebase,
                                                  const double *__restrict__ g64,
                                                  const float *__restrict__ g32)
{
    double a64 = 0.0;
    float a32 = 0.0f;
    if (c.h01.y) slab_spmv<B, IDX8, 0, U>(L, c, make_int2(c.h01.x, c.h01.y), r, lanebase, g64, a64);
    if (c.h01.w) slab_spmv<B, IDX8, 1, U>(L, c, make_int2(c.h01.z, c.h01.w), r, lanebase, g32, a32);
    if (c.h23.y) slab_spmv<B, IDX8, 2, U>(L, c, make_int2(c.h23.x, c.h23.y), r, lanebase, g32, a32);
    if (c.h23.w) slab_spmv<B, IDX8, 3, U>(L, c, make_int2(c.h23.z, c.h23.w), r, lanebase, g32, a32);
    return a64 + (double)a32;
}

template <int B, bool IDX8, int U = 8>
__device__ __forceinline__ double tile_row_spmv(const DevL &L, int tr, int r, int lanebase,
                                                const double *__restrict__ g64,
                                                const float *__restrict__ g32)
{
    return tile_row_spmv_c<B, IDX8, U>(L, row_ctx<B, IDX8>(L, tr, r), r, lanebase, g64, g32);
}

// FP64-accumulated product of NV gathered vectors at once (true residuals and
// recurrence refreshes): every stored value is widened to FP64.
template <int B, bool IDX8, int CLS, int NV>
__device__ __forceinline__ void slab_spmv_f64(const DevL &L, const RowCtx &c, int2 h, int r, int lanebase,
                                              const double *const (&g)[NV], double (&acc)[NV])
{
    using VT = typename ClsT<CLS>::T;
    const int Lc = h.y;
    const uint8_t *base = L.blob + (size_t)(unsigned)h.x * 8;
    const VT *vals = reinterpret_cast<const VT *>(base);
    const uint8_t *idx = base + (size_t)Lc * B * sizeof(VT);
    for (int k = 0; k < Lc; k++) {
        const double v = (double)ClsT<CLS>::dec(vals[(size_t)k * B + r]);
        const int gi = gather_index<B, IDX8>(load_index<IDX8>(idx, (size_t)k * B + r), c, L, lanebase);
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] = __builtin_fma(v, g[q][gi], acc[q]);
    }
}

template <int B, bool IDX8, int NV>
__device__ __forceinline__ void tile_row_spmv_f64_c(const DevL &L, const RowCtx &c, int r, int lanebase,
                                                    const double *const (&g)[NV], double (&acc)[NV])
{
    if (c.h01.y) slab_spmv_f64<B, IDX8, 0, NV>(L, c, make_int2(c.h01.x, c.h01.y), r, lanebase, g, acc);
    if (c.h01.w) slab_spmv_f64<B, IDX8, 1, NV>(L, c, make_int2(c.h01.z, c.h01.w), r, lanebase, g, acc);
    if (c.h23.y) slab_spmv_f64<B, IDX8, 2, NV>(L, c, make_int2(c.h23.x, c.h23.y), r, lanebase, g, acc);
    if (c.h23.w) slab_spmv_f64<B, IDX8, 3, NV>(L, c, make_int2(c.h23.z, c.h23.w), r, lanebase, g, acc);
}

template <int B, bool IDX8, int NV>
__device__ __forceinline__ void tile_row_spmv_f64(const DevL &L, int tr, int r, int lanebase,
                                                  const double *const (&g)[NV], double (&acc)[NV])
{
    tile_row_spmv_f64_c<B, IDX8, NV>(L, row_ctx<B, IDX8>(L, tr, r), r, lanebase, g, acc);
}

// ---------------------------------------------------------- wave64 helpers
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;  // identical on every lane only up to rounding; callers read lane 0
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Maps dispatch order to a chunk index so that the blocks of one XCD (blocks
// are dealt round-robin over the 8 XCDs) get consecutive chunks, which keeps
// the neighbour rows of a chunk in that XCD's L2.
__device__ __forceinline__ int xcd_chunk(int bid, int nb)
{
    const int per = nb >> 3, rem = nb & 7, xcd = bid & 7, k = bid >> 3;
    return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + k;
}

__device__ __forceinline__ unsigned ld_agent(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier over monotone arrival counters (never reset inside a launch,
// so there is no reset race).  Blocks are split into 8 groups by blockIdx % 8
// (the XCD the dispatcher deals them to); each block adds 1 to its group's
// counter, the group's last arriver adds 1 to the global counter, and every
// block waits for the global counter to reach barriers_passed * groups.
// Writes before the barrier are released at agent scope and acquired after
// it.  Spreading the arrivals over 9 counters instead of one is what keeps
// a barrier of 256 blocks at a few microseconds.  The wait is bounded: after
// spin_limit polls the block raises the abort word, and every waiter leaves
// when it sees it.  Returns false on abort (the caller exits the kernel).
// Counter layout: bar[0] global, bar[32 * (1 + x)] group x (separate lines).
__device__ __forceinline__ bool grid_barrier(unsigned *bar, int *abort_word, unsigned ep,
                                             long long spin_limit, int *s_flag)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x, x = blockIdx.x & 7u;
        const unsigned groups = G < 8u ? G : 8u;
        const unsigned nx = (G >> 3) + (x < (G & 7u) ? 1u : 0u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned old =
            __hip_atomic_fetch_add(bar + 32 * (1 + x), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == ep * nx) __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ab = 0;
        long long spins = 0;
        while (ld_agent(bar) < ep * groups) {
            if (ld_agent(abort_word)) { ab = 1; break; }
            if (++spins > spin_limit) {
                __hip_atomic_store(abort_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ab = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_flag = ab;
    }
    __syncthreads();
    return *s_flag == 0;
}

}  // namespace mf
