// Device building blocks for gfx950 (wave64): the slab layout's value and
// index decoders, the per-row tile SpMV used by every kernel, wave64
// reductions and the XCD-aware chunk order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace mf {

// ----------------------------------------------------------------- layout
// Compute layout ("tile-row slabs") of one matrix on the device, built from
// the reference tile arrays by tiles.hip.  Within a tile row the tiles are
// ordered by precision class (reference order inside a class); slot s is a
// tile's position in that order and stc[tile_ptr[tr] + s] its tile column.
// For each class c present in the tile row, row r's nonzeros in the class's
// tiles (slot order, then element order) are entries k = 0 .. L_c - 1 of a
// slab of ceil(L_c / 4) quads: entry k of row r sits at element
// ((k / 4) * B + r) * 4 + k % 4, so one lane reads four consecutive entries of
// its row with one load and the B lanes of a tile row read one contiguous
// block.  All value quads come first, then the index quads in one of three
// widths IW: one byte (slot << log2(B) | column in tile) when every tile row
// has few enough tiles for the slots of a tile row to fit the lanes of its
// group; else two bytes (int16 column - tr * B, the column relative to the
// tile row's first row) when every tile lies within 32768 / B - 1 tile
// columns of the diagonal; else four bytes (the matrix column itself).  Short
// rows are padded with value 0 / index 0 (in every form a valid column).
// Every slab size is a multiple of 16 bytes.  trh[tr * 4 + c] = {slab offset / 8, L_c}.  Class c:
// 0 FP64, 1 FP32, 2 FP16, 3 FP8 E4M3 (OCP).
// Pattern slabs (two- and four-byte forms; bit 0 of the offset word set):
// when the rows of a tile row have their class-c entries in ascending column
// order and the union of their offsets column - row is short (stencils), the
// slab stores one table of those offsets (ascending, L_c of them, in quads of
// 4 x IW bytes) instead of an index per entry; row r's entry with offset d
// sits at the table position of d, and positions it lacks hold value 0 at the
// in-range column row + d.  Entry order per row is unchanged, so the sums
// are bit-identical to the indexed form.
struct DevL {
    const int *tile_ptr;
    const int2 *trh;
    const int *stc;
    const uint8_t *blob;
    // Rows split over lanes (the row-reordered copy of a matrix with long
    // rows): split[tr] = S lanes per row in tile row tr (S = 1, 2, 4, 8 or 16,
    // at most B; entries beyond tilem read 0 = no split).  Lanes j S .. j S +
    // S - 1 hold consecutive segments of one row: the sum of their products
    // goes to lane j S, the others (rows of their own without entries) get 0.
    // nullptr: no split tile rows.
    const uint8_t *split = nullptr;
};

// Sum over the S-lane segment of this lane (S a power of two within the
// lane group, every lane of the segment active); the segment's first lane
// keeps it, the other lanes return 0.
__device__ __forceinline__ double split_sum(double v, int S, int lane)
{
    for (int off = 1; off < S; off <<= 1) v += __shfl_xor(v, off, 64);
    return (lane & (S - 1)) ? 0.0 : v;
}

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

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <int B> __host__ __device__ constexpr int log2_of() { return B == 8 ? 3 : (B == 16 ? 4 : 5); }

// Per-tile-row context of one lane group (B lanes, lane r = row r).
struct RowCtx {
    int t0;      // first tile of the tile row
    int tc_reg;  // index base: tile column of slot r (IW 1), tr * B (IW 2), 0 (IW 4)
    int4 h01, h23;  // slab headers of classes 0,1 and 2,3
    int split;   // lanes per row (DevL::split), 1 when none
    int trB;     // first row of the tile row (tr * B)
};

template <int B, int IW>
__device__ __forceinline__ RowCtx row_ctx(const DevL &L, int tr, int r)
{
    RowCtx c;
    c.t0 = L.tile_ptr[tr];
    const int nt = L.tile_ptr[tr + 1] - c.t0;
    c.tc_reg = IW == 1 ? (r < nt ? L.stc[c.t0 + r] : 0) : IW == 2 ? tr * B : 0;
    const int4 *h = reinterpret_cast<const int4 *>(L.trh + (size_t)tr * 4);
    c.h01 = h[0];
    c.h23 = h[1];
    c.split = L.split ? (int)L.split[tr] : 1;
    c.trB = tr * B;
    return c;
}

// Gather-vector index of one slab entry (as gather_col): base is the row
// context's index base, or the lane's row for a pattern slab.
template <int B, int IW>
__device__ __forceinline__ int gather_index(uint32_t id, int base, int lanebase)
{
    if (IW == 4 || IW == 2) return base + (int)id;
    const int slot = (int)(id >> log2_of<B>()), col = (int)(id & (B - 1));
    return __shfl(base, lanebase + slot, 64) * B + col;
}

// Slab of a header word: byte offset and pattern flag (bit 0 of offset / 8).
__device__ __forceinline__ size_t slab_off(int hx) { return (size_t)((unsigned)hx & ~1u) * 8; }
__device__ __forceinline__ bool slab_pattern(int hx) { return hx & 1; }

// One quad (four consecutive entries of one lane's row) of values and of
// indices, loaded with one or two vector loads.
template <int CLS> struct Quad;
template <> struct Quad<0> {
    double2 a, b;
    __device__ void load(const uint8_t *p, size_t q)
    {
        const double2 *v = reinterpret_cast<const double2 *>(p) + 2 * q;
        a = v[0];
        b = v[1];
    }
    __device__ void zero() { a = b = make_double2(0.0, 0.0); }
    template <int E> __device__ double get() const { return E == 0 ? a.x : E == 1 ? a.y : E == 2 ? b.x : b.y; }
};
template <> struct Quad<1> {
    float4 a;
    __device__ void load(const uint8_t *p, size_t q) { a = reinterpret_cast<const float4 *>(p)[q]; }
    __device__ void zero() { a = make_float4(0.f, 0.f, 0.f, 0.f); }
    template <int E> __device__ float get() const { return E == 0 ? a.x : E == 1 ? a.y : E == 2 ? a.z : a.w; }
};
template <> struct Quad<2> {
    uint2 a;
    __device__ void load(const uint8_t *p, size_t q) { a = reinterpret_cast<const uint2 *>(p)[q]; }
    __device__ void zero() { a = make_uint2(0u, 0u); }
    template <int E> __device__ float get() const
    {
        return f16_to_f32((uint16_t)(((E < 2 ? a.x : a.y) >> (16 * (E & 1))) & 0xffffu));
    }
};
template <> struct Quad<3> {
    uint32_t a;
    __device__ void load(const uint8_t *p, size_t q) { a = reinterpret_cast<const uint32_t *>(p)[q]; }
    __device__ void zero() { a = 0u; }
    // gfx950's FP8 conversion reads OCP E4M3; it equals e4m3_to_f32 on every
    // non-NaN code (the layout never stores NaN)
    template <int E> __device__ float get() const { return __builtin_amdgcn_cvt_f32_fp8((int)a, E); }
};
template <int IW> struct IdxQuad;
template <> struct IdxQuad<1> {
    uint32_t a;
    __device__ void load(const uint8_t *p, size_t q) { a = reinterpret_cast<const uint32_t *>(p)[q]; }
    __device__ void zero() { a = 0u; }
    template <int E> __device__ uint32_t get() const { return (a >> (8 * E)) & 0xffu; }
};
// two-byte form: get() returns the sign-extended int16 as a 32-bit word
template <> struct IdxQuad<2> {
    uint2 a;
    __device__ void load(const uint8_t *p, size_t q) { a = reinterpret_cast<const uint2 *>(p)[q]; }
    __device__ void zero() { a = make_uint2(0u, 0u); }
    template <int E> __device__ uint32_t get() const
    {
        const uint32_t w = E < 2 ? a.x : a.y;
        return (uint32_t)(E & 1 ? (int32_t)w >> 16 : (int32_t)(int16_t)(w & 0xffffu));
    }
};
template <> struct IdxQuad<4> {
    uint4 a;
    __device__ void load(const uint8_t *p, size_t q) { a = reinterpret_cast<const uint4 *>(p)[q]; }
    __device__ void zero() { a = make_uint4(0u, 0u, 0u, 0u); }
    template <int E> __device__ uint32_t get() const { return E == 0 ? a.x : E == 1 ? a.y : E == 2 ? a.z : a.w; }
};

template <int N, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (N > 0) {
        static_for<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

// Context of tile row tr from headers already at hand (t0, nt, slab
// headers); only the tile-column word of the lane's slot is loaded.
template <int B, int IW>
__device__ __forceinline__ RowCtx row_ctx_from(const DevL &L, int tr, int t0, int nt, int4 h01, int4 h23, int r)
{
    RowCtx c;
    c.t0 = t0;
    c.tc_reg = IW == 1 ? (r < nt ? L.stc[t0 + r] : 0) : IW == 2 ? tr * B : 0;
    c.h01 = h01;
    c.h23 = h23;
    c.split = L.split ? (int)L.split[tr] : 1;
    c.trB = tr * B;
    return c;
}

// Element idx of vector g with a 32-bit byte offset from the uniform base
// (vectors are below 4 GB, which the host checks).
template <typename T> __device__ __forceinline__ T ld_elem(const T *__restrict__ g, uint32_t idx)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(g) + (idx * (uint32_t)sizeof(T)));
}

// Element of the gathered vector for one slab index.  One-byte form (slot <<
// log2 B | col): tcb is the lane's copy of tile column * B of slot
// (lane - lanebase), fetched from the slot's lane with ds_bpermute.  Two-byte
// form: tcb = tr * B plus the sign-extended offset.  Four-byte form: the
// index is the column.
template <int B, int IW>
__device__ __forceinline__ uint32_t gather_col(uint32_t id, int tcb, int t0, const DevL &L, int lanebase)
{
    (void)t0;
    (void)L;
    if (IW == 4 || IW == 2) return (uint32_t)tcb + id;
    const uint32_t slot = id >> log2_of<B>(), col = id & (B - 1);
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lanebase + slot) << 2), tcb) + col;
}

// Tile-row headers staged in LDS: hd[2 q] = {off0, L0, off1, L1} and
// hd[2 q + 1] = {off2, L2, off3, L3} of local tile row q; t0s[q] is its first
// tile (t0s[q + 1] - t0s[q] tiles).
template <int CLS> __device__ __forceinline__ int2 lds_header(const int4 *hd, int q)
{
    return reinterpret_cast<const int2 *>(hd + 2 * q)[CLS];
}

// Index base of local tile row q (global tile row tr): tile column * B of
// slot r (one-byte form), tr * B (two-byte form), 0 (four-byte form).
template <int B, int IW>
__device__ __forceinline__ int load_tcr(const DevL &L, const int *t0s, int q, int tr, int r)
{
    if (IW == 2) return tr * B;
    if (IW == 4) return 0;
    const int t0 = t0s[q], nt = t0s[q + 1] - t0;
    return r < nt ? L.stc[t0 + r] * B : 0;
}

// One class slab of NR tile rows at once (local rows q0 + k * QS, this lane's
// row in each): per step the quads of NB rows are loaded together, then every
// gather is issued, then the FMAs run in each row's own entry order, so each
// row's sum is bit-identical to slab_spmv.  FP64 values go 1 quad x NR/2 rows
// per step (their quads are 4x wider), the other classes 2 quads x NR rows.
// With four-byte indices (unstructured columns, where rows of thousands of
// entries occur and one lane walks each row) more quads of a row are in
// flight per step: FP64 UQ64 when one row is (NB = 1), the other classes UQL4.
// A 2000-entry row is then ~125 dependent load -> gather steps, not ~500
// (random SPD, 1M rows: FP64 CG 315 -> 218 ms).
template <int B, int IW, int CLS, int NR, int QS, int UQL4, int UQ64, typename AccT, typename GT>
__device__ __forceinline__ void slab_spmv_rows(const DevL &L, const int4 *hd, const int *t0s, int q0,
                                               const int (&tcr)[NR], int row0, int r, int lanebase,
                                               const GT *__restrict__ g, AccT (&acc)[NR])
{
    // FP64 sums (FP64 tiles, or low-precision tiles under CM_ACC64) gather
    // FP64 entries: half the rows and one quad per step
    constexpr bool wide = CLS == 0 || sizeof(AccT) == 8;
    constexpr int NB = wide && NR > 1 ? NR / 2 : NR;
    constexpr int UQ = wide ? (NB == 1 && IW == 4 ? UQ64 : 1) : (IW == 4 ? UQL4 : 2);
#pragma unroll
    for (int kb = 0; kb < NR; kb += NB) {
        int2 h[NB];
        int nqmax = 0;
#pragma unroll
        for (int k = 0; k < NB; k++) {
            h[k] = lds_header<CLS>(hd, q0 + (kb + k) * QS);
            nqmax = max(nqmax, (h[k].y + 3) >> 2);
        }
        for (int j0 = 0; j0 < nqmax; j0 += UQ) {
            Quad<CLS> v[NB][UQ];
            IdxQuad<IW> id[NB][UQ];
#pragma unroll
            for (int k = 0; k < NB; k++) {
                const int nq = (h[k].y + 3) >> 2;
                const uint8_t *vals = L.blob + slab_off(h[k].x);
                const uint8_t *idx = vals + (size_t)nq * B * 4 * cls_width(CLS);
                // a pattern slab's offset table is one lane's worth of index quads
                const bool pat = IW != 1 && slab_pattern(h[k].x);
#pragma unroll
                for (int u = 0; u < UQ; u++) {
                    if (j0 + u < nq) {
                        v[k][u].load(vals, (size_t)(j0 + u) * B + r);
                        id[k][u].load(idx, pat ? (size_t)(j0 + u) : (size_t)(j0 + u) * B + r);
                    } else {
                        v[k][u].zero();
                        id[k][u].zero();
                    }
                }
            }
            GT xg[NB][UQ][4];
#pragma unroll
            for (int k = 0; k < NB; k++) {
                const int t0 = 0;
                // index base: the lane's row for a pattern slab, else the tile row's
                const int tcb = IW != 1 && slab_pattern(h[k].x)
                                    ? (IW == 2 ? tcr[kb + k] + r : row0 + (kb + k) * QS * B)
                                    : tcr[kb + k];
#pragma unroll
                for (int u = 0; u < UQ; u++)
                    static_for<4>([&](auto e) {
                        xg[k][u][e] =
                            ld_elem(g, gather_col<B, IW>(id[k][u].template get<e>(), tcb, t0, L, lanebase));
                    });
            }
            // padding entries (value 0) add +0 or -0 to a sum that starts at +0
            // and is never -0, so they leave it unchanged
#pragma unroll
            for (int k = 0; k < NB; k++)
#pragma unroll
                for (int u = 0; u < UQ; u++)
                    static_for<4>([&](auto e) {
                        acc[kb + k] = fma_t((AccT)v[k][u].template get<e>(), (AccT)xg[k][u][e], acc[kb + k]);
                    });
        }
    }
}

// Class masks of a layout: bit c set when it holds class-c tiles (0 FP64,
// 1 FP32, 2 FP16, 3 E4M3).  Kernels built for a mask compile only those
// classes' slab loops: fewer live registers in the SpMV.
constexpr int CM_ALL = 0xF, CM_F64 = 0x1, CM_LOW = 0xE, CM_E4M3 = 0x8;
// With CM_ACC64 the FP32, FP16 and E4M3 tiles' products are summed in FP64
// against the FP64 vector instead (kernels for sync-bound systems, whose
// FP64 vector stays in cache): the values keep their storage precision, and
// lossless tiles then give the FP64 product.
constexpr int CM_ACC64 = 0x10;

// y of NR tile rows (local rows q0 + k * QS, global tile rows trq + k * QS,
// headers in LDS).  Per row, slab by slab in class order: the FP64 tiles'
// products summed in FP64, the FP32, FP16 and E4M3 tiles' in FP32 against
// the FP32 copy of the vector (in FP64 against the FP64 vector under
// CM_ACC64); y = a64 + (double)a32.  Only the classes in CM are read (the
// layout holds no others).
template <int B, int IW, int NR, int QS, int UQL4 = 2, int UQ64 = 4, int CM = CM_ALL>
__device__ __forceinline__ void tile_rows_spmv(const DevL &L, const int4 *hd, const int *t0s, int q0, int trq,
                                               int r, int lanebase, const double *__restrict__ g64,
                                               const float *__restrict__ g32, double (&y)[NR])
{
    int tcr[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) tcr[k] = load_tcr<B, IW>(L, t0s, q0 + k * QS, trq + k * QS, r);
    double a64[NR];
    float a32[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        a64[k] = 0.0;
        a32[k] = 0.0f;
    }
    const int row0 = trq * B + r;
    if constexpr (CM & 1) slab_spmv_rows<B, IW, 0, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g64, a64);
    if constexpr ((CM & CM_ACC64) != 0) {
        if constexpr (CM & 2) slab_spmv_rows<B, IW, 1, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g64, a64);
        if constexpr (CM & 4) slab_spmv_rows<B, IW, 2, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g64, a64);
        if constexpr (CM & 8) slab_spmv_rows<B, IW, 3, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g64, a64);
    } else {
        if constexpr (CM & 2) slab_spmv_rows<B, IW, 1, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g32, a32);
        if constexpr (CM & 4) slab_spmv_rows<B, IW, 2, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g32, a32);
        if constexpr (CM & 8) slab_spmv_rows<B, IW, 3, NR, QS, UQL4, UQ64>(L, hd, t0s, q0, tcr, row0, r, lanebase, g32, a32);
    }
#pragma unroll
    for (int k = 0; k < NR; k++) y[k] = a64[k] + (double)a32[k];
    if (L.split) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int S = L.split[trq + k * QS];
            if (S > 1) y[k] = split_sum(y[k], S, r);
        }
    }
}

// FP64-accumulated product of NV gathered vectors at once (true residuals and
// recurrence refreshes): every stored value is widened to FP64.
template <int B, int IW, int CLS, int NV>
__device__ __forceinline__ void slab_spmv_f64(const DevL &L, const RowCtx &c, int2 h, int r, int lanebase,
                                              const double *const (&g)[NV], double (&acc)[NV])
{
    const int Lc = h.y, nq = (Lc + 3) >> 2;
    const uint8_t *vals = L.blob + slab_off(h.x);
    const uint8_t *idx = vals + (size_t)nq * B * 4 * cls_width(CLS);
    const bool pat = IW != 1 && slab_pattern(h.x);
    const int base = pat ? c.trB + r : c.tc_reg;
    for (int j = 0; j < nq; j++) {
        Quad<CLS> v;
        IdxQuad<IW> id;
        v.load(vals, (size_t)j * B + r);
        id.load(idx, pat ? (size_t)j : (size_t)j * B + r);
        static_for<4>([&](auto e) {
            if (j * 4 + e < Lc) {
                const double a = (double)v.template get<e>();
                const int gi = gather_index<B, IW>(id.template get<e>(), base, lanebase);
#pragma unroll
                for (int q = 0; q < NV; q++) acc[q] = __builtin_fma(a, g[q][gi], acc[q]);
            }
        });
    }
}

template <int B, int IW, int NV>
__device__ __forceinline__ void tile_row_spmv_f64_c(const DevL &L, const RowCtx &c, int r, int lanebase,
                                                    const double *const (&g)[NV], double (&acc)[NV])
{
    if (c.h01.y) slab_spmv_f64<B, IW, 0, NV>(L, c, make_int2(c.h01.x, c.h01.y), r, lanebase, g, acc);
    if (c.h01.w) slab_spmv_f64<B, IW, 1, NV>(L, c, make_int2(c.h01.z, c.h01.w), r, lanebase, g, acc);
    if (c.h23.y) slab_spmv_f64<B, IW, 2, NV>(L, c, make_int2(c.h23.x, c.h23.y), r, lanebase, g, acc);
    if (c.h23.w) slab_spmv_f64<B, IW, 3, NV>(L, c, make_int2(c.h23.z, c.h23.w), r, lanebase, g, acc);
    if (c.split > 1) {
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] = split_sum(acc[q], c.split, r);
    }
}

template <int B, int IW, int NV>
__device__ __forceinline__ void tile_row_spmv_f64(const DevL &L, int tr, int r, int lanebase,
                                                  const double *const (&g)[NV], double (&acc)[NV])
{
    tile_row_spmv_f64_c<B, IW, NV>(L, row_ctx<B, IW>(L, tr, r), r, lanebase, g, acc);
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

}  // namespace mf
