// GPU tile builder and compute-layout builder.
//
// Reference layout (PC/csr2block.h:267-422, bit-exact):  within a tile row the
// tiles are the distinct tile columns in ascending order, and each tile holds
// its nonzeros in CSR order (row-major, file order inside a row).  That order
// is a stable sort of the tile row's CSR segment by tile column, so the
// builder sorts (tile column, position) keys per tile row -- in LDS for
// segments up to SORT_CAP entries, in global scratch above that -- and every
// Blockcsr array of the tile row is then written at rowptr[first row] + k.
// Row starts are the uint8 exclusive scan of the per-row counts
// (PC/csr2block.h:85, :237).  Where those wrap (duplicate entries) or the
// scan is a no-op (a one-row tile row), the reference copy loop (:239-253)
// leaves positions unwritten; the device keeps every value (the compute
// layout is built from them) and mf_matrix_export zeroes those positions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "mf_matrix.h"

namespace mf {

static constexpr int SORT_CAP = 4096;  // LDS sort capacity per tile row (32 KB of keys)

// ------------------------------------------------------------- device scan
template <typename T>
__global__ __launch_bounds__(256) void k_scan_blocks(const T *in, T *out, T *block_tot, long long n)
{
    __shared__ T s_w[4];
    const long long base = (long long)blockIdx.x * 1024;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    T v[4];
    T tsum = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        long long i = base + tid * 4 + q;
        v[q] = i < n ? in[i] : T(0);
        tsum += v[q];
    }
    T inc = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    T wpre = 0;
    for (int w = 0; w < wv; w++) wpre += s_w[w];
    T run = wpre + inc - tsum;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        long long i = base + tid * 4 + q;
        if (i < n) out[i] = run;
        run += v[q];
    }
    if (tid == 255) block_tot[blockIdx.x] = wpre + inc;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_scan_totals(T *tot, long long nb)
{
    __shared__ T s_w[16];
    __shared__ T s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (long long base = 0; base < nb; base += 1024) {
        long long i = base + tid;
        T v = i < nb ? tot[i] : T(0);
        T inc = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            T u = __shfl_up(inc, off, 64);
            if (lane >= off) inc += u;
        }
        if (lane == 63) s_w[wv] = inc;
        __syncthreads();
        T wpre = 0;
        for (int w = 0; w < wv; w++) wpre += s_w[w];
        T carry = s_carry;
        if (i < nb) tot[i] = carry + wpre + inc - v;
        __syncthreads();
        if (tid == 1023) s_carry = carry + wpre + inc;
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_scan_add(T *out, const T *tot, long long n)
{
    const long long base = (long long)blockIdx.x * 1024;
    const T add = tot[blockIdx.x];
    for (int q = threadIdx.x; q < 1024; q += 256) {
        long long i = base + q;
        if (i < n) out[i] += add;
    }
}

// Exclusive scan of n elements; out may alias in.
template <typename T>
static int scan_exclusive(const T *in, T *out, long long n, hipStream_t st)
{
    if (n <= 0) return MF_OK;
    long long nb = (n + 1023) / 1024;
    T *tot = nullptr;
    MF_HIP(hipMallocAsync((void **)&tot, sizeof(T) * nb, st));
    k_scan_blocks<T><<<(unsigned)nb, 256, 0, st>>>(in, out, tot, n);
    k_scan_totals<T><<<1, 1024, 0, st>>>(tot, nb);
    k_scan_add<T><<<(unsigned)nb, 256, 0, st>>>(out, tot, n);
    MF_HIP(hipGetLastError());
    MF_HIP(hipFreeAsync(tot, st));
    return MF_OK;
}

// ------------------------------------------------------------ sort + count
template <int B>
__device__ __forceinline__ int ilog2B()
{
    return B == 8 ? 3 : (B == 16 ? 4 : 5);
}

__device__ __forceinline__ void cmp_swap(unsigned long long &a, unsigned long long &b, bool asc)
{
    if ((a > b) == asc) {
        unsigned long long t = a;
        a = b;
        b = t;
    }
}

// One block per tile row: sort (tile column << 32 | position) in LDS,
// write the permutation and the number of distinct tile columns.
template <int B>
__global__ __launch_bounds__(256) void k_sort_count(const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, int rowA,
                                                    int tilem, int *perm, int *tcount,
                                                    int *big_list, int *big_count)
{
    __shared__ unsigned long long keys[SORT_CAP];
    __shared__ int s_cnt[4];
    const int tid = threadIdx.x;
    for (int bi = blockIdx.x; bi < tilem; bi += gridDim.x) {
        const int r0 = bi * B, r1 = min(r0 + B, rowA);
        const int lo = rowptr[r0], hi = rowptr[r1], L = hi - lo;
        if (L > SORT_CAP) {
            if (tid == 0) {
                big_list[atomicAdd(big_count, 1)] = bi;
                tcount[bi] = 0;
            }
            continue;
        }
        int P = 1;
        while (P < L) P <<= 1;
        for (int k = tid; k < P; k += 256)
            keys[k] = k < L ? ((unsigned long long)(unsigned)(col[lo + k] >> ilog2B<B>()) << 32) | (unsigned)k
                            : ~0ull;
        __syncthreads();
        for (int size = 2; size <= P; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int k = tid; k < (P >> 1); k += 256) {
                    const int pos = 2 * k - (k & (stride - 1));
                    cmp_swap(keys[pos], keys[pos + stride], (pos & size) == 0);
                }
                __syncthreads();
            }
        int cnt = 0;
        for (int k = tid; k < L; k += 256) {
            perm[lo + k] = lo + (int)(unsigned)(keys[k] & 0xffffffffu);
            cnt += (k == 0 || (keys[k] >> 32) != (keys[k - 1] >> 32)) ? 1 : 0;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
        if ((tid & 63) == 0) s_cnt[tid >> 6] = cnt;
        __syncthreads();
        if (tid == 0) tcount[bi] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        __syncthreads();
    }
}

// Tile rows with more than SORT_CAP entries: one block each, bitonic sort in
// global scratch (keys_base + off[i], P[i] a power of two).
template <int B>
__global__ __launch_bounds__(1024) void k_sort_big(const int *__restrict__ rowptr,
                                                   const int *__restrict__ col, int rowA,
                                                   const int *big_list, const long long *koff,
                                                   unsigned long long *keys_base, int *perm,
                                                   int *tcount)
{
    __shared__ int s_cnt[16];
    const int tid = threadIdx.x;
    const int bi = big_list[blockIdx.x];
    unsigned long long *keys = keys_base + koff[blockIdx.x];
    const int r0 = bi * B, r1 = min(r0 + B, rowA);
    const int lo = rowptr[r0], hi = rowptr[r1], L = hi - lo;
    int P = 1;
    while (P < L) P <<= 1;
    for (int k = tid; k < P; k += 1024)
        keys[k] = k < L ? ((unsigned long long)(unsigned)(col[lo + k] >> ilog2B<B>()) << 32) | (unsigned)k
                        : ~0ull;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = tid; k < (P >> 1); k += 1024) {
                const int pos = 2 * k - (k & (stride - 1));
                unsigned long long a = keys[pos], b = keys[pos + stride];
                cmp_swap(a, b, (pos & size) == 0);
                keys[pos] = a;
                keys[pos + stride] = b;
            }
            __syncthreads();
        }
    int cnt = 0;
    for (int k = tid; k < L; k += 1024) {
        perm[lo + k] = lo + (int)(unsigned)(keys[k] & 0xffffffffu);
        cnt += (k == 0 || (keys[k] >> 32) != (keys[k - 1] >> 32)) ? 1 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((tid & 63) == 0) s_cnt[tid >> 6] = cnt;
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int w = 0; w < 16; w++) s += s_cnt[w];
        tcount[bi] = s;
    }
}

// ------------------------------------------------------------------- fill
// One block per tile row, in chunks of 256 sorted entries with a running tile
// counter: writes values, low values, tile-local columns and each entry's row
// inside the tile, and the tile columns and first positions of the tiles.
template <int B>
__global__ __launch_bounds__(256) void k_fill(const int *__restrict__ rowptr, const int *__restrict__ col,
                                              const double *__restrict__ val,
                                              const float *__restrict__ val_low, int rowA, int tilem,
                                              const int *__restrict__ perm, const int *__restrict__ tile_ptr,
                                              int *tile_columnidx, int *tile_start, double *bval,
                                              float *bval_low, uint8_t *bcol, uint8_t *rowtmp)
{
    __shared__ int s_rp[B + 1];
    __shared__ int s_w[4];
    __shared__ int s_carry, s_lasttc;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int bi = blockIdx.x; bi < tilem; bi += gridDim.x) {
        const int r0 = bi * B, r1 = min(r0 + B, rowA), nr = r1 - r0;
        const int lo = rowptr[r0], hi = rowptr[r1], L = hi - lo;
        if (tid <= nr) s_rp[tid] = rowptr[r0 + tid];
        if (tid == 0) { s_carry = 0; s_lasttc = -1; }
        __syncthreads();
        const int tb = tile_ptr[bi];
        for (int base = 0; base < L; base += 256) {
            const int k = base + tid;
            const bool ok = k < L;
            int e = 0, c = 0, tc = -1, prev = -1;
            if (ok) {
                e = perm[lo + k];
                c = col[e];
                tc = c >> ilog2B<B>();
                prev = k == 0 ? -1 : (col[perm[lo + k - 1]] >> ilog2B<B>());
            }
            const int f = ok && tc != prev ? 1 : 0;
            int inc = wave_incl_scan(f, lane);
            if (lane == 63) s_w[wv] = inc;
            __syncthreads();
            int wpre = 0;
            for (int w = 0; w < wv; w++) wpre += s_w[w];
            const int carry = s_carry;
            if (ok) {
                const int tix = carry + wpre + inc - 1;
                const int t = tb + tix;
                const int pos = lo + k;
                bval[pos] = val[e];
                bval_low[pos] = val_low ? val_low[e] : (float)val[e];
                bcol[pos] = (uint8_t)(c - tc * B);
                int rr = 0;  // row of entry e inside the tile row
                while (rr + 1 < nr && s_rp[rr + 1] <= e) ++rr;
                rowtmp[pos] = (uint8_t)rr;
                if (f) {
                    tile_columnidx[t] = tc;
                    tile_start[t] = pos;
                }
            }
            __syncthreads();
            if (tid == 0) s_carry = carry + s_w[0] + s_w[1] + s_w[2] + s_w[3];
            __syncthreads();
        }
    }
}

// One thread per tile: uint8 row starts, nnz mod 256 and row-slot offsets.
// The row starts are the reference's in-place exclusive scan of uint8
// per-row counts (PC/csr2block.h:85, :237), including its no-op for a
// single-row tile row (PC/utils.h:50).  The values stay complete on the
// device; the export applies the reference copy loop's coverage.
template <int B>
__global__ __launch_bounds__(256) void k_tiles(const int *__restrict__ tile_ptr, int tilem, int tilenum,
                                               int rowA, const int *__restrict__ tile_start,
                                               const uint8_t *__restrict__ rowtmp, uint8_t *bptr,
                                               uint8_t *blknnznnz, int *csrptr_offset)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int last = tile_ptr[tilem - 1];
    const int rowlen_last = rowA - (tilem - 1) * B;
    if (t == tilenum) {
        blknnznnz[t] = 0;
        csrptr_offset[t] = t <= last ? B * t : B * last + rowlen_last * (t - last);
    }
    if (t >= tilenum) return;
    int lo = 0, hi = tilem;  // tile row: largest bi with tile_ptr[bi] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tile_ptr[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int rowlen = lo == tilem - 1 ? rowlen_last : B;
    const int s = tile_start[t], e = tile_start[t + 1], nz = e - s;
    const int po = t <= last ? B * t : B * last + rowlen_last * (t - last);
    csrptr_offset[t] = po;
    blknnznnz[t] = (uint8_t)nz;
    unsigned cnt[B];
#pragma unroll
    for (int r = 0; r < B; r++) cnt[r] = 0;
    for (int k = s; k < e; k++) {
        const int rr = rowtmp[k];
#pragma unroll
        for (int r = 0; r < B; r++) cnt[r] += rr == r;
    }
    if (rowlen == 1) {
        bptr[po] = (uint8_t)cnt[0];
        return;
    }
    unsigned run = 0;
#pragma unroll
    for (int r = 0; r < B; r++) {
        if (r < rowlen) bptr[po + r] = (uint8_t)run;
        run = (run + (cnt[r] & 255u)) & 255u;
    }
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// ------------------------------------------------------------------ encode
// PC/encode.h:29 over the whole column array: byte i = col[2i] << 4 | col[2i+1],
// odd tail col[n-1] << 4.
__global__ void k_encode(const uint8_t *__restrict__ c, uint8_t *out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nb = (n + 1) / 2;
    if (i >= nb) return;
    const uint8_t a = c[2 * i];
    out[i] = 2 * i + 1 < n ? (uint8_t)((a << 4) + c[2 * i + 1]) : (uint8_t)(a << 4);
}

// ---------------------------------------------------- precision classes
// Round-to-nearest-even encoders written from the format definitions
// (OCP E4M3: bias 7, max 448, min normal 2^-6, subnormal step 2^-9, saturating;
// binary16: bias 15, max 65504, min normal 2^-14, subnormal step 2^-24).
__device__ uint8_t f64_to_e4m3(double v)
{
    if (isnan(v)) return 0x7F;
    const uint8_t s = signbit(v) ? 0x80 : 0;
    const double a = fabs(v);
    if (a >= 448.0) return s | 0x7E;
    if (a == 0.0) return s;
    int e2;
    frexp(a, &e2);
    int ue = e2 - 1;
    if (ue < -6) {
        const double r = rint(a * 512.0);
        return r >= 8.0 ? (uint8_t)(s | 8) : (uint8_t)(s | (int)r);
    }
    const double r = rint(ldexp(a, 3 - ue));
    int mant = (int)r - 8;
    if (mant == 8) { ue += 1; mant = 0; }
    int code = ((ue + 7) << 3) | mant;
    if (code > 0x7E) code = 0x7E;
    return (uint8_t)(s | code);
}

__device__ uint16_t f64_to_f16(double v)
{
    if (isnan(v)) return 0x7E00;
    const uint16_t s = signbit(v) ? 0x8000 : 0;
    const double a = fabs(v);
    if (a >= 65520.0) return s | 0x7C00;
    if (a == 0.0) return s;
    int e2;
    frexp(a, &e2);
    int ue = e2 - 1;
    if (ue < -14) {
        const double r = rint(ldexp(a, 24));
        return r >= 1024.0 ? (uint16_t)(s | 0x400) : (uint16_t)(s | (int)r);
    }
    const double r = rint(ldexp(a, 10 - ue));
    int mant = (int)r - 1024;
    if (mant == 1024) { ue += 1; mant = 0; }
    if (ue > 15) return s | 0x7C00;
    return (uint16_t)(s | ((ue + 15) << 10) | mant);
}

__device__ double e4m3_to_f64(uint8_t c) { return (double)e4m3_to_f32(c); }

__device__ double rounded(int code, double v)
{
    switch (code) {
    case 2: return e4m3_to_f64(f64_to_e4m3(v));
    case 3: return (double)f16_to_f32(f64_to_f16(v));
    case 4: return (double)(float)v;
    default: return v;
    }
}

__device__ bool fits(int code, double v, double tau)
{
    if (!isfinite(v)) return false;
    const double a = fabs(v);
    if (a == 0.0) return true;
    const double maxv = code == 2 ? 448.0 : (code == 3 ? 65504.0 : 3.4028234663852886e38);
    const double minn = code == 2 ? 0.015625 : (code == 3 ? 6.103515625e-05 : 1.1754943508222875e-38);
    if (a > maxv || a < minn) return false;
    return fabs(rounded(code, v) - v) <= tau * a;
}

// code (reference enumeration) -> cls (device width index)
__host__ __device__ inline int code_to_cls(int code) { return code == 4 ? 1 : code == 3 ? 2 : code == 2 ? 3 : 0; }

__global__ __launch_bounds__(256) void k_classes(const int *__restrict__ tile_start, int tilenum,
                                                 const double *__restrict__ bval, double tau,
                                                 int allow_mask, uint8_t *codes)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= tilenum) return;
    const int s = tile_start[t], e = tile_start[t + 1];
    const int order[3] = {2, 3, 4};
    uint8_t code = 0;
    for (int o = 0; o < 3 && !code; o++) {
        const int c = order[o];
        if (!(allow_mask & (1 << c))) continue;
        bool ok = true;
        for (int k = s; k < e && ok; k++) ok = fits(c, bval[k], tau);
        if (ok) code = (uint8_t)c;
    }
    codes[t] = code;
}

// out[0] = most tiles in a tile row, out[1] = largest |tile column - tile row|
__global__ __launch_bounds__(256) void k_max_tiles(const int *__restrict__ tile_ptr,
                                                   const int *__restrict__ tile_columnidx, int tilem, int *out)
{
    const int tr = blockIdx.x * 256 + threadIdx.x;
    if (tr >= tilem) return;
    const int t0 = tile_ptr[tr], t1 = tile_ptr[tr + 1];
    int d = 0;
    for (int t = t0; t < t1; t++) d = max(d, abs(tile_columnidx[t] - tr));
    atomicMax(out, t1 - t0);
    atomicMax(out + 1, d);
}

// ----------------------------------------------------------- pattern slabs
// A class slab of a tile row becomes a pattern slab (one table of offsets
// column - row shared by the B rows instead of an index per entry) when
// every row's entries of the class are in strictly ascending column order,
// the union of the offsets has at most PAT_MAX entries, every row + offset
// of the union (and row + 0 for the table's padding) is a valid column, and
// the slab gets smaller.
constexpr int PAT_MAX = 64;

// Union of the offsets of tile row tr's class-c entries, ascending, into U;
// returns its length, or 0 when some row is not strictly ascending or the
// union exceeds PAT_MAX.  lmax gets the longest row's entry count.
template <int B>
__device__ int pattern_union(int tr, int c, const int *__restrict__ tile_ptr, const int *__restrict__ tile_start,
                             const int *__restrict__ tile_columnidx, const uint8_t *__restrict__ rowtmp,
                             const uint8_t *__restrict__ bcol, const uint8_t *__restrict__ codes, int *U,
                             int *lmax)
{
    int prev[B], len[B];
    for (int r = 0; r < B; r++) {
        prev[r] = INT_MIN;
        len[r] = 0;
    }
    int n = 0;
    bool ok = true;
    for (int t = tile_ptr[tr]; t < tile_ptr[tr + 1]; t++) {
        if (code_to_cls(codes ? codes[t] : 0) != c) continue;
        const int cb = (tile_columnidx[t] - tr) * B;
        for (int k = tile_start[t]; k < tile_start[t + 1]; k++) {
            const int r = rowtmp[k], d = cb + bcol[k] - r;
            len[r]++;
            if (!ok) continue;
            if (d <= prev[r]) { ok = false; continue; }
            prev[r] = d;
            int lo = 0, hi = n;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (U[mid] < d) lo = mid + 1; else hi = mid;
            }
            if (lo < n && U[lo] == d) continue;
            if (n == PAT_MAX) { ok = false; continue; }
            for (int q = n; q > lo; q--) U[q] = U[q - 1];
            U[lo] = d;
            n++;
        }
    }
    int m = 0;
    for (int r = 0; r < B; r++) m = max(m, len[r]);
    *lmax = m;
    return ok ? n : 0;
}

__host__ __device__ inline unsigned long long slab_bytes(int L, int w, int iw, int B)
{
    return (unsigned long long)((L + 3) >> 2) * B * 4 * (w + iw);
}
__host__ __device__ inline unsigned long long pattern_bytes(int L, int w, int iw, int B)
{
    const unsigned long long nq = (unsigned long long)((L + 3) >> 2);
    return nq * B * 4 * w + ((nq * 4 * iw + 15) & ~15ull);
}

// pat[tr * 4 + c] = union length of a pattern slab, 0 for an indexed slab.
template <int B>
__global__ __launch_bounds__(256) void k_slab_pattern(const int *__restrict__ tile_ptr, int tilem, int ncolv,
                                                      const int *__restrict__ tile_start,
                                                      const int *__restrict__ tile_columnidx,
                                                      const uint8_t *__restrict__ rowtmp,
                                                      const uint8_t *__restrict__ bcol,
                                                      const uint8_t *__restrict__ codes, int iw, int *pat)
{
    const int tr = blockIdx.x * 256 + threadIdx.x;
    if (tr >= tilem) return;
    int U[PAT_MAX];
    for (int c = 0; c < 4; c++) {
        int lmax = 0;
        const int n = pattern_union<B>(tr, c, tile_ptr, tile_start, tile_columnidx, rowtmp, bcol, codes, U, &lmax);
        bool ok = n > 0;
        if (ok) {
            const long long lo = (long long)tr * B + min(0, U[0]), hi = (long long)tr * B + (B - 1) + max(0, U[n - 1]);
            ok = lo >= 0 && hi < ncolv;
        }
        ok = ok && pattern_bytes(n, cls_width(c), iw, B) < slab_bytes(lmax, cls_width(c), iw, B);
        pat[(size_t)tr * 4 + c] = ok ? n : 0;
    }
}

// Per tile row: row lengths per class -> slab sizes (8-byte units) and the
// tile / nonzero counts per code (stats[10]: pattern slabs).
template <int B>
__global__ __launch_bounds__(256) void k_slab_sizes(const int *__restrict__ tile_ptr, int tilem,
                                                    const int *__restrict__ tile_start,
                                                    const uint8_t *__restrict__ rowtmp,
                                                    const uint8_t *__restrict__ codes, int iw,
                                                    const int *__restrict__ pat,
                                                    unsigned long long *units, unsigned long long *stats)
{
    // tiles and nonzeros per code: per-thread counts, summed in LDS, then one
    // global add per block and code
    __shared__ unsigned long long s_stats[11];
    if (threadIdx.x < 11) s_stats[threadIdx.x] = 0ull;
    __syncthreads();
    const int tr = blockIdx.x * 256 + threadIdx.x;
    unsigned long long u = 0;
    if (tr < tilem) {
        int len[4][B];
        unsigned long long cnt[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < B; r++) len[c][r] = 0;
        for (int t = tile_ptr[tr]; t < tile_ptr[tr + 1]; t++) {
            const int code = codes ? codes[t] : 0, c = code_to_cls(code);
            const int s = tile_start[t], e = tile_start[t + 1];
            for (int k = s; k < e; k++) len[c][rowtmp[k]]++;
            cnt[code]++;
            cnt[5 + code] += (unsigned long long)(e - s);
        }
        for (int k = 0; k < 10; k++)
            if (cnt[k]) atomicAdd(&s_stats[k], cnt[k]);
        unsigned long long npat = 0;
        for (int c = 0; c < 4; c++) {
            int Lc = 0;
            for (int r = 0; r < B; r++) Lc = max(Lc, len[c][r]);
            const int Lp = pat ? pat[(size_t)tr * 4 + c] : 0;
            u += (Lp ? pattern_bytes(Lp, cls_width(c), iw, B) : slab_bytes(Lc, cls_width(c), iw, B)) / 8;
            npat += Lp ? 1 : 0;
        }
        if (npat) atomicAdd(&s_stats[10], npat);
        units[tr] = u;
    }
    __syncthreads();
    if (threadIdx.x < 11 && s_stats[threadIdx.x]) atomicAdd(&stats[threadIdx.x], s_stats[threadIdx.x]);
}


template <int B>
__global__ __launch_bounds__(256) void k_slab_pack(const int *__restrict__ tile_ptr, int tilem,
                                                   const int *__restrict__ tile_start,
                                                   const int *__restrict__ tile_columnidx,
                                                   const uint8_t *__restrict__ rowtmp,
                                                   const double *__restrict__ bval,
                                                   const uint8_t *__restrict__ bcol,
                                                   const uint8_t *__restrict__ codes, int iw,
                                                   const int *__restrict__ pat,
                                                   const unsigned long long *__restrict__ off8,
                                                   int2 *trh, int *stc, uint8_t *blob, int *err)
{
    const int tr = blockIdx.x * 256 + threadIdx.x;
    if (tr >= tilem) return;
    int U[PAT_MAX];
    const int t0 = tile_ptr[tr], t1 = tile_ptr[tr + 1];
    int len[4][B];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < B; r++) len[c][r] = 0;
    for (int t = t0; t < t1; t++) {
        const int c = code_to_cls(codes ? codes[t] : 0);
        for (int k = tile_start[t]; k < tile_start[t + 1]; k++) len[c][rowtmp[k]]++;
    }
    unsigned long long o = off8[tr];
    int slot = 0;
    for (int c = 0; c < 4; c++) {
        int Lc = 0;
        for (int r = 0; r < B; r++) { Lc = max(Lc, len[c][r]); len[c][r] = 0; }
        if (o > 0xFFFFFFFFull) { atomicOr(err, 2); return; }
        const int w = cls_width(c);
        const int Lp = pat ? pat[(size_t)tr * 4 + c] : 0;
        if (Lp) {
            // pattern slab: the offset table, then each entry at its offset's position
            int lmax = 0;
            const int n = pattern_union<B>(tr, c, tile_ptr, tile_start, tile_columnidx, rowtmp, bcol, codes, U, &lmax);
            if (n != Lp) { atomicOr(err, 4); return; }
            trh[(size_t)tr * 4 + c] = make_int2((int)((unsigned)o | 1u), Lp);
            const int nq = (Lp + 3) >> 2;
            uint8_t *vals = blob + o * 8;
            uint8_t *tab = vals + (size_t)nq * B * 4 * w;
            for (int q = 0; q < nq * 4; q++) {
                const int d = q < Lp ? U[q] : 0;
                if (iw == 2) reinterpret_cast<int16_t *>(tab)[q] = (int16_t)d;
                else reinterpret_cast<int32_t *>(tab)[q] = d;
            }
            for (int t = t0; t < t1; t++) {
                if (code_to_cls(codes ? codes[t] : 0) != c) continue;
                stc[t0 + slot] = tile_columnidx[t];
                const int cb = (tile_columnidx[t] - tr) * B;
                for (int k = tile_start[t]; k < tile_start[t + 1]; k++) {
                    const int r = rowtmp[k], d = cb + bcol[k] - r;
                    int lo = 0, hi = Lp;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (U[mid] < d) lo = mid + 1; else hi = mid;
                    }
                    const size_t e = ((size_t)(lo >> 2) * B + r) * 4 + (lo & 3);
                    const double v = bval[k];
                    switch (c) {
                    case 0: reinterpret_cast<double *>(vals)[e] = v; break;
                    case 1: reinterpret_cast<float *>(vals)[e] = (float)v; break;
                    case 2: reinterpret_cast<uint16_t *>(vals)[e] = f64_to_f16(v); break;
                    default: vals[e] = f64_to_e4m3(v); break;
                    }
                }
                slot++;
            }
            o += pattern_bytes(Lp, w, iw, B) / 8;
            continue;
        }
        trh[(size_t)tr * 4 + c] = make_int2((int)(unsigned)o, Lc);
        if (!Lc) continue;
        const int nq = (Lc + 3) >> 2;
        uint8_t *vals = blob + o * 8;
        uint8_t *idx = vals + (size_t)nq * B * 4 * w;
        for (int t = t0; t < t1; t++) {
            if (code_to_cls(codes ? codes[t] : 0) != c) continue;
            stc[t0 + slot] = tile_columnidx[t];
            for (int k = tile_start[t]; k < tile_start[t + 1]; k++) {
                const int r = rowtmp[k], q = len[c][r]++;
                const size_t e = ((size_t)(q >> 2) * B + r) * 4 + (q & 3);
                const double v = bval[k];
                switch (c) {
                case 0: reinterpret_cast<double *>(vals)[e] = v; break;
                case 1: reinterpret_cast<float *>(vals)[e] = (float)v; break;
                case 2: reinterpret_cast<uint16_t *>(vals)[e] = f64_to_f16(v); break;
                default: vals[e] = f64_to_e4m3(v); break;
                }
                if (iw == 1)
                    idx[e] = (uint8_t)(((unsigned)slot << log2_of<B>()) | bcol[k]);
                else if (iw == 2)
                    reinterpret_cast<int16_t *>(idx)[e] = (int16_t)((tile_columnidx[t] - tr) * B + bcol[k]);
                else
                    reinterpret_cast<uint32_t *>(idx)[e] = (uint32_t)tile_columnidx[t] * B + bcol[k];
            }
            slot++;
        }
        o += (unsigned long long)nq * B * 4 * (w + iw) / 8;
    }
}

// -------------------------------------------------------------- host side
template <int B>
static int build_reference_tiles_B(mf_matrix m)
{
    hipStream_t st = stream();
    const int tilem = m->tilem;
    int *tcount = nullptr, *perm = nullptr, *big_list = nullptr, *big_count = nullptr;
    MF_HIP(hipMallocAsync((void **)&tcount, sizeof(int) * (tilem + 1), st));
    MF_HIP(hipMallocAsync((void **)&perm, sizeof(int) * (size_t)std::max(m->csrsize, 1), st));
    MF_HIP(hipMallocAsync((void **)&big_list, sizeof(int) * (tilem + 1), st));
    MF_HIP(hipMallocAsync((void **)&big_count, sizeof(int), st));
    MF_HIP(hipMemsetAsync(big_count, 0, sizeof(int), st));
    MF_HIP(hipMemsetAsync(tcount, 0, sizeof(int) * (tilem + 1), st));
    const int grid = std::min(tilem, 65535 * 4);
    if (tilem > 0)
        k_sort_count<B><<<grid, 256, 0, st>>>(m->d_rowptr, m->d_col, m->rowA, tilem, perm, tcount,
                                              big_list, big_count);
    MF_HIP(hipGetLastError());
    int nbig = 0;
    MF_HIP(hipMemcpyAsync(&nbig, big_count, sizeof(int), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    unsigned long long *gkeys = nullptr;
    long long *koff = nullptr;
    if (nbig > 0) {
        std::vector<int> bl(nbig), rp_host;
        MF_HIP(hipMemcpy(bl.data(), big_list, sizeof(int) * nbig, hipMemcpyDeviceToHost));
        std::sort(bl.begin(), bl.end());
        std::vector<long long> off(nbig + 1, 0);
        for (int i = 0; i < nbig; i++) {
            int r0 = bl[i] * B, r1 = std::min(r0 + B, m->rowA);
            int a = 0, b = 0;
            MF_HIP(hipMemcpy(&a, m->d_rowptr + r0, sizeof(int), hipMemcpyDeviceToHost));
            MF_HIP(hipMemcpy(&b, m->d_rowptr + r1, sizeof(int), hipMemcpyDeviceToHost));
            long long P = 1;
            while (P < b - a) P <<= 1;
            off[i + 1] = off[i] + P;
        }
        MF_HIP(hipMemcpy(big_list, bl.data(), sizeof(int) * nbig, hipMemcpyHostToDevice));
        MF_HIP(hipMalloc((void **)&gkeys, sizeof(unsigned long long) * off[nbig]));
        MF_HIP(hipMalloc((void **)&koff, sizeof(long long) * (nbig + 1)));
        MF_HIP(hipMemcpy(koff, off.data(), sizeof(long long) * (nbig + 1), hipMemcpyHostToDevice));
        k_sort_big<B><<<nbig, 1024, 0, st>>>(m->d_rowptr, m->d_col, m->rowA, big_list, koff, gkeys,
                                            perm, tcount);
        MF_HIP(hipGetLastError());
    }
    // tile_ptr = exclusive scan of the distinct-column counts
    MF_HIP(hipMallocAsync((void **)&m->d_tile_ptr, sizeof(int) * (tilem + 1), st));
    int rc = scan_exclusive<int>(tcount, m->d_tile_ptr, tilem + 1, st);
    if (rc) return rc;
    MF_HIP(hipMemcpyAsync(&m->tilenum, m->d_tile_ptr + tilem, sizeof(int), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    if (gkeys) { MF_HIP(hipFree(gkeys)); MF_HIP(hipFree(koff)); }
    const int tilenum = m->tilenum;
    const size_t nzs = (size_t)std::max(m->csrsize, 1);
    MF_HIP(hipMallocAsync((void **)&m->d_tile_columnidx, sizeof(int) * (tilenum + 1), st));
    MF_HIP(hipMallocAsync((void **)&m->d_tile_start, sizeof(int) * (tilenum + 1), st));
    MF_HIP(hipMallocAsync((void **)&m->d_csrptr_offset, sizeof(int) * (tilenum + 1), st));
    MF_HIP(hipMallocAsync((void **)&m->d_blknnznnz, tilenum + 1, st));
    MF_HIP(hipMallocAsync((void **)&m->d_Blockcsr_Val, sizeof(double) * nzs, st));
    MF_HIP(hipMallocAsync((void **)&m->d_Blockcsr_Val_Low, sizeof(float) * nzs, st));
    MF_HIP(hipMallocAsync((void **)&m->d_Tile_csr_Col, nzs, st));
    MF_HIP(hipMallocAsync((void **)&m->d_rowtmp, nzs, st));
    // csrptrlen = sum of row lengths over tiles
    int last_tiles = 0;
    {
        int a = 0;
        MF_HIP(hipMemcpyAsync(&a, m->d_tile_ptr + tilem - 1, sizeof(int), hipMemcpyDeviceToHost, st));
        MF_HIP(hipStreamSynchronize(st));
        last_tiles = tilenum - a;
    }
    const int rowlen_last = m->rowA - (tilem - 1) * B;
    m->csrptrlen = B * (tilenum - last_tiles) + rowlen_last * last_tiles;
    MF_HIP(hipMallocAsync((void **)&m->d_Blockcsr_Ptr, (size_t)std::max(m->csrptrlen, 1), st));
    if (tilem > 0)
        k_fill<B><<<grid, 256, 0, st>>>(m->d_rowptr, m->d_col, m->d_val, m->d_val_low, m->rowA, tilem,
                                        perm, m->d_tile_ptr, m->d_tile_columnidx, m->d_tile_start,
                                        m->d_Blockcsr_Val, m->d_Blockcsr_Val_Low, m->d_Tile_csr_Col,
                                        m->d_rowtmp);
    k_set_int<<<1, 1, 0, st>>>(m->d_tile_start + tilenum, m->csrsize);
    if (tilem > 0)
        k_tiles<B><<<(tilenum + 1 + 255) / 256, 256, 0, st>>>(
            m->d_tile_ptr, tilem, tilenum, m->rowA, m->d_tile_start, m->d_rowtmp, m->d_Blockcsr_Ptr,
            m->d_blknnznnz, m->d_csrptr_offset);
    MF_HIP(hipGetLastError());
    MF_HIP(hipFreeAsync(tcount, st));
    MF_HIP(hipFreeAsync(perm, st));
    MF_HIP(hipFreeAsync(big_list, st));
    MF_HIP(hipFreeAsync(big_count, st));
    MF_HIP(hipStreamSynchronize(st));
    return MF_OK;
}

int build_reference_tiles(mf_matrix m)
{
    switch (m->B) {
    case 8: return build_reference_tiles_B<8>(m);
    case 16: return build_reference_tiles_B<16>(m);
    case 32: return build_reference_tiles_B<32>(m);
    default: return fail(MF_E_ARG, "tile size must be 8, 16 or 32");
    }
}

// The reference's column array: tile-local columns with the positions its
// copy loop leaves unwritten set to 0 (tiles whose uint8 row starts wrapped,
// or tiles of a one-row last tile row).  One thread per tile.
template <int B>
__global__ __launch_bounds__(256) void k_ref_cols(const int *__restrict__ tile_ptr, int tilem, int tilenum,
                                                  int rowA, const int *__restrict__ tile_start,
                                                  const int *__restrict__ csrptr_offset,
                                                  const uint8_t *__restrict__ bptr,
                                                  const uint8_t *__restrict__ cols, uint8_t *out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= tilenum) return;
    const int s = tile_start[t], nz = tile_start[t + 1] - s;
    const int rowlen = t >= tile_ptr[tilem - 1] ? rowA - (tilem - 1) * B : B;
    const bool holes = nz > 255 || rowlen == 1;
    const int po = csrptr_offset[t];
    for (int k = 0; k < nz; k++) {
        bool cov = !holes;
        for (int ri = 0; ri < rowlen && !cov; ri++) {
            const int a = bptr[po + ri], b = ri == rowlen - 1 ? nz : bptr[po + ri + 1];
            cov = a <= k && k < b;
        }
        out[s + k] = cov ? cols[s + k] : (uint8_t)0;
    }
}

int encode_cols_device(mf_matrix m, uint8_t *d_out)
{
    hipStream_t st = stream();
    long long n = m->csrsize, nb = (n + 1) / 2;
    if (nb <= 0) return MF_OK;
    uint8_t *cols = nullptr;
    MF_HIP(hipMallocAsync((void **)&cols, (size_t)n, st));
    const unsigned g = (unsigned)((m->tilenum + 255) / 256);
#define MF_REFCOLS(BB)                                                                                 \
    k_ref_cols<BB><<<g, 256, 0, st>>>(m->d_tile_ptr, m->tilem, m->tilenum, m->rowA, m->d_tile_start, \
                                      m->d_csrptr_offset, m->d_Blockcsr_Ptr, m->d_Tile_csr_Col, cols)
    if (m->B == 8) MF_REFCOLS(8);
    else if (m->B == 16) MF_REFCOLS(16);
    else MF_REFCOLS(32);
#undef MF_REFCOLS
    k_encode<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(cols, d_out, n);
    MF_HIP(hipGetLastError());
    MF_HIP(hipFreeAsync(cols, st));
    MF_HIP(hipStreamSynchronize(st));
    return MF_OK;
}

int classes_device(mf_matrix m, double tau, int allow_mask, uint8_t *d_codes)
{
    hipStream_t st = stream();
    if (m->tilenum > 0)
        k_classes<<<(m->tilenum + 255) / 256, 256, 0, st>>>(m->d_tile_start, m->tilenum,
                                                            m->d_Blockcsr_Val, tau, allow_mask, d_codes);
    MF_HIP(hipGetLastError());
    MF_HIP(hipStreamSynchronize(st));
    return MF_OK;
}

void Layout::release()
{
    if (trh) (void)hipFree(trh);
    if (stc) (void)hipFree(stc);
    if (blob) (void)hipFree(blob);
    trh = nullptr;
    stc = nullptr;
    blob = nullptr;
    built = false;
}

template <int B>
static int build_layout_B(mf_matrix m, Layout &L, double tau, int allow_mask, int want, bool pattern)
{
    hipStream_t st = stream();
    L.release();
    const int tilenum = m->tilenum, tilem = m->tilem;
    uint8_t *codes = nullptr;
    unsigned long long *units = nullptr, *stats = nullptr;
    int *err = nullptr;
    MF_HIP(hipMallocAsync((void **)&codes, (size_t)tilenum + 1, st));
    MF_HIP(hipMallocAsync((void **)&units, sizeof(unsigned long long) * (tilem + 1), st));
    MF_HIP(hipMallocAsync((void **)&stats, sizeof(unsigned long long) * 11, st));
    MF_HIP(hipMallocAsync((void **)&err, sizeof(int), st));
    MF_HIP(hipMemsetAsync(stats, 0, sizeof(unsigned long long) * 11, st));
    MF_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
    MF_HIP(hipMemsetAsync(units, 0, sizeof(unsigned long long) * (tilem + 1), st));
    const unsigned nbt = (unsigned)((tilenum + 255) / 256), nbr = (unsigned)((tilem + 255) / 256);
    if (allow_mask) {
        if (tilenum) k_classes<<<nbt, 256, 0, st>>>(m->d_tile_start, tilenum, m->d_Blockcsr_Val, tau, allow_mask, codes);
    } else {
        MF_HIP(hipMemsetAsync(codes, 0, (size_t)tilenum + 1, st));
    }
    int fits = 0;
    int rc = index_fits(m, &fits);
    if (rc) return rc;
    const int iw = resolve_index_width(want, fits);
    // pattern slabs: automatic index form only, two- or four-byte widths
    int *pat = nullptr;
    if (pattern && want == 0 && iw != 1 && tilem > 0) {
        MF_HIP(hipMallocAsync((void **)&pat, sizeof(int) * 4 * (size_t)tilem, st));
        k_slab_pattern<B><<<nbr, 256, 0, st>>>(m->d_tile_ptr, tilem, m->tilen * B, m->d_tile_start,
                                               m->d_tile_columnidx, m->d_rowtmp, m->d_Tile_csr_Col, codes, iw, pat);
        MF_HIP(hipGetLastError());
    }
    k_slab_sizes<B><<<nbr, 256, 0, st>>>(m->d_tile_ptr, tilem, m->d_tile_start, m->d_rowtmp, codes, iw, pat,
                                         units, stats);
    rc = scan_exclusive<unsigned long long>(units, units, tilem + 1, st);
    if (rc) return rc;
    unsigned long long total8 = 0, hs[11];
    MF_HIP(hipMemcpyAsync(&total8, units + tilem, sizeof(total8), hipMemcpyDeviceToHost, st));
    MF_HIP(hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    L.blob_bytes = (size_t)total8 * 8;
    MF_HIP(hipMalloc((void **)&L.trh, sizeof(int2) * 4 * ((size_t)tilem + 1)));
    MF_HIP(hipMalloc((void **)&L.stc, sizeof(int) * ((size_t)tilenum + 1)));
    MF_HIP(hipMalloc((void **)&L.blob, std::max<size_t>(L.blob_bytes, 64)));
    MF_HIP(hipMemsetAsync(L.blob, 0, std::max<size_t>(L.blob_bytes, 64), st));
    MF_HIP(hipMemsetAsync(L.stc, 0, sizeof(int) * ((size_t)tilenum + 1), st));
    MF_HIP(hipMemsetAsync(L.trh, 0, sizeof(int2) * 4 * ((size_t)tilem + 1), st));
    k_slab_pack<B><<<nbr, 256, 0, st>>>(m->d_tile_ptr, tilem, m->d_tile_start, m->d_tile_columnidx,
                                        m->d_rowtmp, m->d_Blockcsr_Val, m->d_Tile_csr_Col, codes, iw, pat,
                                        units, L.trh, L.stc, L.blob, err);
    MF_HIP(hipGetLastError());
    int herr = 0;
    MF_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    MF_HIP(hipFree(codes));
    if (pat) MF_HIP(hipFree(pat));
    MF_HIP(hipFree(units));
    MF_HIP(hipFree(stats));
    MF_HIP(hipFree(err));
    if (herr) {
        L.release();
        return fail(MF_E_FORMAT, herr & 4 ? "pattern slab rebuilt differently" : "layout larger than 32 GB");
    }
    L.pattern_slabs = (long long)hs[10];
    for (int c = 0; c < 5; c++) {
        L.tiles_per_code[c] = (long long)hs[c];
        L.nnz_per_code[c] = (long long)hs[5 + c];
    }
    L.iw = iw;
    L.want_iw = want;
    L.has64 = L.tiles_per_code[0] > 0;
    L.haslow = L.tiles_per_code[2] + L.tiles_per_code[3] + L.tiles_per_code[4] > 0;
    // algorithmic bytes: sum nnz (w+1) + tilenum (8+B) + 4 (tilem+1) + n (w_x + 8),
    // against the CSR form sum nnz (w+4) + 4 (n+1) + n (w_x + 8)
    const double wv[5] = {8, 0, 1, 2, 4};
    double val_tile = 0, val_csr = 0;
    for (int c = 0; c < 5; c++) {
        val_tile += L.nnz_per_code[c] * (wv[c] + 1);
        val_csr += L.nnz_per_code[c] * (wv[c] + 4);
    }
    const double n = (double)tilem * B;
    const double wx = (L.has64 ? 8.0 : 0.0) + (L.haslow ? 4.0 : 0.0);
    const double bt = val_tile + (double)tilenum * (8 + B) + 4.0 * (tilem + 1) + n * (wx + 8);
    const double bc = val_csr + 4.0 * (n + 1) + n * (wx + 8);
    L.bytes_alg = std::min(bt, bc);
    L.bytes_actual = (double)L.blob_bytes + 32.0 * tilem + 4.0 * tilenum + 4.0 * (tilem + 1) + n * (wx + 8);
    L.tau = tau;
    L.allow_mask = allow_mask;
    L.built = true;
    return MF_OK;
}

int build_layout(mf_matrix m, Layout &L, double tau, int allow_mask, int want_iw, bool pattern)
{
    double t0 = now_ms();
    int rc;
    if (!want_iw) want_iw = m->want_iw;
    switch (m->B) {
    case 8: rc = build_layout_B<8>(m, L, tau, allow_mask, want_iw, pattern); break;
    case 16: rc = build_layout_B<16>(m, L, tau, allow_mask, want_iw, pattern); break;
    case 32: rc = build_layout_B<32>(m, L, tau, allow_mask, want_iw, pattern); break;
    default: return fail(MF_E_ARG, "tile size");
    }
    L.split = m->d_split;
    m->ms_layout += now_ms() - t0;
    return rc;
}

int index_fits(mf_matrix m, int *fits)
{
    hipStream_t st = stream();
    const int B = m->B, tilem = m->tilem;
    int *d = nullptr;
    MF_HIP(hipMallocAsync((void **)&d, 2 * sizeof(int), st));
    MF_HIP(hipMemsetAsync(d, 0, 2 * sizeof(int), st));
    if (tilem) k_max_tiles<<<(unsigned)((tilem + 255) / 256), 256, 0, st>>>(m->d_tile_ptr, m->d_tile_columnidx, tilem, d);
    MF_HIP(hipGetLastError());
    int h[2] = {0, 0};
    MF_HIP(hipMemcpyAsync(h, d, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    MF_HIP(hipStreamSynchronize(st));
    MF_HIP(hipFree(d));
    // one byte: slot and column fit 8 bits and every slot has a lane of the
    // group to hold its tile column (B = 8: 8 slots, 16: 16, 32: 8); two
    // bytes: every column - tr * B fits an int16 (|tile column - tr| <= 32768 / B - 1)
    const int lb = B == 8 ? 3 : (B == 16 ? 4 : 5);
    *fits = (h[0] <= std::min(B, 1 << (8 - lb)) ? 1 : 0) | ((long long)(h[1] + 1) * B <= 32768 ? 2 : 0);
    return MF_OK;
}

int resolve_index_width(int want, int fits)
{
    const bool f1 = fits & 1, f2 = fits & 2;
    switch (want) {
    case 1: return f1 ? 1 : f2 ? 2 : 4;
    case 2: return f2 ? 2 : 4;
    case 4: return 4;
    default: return f2 ? 2 : f1 ? 1 : 4;
    }
}

}  // namespace mf
