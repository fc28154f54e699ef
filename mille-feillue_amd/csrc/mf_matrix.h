// Device-resident matrix object behind the opaque mf_matrix handle.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "mf_device.h"
#include "mf_internal.h"

namespace mf {

int hip_fail(hipError_t e, const char *what);
#define MF_HIP(call)                                         \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return ::mf::hip_fail(e_, #call); \
    } while (0)

int ensure_device();
int device_cu_count();

// One compute layout (FP64-only, or per-tile mixed precision).
struct Layout {
    bool built = false;
    double tau = -1.0;
    int allow_mask = -1;
    int2 *trh = nullptr;
    int *stc = nullptr;
    uint8_t *blob = nullptr;
    int iw = 1;                 // slab index bytes: 1 slot form, 2 tile-row relative, 4 matrix column
    int want_iw = 0;            // width asked for (0 automatic; see resolve_index_width)
    size_t blob_bytes = 0;
    long long tiles_per_code[5] = {0, 0, 0, 0, 0};
    long long nnz_per_code[5] = {0, 0, 0, 0, 0};
    bool has64 = false, haslow = false;
    double bytes_alg = 0.0;     // algorithmic bytes per SpMV
    double bytes_actual = 0.0;  // bytes the layout occupies (+ gather + y)
    long long pattern_slabs = 0;  // (tile row, class) slabs in pattern form
    const uint8_t *split = nullptr;  // the matrix's split tile rows (not owned)
    DevL view(const int *tile_ptr) const { return DevL{tile_ptr, trh, stc, blob, split}; }
    void release();
};

// Solver scratch, sized for the matrix and grid on first use.
struct Workspace {
    size_t ncol = 0;
    int G = 0;
    double *g64 = nullptr, *h64 = nullptr, *q64 = nullptr;
    float *g32 = nullptr;
    unsigned long long *slots = nullptr;
    int *ctl = nullptr;
    double *out = nullptr;
    double *r_outer = nullptr, *d = nullptr, *tmp = nullptr;
    double *red = nullptr;
    // neighbour syncs of the persistent solvers, for the grid (dep_G chunks
    // of dep_trw tile rows): per chunk the other chunks whose rows its tile
    // columns reference (dep_ptr[dep_G + 1], dep_idx), the longest such list
    // (dep_max), and one epoch word per chunk, 64 bytes apart (flags)
    int dep_G = 0, dep_trw = 0, dep_max = -1;
    int *dep_ptr = nullptr, *dep_idx = nullptr;
    unsigned *flags = nullptr;
    // vectors of the streamed solver kernels (vec_n doubles)
    double *vec = nullptr;
    size_t vec_n = 0;
    void release();
};

}  // namespace mf

namespace mf {
struct VendorCtx;
}

struct mf_matrix_s {
    int B = 16, rowA = 0, colA = 0, nnz = 0;
    int tilem = 0, tilen = 0, tilenum = 0, csrsize = 0, csrptrlen = 0;
    // CSR input on the device
    int *d_rowptr = nullptr, *d_col = nullptr;
    double *d_val = nullptr;
    float *d_val_low = nullptr;
    // reference tile layout on the device
    int *d_tile_ptr = nullptr, *d_tile_columnidx = nullptr, *d_tile_start = nullptr;
    int *d_csrptr_offset = nullptr;
    uint8_t *d_blknnznnz = nullptr, *d_Blockcsr_Ptr = nullptr, *d_Tile_csr_Col = nullptr;
    uint8_t *d_rowtmp = nullptr;
    double *d_Blockcsr_Val = nullptr;
    float *d_Blockcsr_Val_Low = nullptr;
    // compute layouts and scratch
    mf::Layout L64, Lmix;
    mf::Workspace ws;
    double ms_build = 0.0, ms_layout = 0.0;
    int want_iw = 0;  // slab index width asked for this matrix's layouts (0 automatic)
    // solvers on a row-reordered copy (rows sorted by length, tile rows dealt
    // to the persistent workgroups' chunks): child = P A P^T, perm[new] = old
    mf_matrix_s *child = nullptr;
    mf::VendorCtx *vendor = nullptr;  // vendor-library baseline state (vendor.hip)
    int *d_perm = nullptr;  // child row -> row of this matrix, -1 for split segments and padding
    double *d_pb = nullptr, *d_px = nullptr;
    int child_grid = 0;
    int child_split_rows = 0;  // rows of this matrix split over lanes in the child
    // rows split over lanes (set on a reordered child): lanes per row of each
    // tile row, tilem + SPLIT_PAD entries (see DevL::split)
    uint8_t *d_split = nullptr;
};

namespace mf {
int build_reference_tiles(mf_matrix m);              // tiles.hip
// want_iw: slab index width asked for (0: the matrix's own setting);
// pattern: pattern slabs where they fit (automatic index form only)
int build_layout(mf_matrix m, Layout &L, double tau, int allow_mask, int want_iw = 0,
                 bool pattern = true);  // tiles.hip
// fits bit 0: the one-byte slot form fits, bit 1: the two-byte form fits
int index_fits(mf_matrix m, int *fits);  // tiles.hip
// the width used for a request (0 automatic: two bytes, else one, else four;
// 1, 2 or 4: that form if it fits, else the next wider one that fits)
int resolve_index_width(int want, int fits);  // tiles.hip
void vendor_release(mf_matrix m);             // vendor.hip
int vendor_solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts &o, mf_report *rep);
int encode_cols_device(mf_matrix m, uint8_t *d_out);  // tiles.hip
int classes_device(mf_matrix m, double tau, int allow_mask, uint8_t *d_codes);  // tiles.hip
hipStream_t stream();
// Live handles, so that mf_finalize can release what callers still hold and
// later destroy calls on released handles do nothing (capi.cpp).
void live_add(void *h, int kind);     // kind 0 matrix, 1 batch
bool live_remove(void *h, int kind);  // false when h is not live
bool live_has(void *h, int kind);
void release_phase_buffer();          // solver.hip
}  // namespace mf
