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
    int min_iw = 1;             // narrowest width the build was allowed
    size_t blob_bytes = 0;
    long long tiles_per_code[5] = {0, 0, 0, 0, 0};
    long long nnz_per_code[5] = {0, 0, 0, 0, 0};
    bool has64 = false, haslow = false;
    double bytes_alg = 0.0;     // algorithmic bytes per SpMV
    double bytes_actual = 0.0;  // bytes the layout occupies (+ gather + y)
    DevL view(const int *tile_ptr) const { return DevL{tile_ptr, trh, stc, blob}; }
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
    void release();
};

}  // namespace mf

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
    int min_iw = 1;  // narrowest slab index width of this matrix's layouts
};

namespace mf {
int build_reference_tiles(mf_matrix m);              // tiles.hip
int build_layout(mf_matrix m, Layout &L, double tau, int allow_mask,
                 int min_iw = 1);  // tiles.hip (narrowest slab index width allowed)
int encode_cols_device(mf_matrix m, uint8_t *d_out);  // tiles.hip
int classes_device(mf_matrix m, double tau, int allow_mask, uint8_t *d_codes);  // tiles.hip
hipStream_t stream();
}  // namespace mf
