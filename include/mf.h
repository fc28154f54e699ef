/*
 * mf.h -- C-ABI of the MI355X (gfx950) Mille-feuille solver library
 * (libmfamd.so).  Plain pointers and sizes only.
 *
 * Two groups of entry points:
 *  1. Drop-in replacements for the reference's host-visible functions, same
 *     names, argument meaning and side effects (cited per declaration).
 *  2. The library's own API (mf_*): device-resident matrices whose
 *     reference tile layout is built on the GPU, tiled SpMV, persistent
 *     CG/BiCGSTAB (FP64 or per-tile mixed precision with FP64 iterative
 *     refinement), batched solves.
 *
 * Reference citations use PC/ = Performance-Comparison/, MP/ = Mixed-Precision/,
 * MC/ = Memory-Cost/.
 */
#ifndef MF_H
#define MF_H
#include <stddef.h>
#include <stdint.h>
#include "mf_tile_matrix.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum {
    MF_OK = 0,
    MF_E_IO = -1,         /* fopen failed (mmio_info/mmio_data return -1) */
    MF_E_BANNER = -2,     /* Matrix Market banner (mmio returns -2) */
    MF_E_SIZE = -4,       /* size line (mmio returns -4) */
    MF_E_ARG = -10,       /* invalid argument or shape */
    MF_E_HIP = -11,       /* HIP runtime error (message in mf_last_error) */
    MF_E_NOMEM = -12,
    MF_E_FORMAT = -13,    /* matrix not representable in the device layout */
    MF_E_NODEV = -14,     /* no usable gfx950 device */
    MF_E_UNSUPPORTED = -15,
    MF_E_BREAKDOWN = -20, /* Krylov breakdown (zero/negative curvature, rho=0) */
    MF_E_MAXIT = -21,     /* iteration limit before tolerance */
    MF_E_TIMEOUT = -22,   /* grid barrier spin limit reached; kernel aborted */
    MF_E_NONFINITE = -23, /* NaN/Inf in a reduction */
    MF_E_DIVERGED = -24   /* lossy tiles: the true residual grew (rounded operator
                             not definite); solves then fall back to FP64 tiles */
} mf_status;

const char *mf_status_str(int status);
const char *mf_last_error(void);

/* ------------------------------------------- 1. reference drop-ins ------- */

/* PC/mmio_highlevel.h:17 -- header pass: sizes, nnz after symmetric
 * expansion.  Returns 0, -1 (fopen), -2 (banner), -4 (size line). */
int mmio_info(int *m, int *n, int *nnz, int *isSymmetric, char *filename);
/* PC/mmio_highlevel.h:275 -- CSR in file order (no column sort, duplicates
 * kept), symmetric/hermitian off-diagonals mirrored. */
int mmio_data(int *csrRowPtr, int *csrColIdx, double *csrVal, char *filename);
/* PC/mmio_highlevel.h:594 -- mmio_data with float values (each value
 * converted from its double parse, round to nearest). */
int mmio_data_Low(int *csrRowPtr, int *csrColIdx, float *csrVal, char *filename);
/* Binary CSR cache, PC/biio2.0/src/biio.h (format: int rows, int cols, int64
 * nnz, int64 row_ptr[rows+1], int col_idx[nnz], double val[nnz]).
 * binary_read_csr (:22) returns 0 or -1 (missing or short file) and mallocs
 * the arrays; binary_write_csr (:75) writes that layout; read_Dmatrix (:126)
 * reads a .cbd or .mtx by extension; read_Dmatrix_32 (:153) likewise with
 * 32-bit row pointers (left untouched beyond INT_MAX nonzeros);
 * read_Dmatrix_convert (:97) reads <name>.cbd when present, else the .mtx,
 * writing the .cbd cache.  Arrays are freed with free() or mf_free(). */
int binary_read_csr(int *row, int *col, int64_t *nnz, int64_t **row_ptr, int **col_idx, double **val,
                    char *filename);
void binary_write_csr(int row, int col, int64_t nnz, int64_t *row_ptr, int *col_idx, double *val, char *filename);
void read_Dmatrix(int *row, int *col, int64_t *nnz, int64_t **row_ptr, int **col_idx, double **val,
                  int *isSymmeticeR, char *filename);
void read_Dmatrix_32(int *row, int *col, int *nnz, int **row_ptr, int **col_idx, double **val, int *isSymmeticeR,
                     char *filename);
void read_Dmatrix_convert(int *row, int *col, int64_t *nnz, int64_t **row_ptr, int **col_idx, double **val,
                          int *isSymmeticeR, char *filename);
/* Matrix Market writer ("coordinate real general", 17 significant digits):
 * mmio_data reads the file back to the same CSR bit for bit.  No reference
 * counterpart (the reference only reads); for the generated configurations. */
int mf_write_mtx(const char *path, int m, int n, const int *rowptr, const int *colidx, const double *val);
/* PC/encode.h:29 / :52 -- 4-bit column packing, two per byte. */
void encode(unsigned char *str, unsigned char *res, int str_len, int offset);
void decode(unsigned char *str, unsigned char *res, int res_len);
/* PC/utils.h:34 / :50 -- in-place exclusive scans (char form wraps mod 256). */
void exclusive_scan(int *input, int length);
void exclusive_scan_char(unsigned char *input, int length);
/* PC/csr2block.h:267 -- builds the 16x16 (MF_TILE_BLOCK_SIZE) tile layout.
 * Here the build runs on the GPU and the CSR fields of Tile_matrix are filled
 * with host copies (malloc'ed, release with Tile_destroy).  On error the
 * struct is zeroed and mf_last_error() says why. */
void Tile_create(Tile_matrix *matrix, int rowA, int colA, int nnzA, int *csrRowPtrA,
                 int *csrColIdxA, double *csrValA, float *csrValA_Low);
/* PC/format.h:88 -- frees what Tile_create allocated. */
void Tile_destroy(Tile_matrix *matrix);
/* PC/blockspmv_cpu.h:2 -- the reference's host set-up routine over a
 * Tile_create'd matrix (16-row tiles): ptroffset1/ptroffset2[tilenum]
 * (nonzeros / row slots before each tile), the row-block split of tile rows
 * with more than 4 tiles (*rowblkblock entries in three malloc'ed arrays,
 * bit 31 marking split tile rows, free with free/mf_free), and y[tilem*16] =
 * the tile product A x summed per tile.  Columns are decoded from
 * Tile_csr_Col (the reference reads the never-filled csr_compressedIdx).
 * The CSR arguments and y_golden are unused, as in the reference. */
void blockspmv_cpu(Tile_matrix *matrix, int *ptroffset1, int *ptroffset2, int *rowblkblock,
                   unsigned int **blkcoostylerowidx, int **blkcoostylerowidx_colstart,
                   int **blkcoostylerowidx_colstop, int rowA, int colA, int nnzA, int *csrRowPtrA,
                   int *csrColIdxA, double *csrValA, double *x, double *y, double *y_golden);
/* PC/Mille-feuille_CG_AMD.cu:1240 -- FP64 CG.  Rows are truncated to a
 * multiple of 16 as in the reference (:1246-1248).  Solves to true relative
 * residual MF_DROPIN_TOL (1e-10) with at most max(maxiter, 10000) iterations
 * on the GPU persistent kernel, zeroes then writes x, stores the iteration
 * count in *iter and appends one row to cg_syncfree_amd_mi200.csv in the
 * reference format (:2021-2030). threshold is ignored as in the reference. */
void cg_solve_inc(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b,
                  int n, int *iter, int maxiter, double threshold, char *filename, int nnzR,
                  int ori, int block_nnz);
/* PC/Mille-feuille_BiCGSTAB_AMD.cu:837 -- FP64 BiCGSTAB; overwrites rhs with
 * A*1 as the reference does (:1343-1350), solves to 1e-10, appends one row
 * to bicg_syncfree_amd_mi200.csv (:1606-1615).  Returns 1 if not converged. */
int bicgstab(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR,
             char *filename, int n, int block_nnz);

/* ------------------------------------------------- 2. library API -------- */

/* Host copy of the reference tile layout (field meaning as Tile_matrix,
 * PC/format.h:14-86, CSR-tile fields only). */
typedef struct {
    int B, rowA, colA, tilem, tilen, tilenum, csrsize, csrptrlen;
    int *tile_ptr;          /* tilem+1 */
    int *tile_columnidx;    /* tilenum */
    int *tile_nnz;          /* tilenum+1, scanned */
    char *Format;           /* tilenum, all 0 (CSR tiles) */
    int *blknnz;            /* tilenum+1, scanned */
    unsigned char *blknnznnz; /* tilenum+1, per-tile nnz mod 256 */
    int *csr_offset;        /* tilenum+1 */
    int *csrptr_offset;     /* tilenum+1 */
    double *Blockcsr_Val;   /* csrsize */
    float *Blockcsr_Val_Low;/* csrsize */
    unsigned char *Blockcsr_Ptr; /* csrptrlen, uint8 row starts */
    unsigned char *Tile_csr_Col; /* csrsize, tile-local column */
    unsigned char *csr_compressedIdx; /* ceil(csrsize/2); zero, as the reference leaves it */
} mf_ref_tiles;

typedef struct mf_matrix_s *mf_matrix;

/* Library set-up on a device (also done lazily by every entry point). */
int mf_init(int device);
/* Releases every matrix and batch still held, the library's device buffers
 * and its stream (e.g. at process exit, before the runtime and any profiler
 * attached to it shut down).  Released handles may still be passed to the
 * destroy functions, which then do nothing; the next call sets the device up
 * again.  No reference counterpart (the reference frees per solve and leaks
 * some device buffers). */
void mf_finalize(void);

/* Uploads a CSR matrix and builds the reference tile layout on the GPU
 * (PC/csr2block.h:267 semantics, bit-exact).  B in {8, 16, 32}.  val_low may
 * be NULL ((float)val is used, as the reference's callers pass).  The caller's
 * arrays are not retained. */
int mf_matrix_create(int rowA, int colA, int nnz, const int *rowptr, const int *colidx,
                     const double *val, const float *val_low, int B, mf_matrix *out);
void mf_matrix_destroy(mf_matrix m);
/* Copies the reference-layout arrays to host (malloc'ed; mf_ref_tiles_free). */
int mf_matrix_export(mf_matrix m, mf_ref_tiles *out);
void mf_ref_tiles_free(mf_ref_tiles *t);
/* PC/encode.h:29 over the whole Tile_csr_Col array (pairing follows the
 * global nonzero index as PC/blockspmv_cpu.h:104 decodes it), computed on the
 * GPU; packed has ceil(csrsize/2) bytes. */
int mf_matrix_encode_cols(mf_matrix m, unsigned char *packed);

/* Per-tile storage class, computed on the GPU.  Codes follow the reference
 * enumeration of MP/Mille-feuille_CG.cu:257-352 (0=FP64, 4=FP32, 3=FP16,
 * 2=FP8 E4M3).  A tile takes the narrowest enabled format whose normal range
 * covers all its values and whose round-to-nearest error is <= tau*|v| for
 * every value (tau = 0: lossless only).  allow_mask bit k enables code k. */
int mf_matrix_classes(mf_matrix m, double tau, int allow_mask, unsigned char *codes);

typedef struct {
    int rowA, colA, nnz, B, tilem, tilen, tilenum;
    long long tiles_per_code[5];  /* index = code */
    long long nnz_per_code[5];
    double bytes_alg_fp64;        /* algorithmic bytes of one FP64 SpMV */
    double bytes_alg_mixed;       /* same for the last mixed layout built */
    double bytes_layout_mixed;    /* actual bytes of the mixed device layout */
    int index_width;              /* slab index bytes of that layout (1, 2 or 4) */
    long long pattern_slabs_fp64; /* (tile row, class) slabs in pattern form: FP64 layout */
    long long pattern_slabs_mixed;/* and the mixed layout (pattern slabs: one offset table
                                     per slab instead of an index per entry, where the
                                     rows of a tile row share their column offsets) */
} mf_matrix_info;
int mf_matrix_get_info(mf_matrix m, mf_matrix_info *info);

/* Slab index width (bytes) of this matrix's compute layouts.  0 (default)
 * = automatic: the two-byte tile-row-relative form where every tile lies
 * within 32768/B - 1 tile columns of the diagonal, else the one-byte slot
 * form where every tile row has few enough tiles (B = 8: 8, 16: 16, 32: 8),
 * else the four-byte column; with the two- and four-byte forms, a class slab
 * whose rows share their column offsets (stencils) stores one table of
 * offsets instead of an index per entry (pattern slab).  1, 2 or 4 = that
 * form where it fits, else the next wider one, and no pattern slabs.
 * Layouts already built are released.  Results are bit-identical across
 * forms.  No reference counterpart (the reference stores four-byte columns
 * only). */
int mf_matrix_set_index_width(mf_matrix m, int width);

/* Precision / solver options. */
enum { MF_CG = 0, MF_BICGSTAB = 1 };
enum { MF_FP64 = 0, MF_MIXED = 1 };
enum { MF_REORDER_AUTO = 0, MF_REORDER_OFF = 1, MF_REORDER_ON = 2 };
enum { MF_SYNC_AUTO = 0, MF_SYNC_GLOBAL = 1 };
enum { MF_VECTORS_AUTO = 0, MF_VECTORS_HBM = 1, MF_VECTORS_ONCHIP = 2 };
typedef struct {
    int method;         /* MF_CG or MF_BICGSTAB */
    int precision;      /* MF_FP64 or MF_MIXED (FP32/FP16/E4M3 tiles summed in FP32
                           against an FP32 copy of the vector; on sync-bound
                           systems -- one row per thread, or one workgroup --
                           in FP64 against the FP64 vector) */
    double tol;         /* true relative residual target (default 1e-10) */
    int maxit;          /* per inner solve (default 10000) */
    double tau;         /* mixed: per-tile rounding tolerance (default 0 = lossless) */
    int allow_mask;     /* mixed: enabled codes (default FP32|FP16|FP8) */
    double rr_factor;   /* residual replacement when ||r|| drops by this factor
                           since the last replacement (0: only at convergence) */
    int ir_max_outer;   /* lossy tiles: 0 (default) = one launch whose true-residual
                           passes use the FP64 layout (reliable updates against A);
                           > 0 = at most this many FP64 refinement steps around
                           inner solves on the lossy layout */
    double ir_inner_tol;/* relative tolerance of each inner solve of the refinement */
    int grid;           /* workgroups (0: one per CU) */
    int block;          /* threads per workgroup (0: auto, 512 or 1024) */
    long long spin_limit; /* grid-barrier polls before MF_E_TIMEOUT */
    int index_width;    /* slab index width, as mf_matrix_set_index_width
                           (0 = the matrix's setting; batched: 0 = automatic,
                           one width that every system fits) */
    int profile;        /* 1 = time the phases of the persistent launch
                           (workgroup 0's clock) into mf_report.phase_us */
    int reorder;        /* MF_REORDER_AUTO (default): the solvers run on P A P^T
                           with rows sorted by length and tile rows dealt to the
                           workgroups when the natural slabs hold more than
                           twice the nonzeros; _OFF: never; _ON: always */
    int sync;           /* BiCGSTAB, MF_SYNC_AUTO (default): after publishing p
                           or s a workgroup waits only for the workgroups whose
                           rows its tile columns reference (when none references
                           more than 64); MF_SYNC_GLOBAL: a grid barrier */
    int vectors;        /* MF_VECTORS_AUTO (default): the Krylov vectors of a
                           workgroup's rows stay in registers and LDS where they
                           fit (CG up to 16 rows per thread at B >= 16, else 8;
                           BiCGSTAB 8) and are streamed from HBM beyond that;
                           MF_VECTORS_HBM: single systems: MF_E_UNSUPPORTED where
                           the vectors fit on chip (forcing the streamed kernels
                           there is switched off);
                           MF_VECTORS_ONCHIP: single systems: MF_E_UNSUPPORTED
                           where they do not fit on chip.
                           Batches: AUTO and HBM keep the vectors in HBM (one
                           workgroup per system where its share of the work
                           allows); ONCHIP runs every system on a gang of
                           workgroups that keeps x, r, p and s of 8 columns in
                           LDS (2 to 8 right-hand sides; MF_E_UNSUPPORTED for
                           one) */
} mf_opts;
void mf_opts_default(mf_opts *o);

typedef struct {
    int status;          /* mf_status of the solve */
    int iterations;      /* Krylov iterations, all inner solves */
    int outer;           /* IR steps */
    int replacements;    /* true-residual replacements inside the kernel */
    double relres_true;  /* ||b - A x|| / ||b||, FP64, computed on the GPU */
    double ms_solve;     /* time to tolerance (device events, inputs resident) */
    double ms_setup;     /* layout build */
    int grid, block, rows_per_thread;
    double phase_us[5];  /* opts.profile, per iteration of the last persistent
                            launch: CG: SpMV, sync after it, vector update, sync
                            after it; BiCGSTAB: both SpMVs, the reductions after
                            them, vector updates, the other three syncs;
                            [4] = true-residual passes, total */
    int reordered;       /* 1: solved on the row-reordered copy */
    double pad_factor;   /* stored slab entries / nonzeros of the natural FP64 layout */
    int sync_deps;       /* BiCGSTAB neighbour syncs: most workgroups one waits
                            for after a publish; -1: grid barriers */
    int streamed;        /* 1: the Krylov vectors lived in HBM (systems beyond
                            register/LDS residency, or opts.vectors) */
    int fallback;        /* 1: the solve on lossy tiles diverged, broke down or
                            hit maxit, and was redone on the FP64 layout (its
                            time and iterations are included); batched: the
                            systems that left their lossy tiles for the FP64
                            layout inside the launch */
    int split_rows;      /* reordered solves: rows longer than 64 entries,
                            each split over 2..16 lanes of its tile row */
} mf_report;

/* Host buffers in/out (copies included; the timed ms_solve excludes them). */
int mf_solve(mf_matrix m, const double *b, double *x, const mf_opts *o, mf_report *rep);
/* Device pointers (n = tilem*B entries, x zero-initialised by the call),
 * on the library stream. */
int mf_solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts *o,
                    mf_report *rep);

/* mf_solve_device without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The launch's
 * buffers (sync and control words, published vector, true-residual slots)
 * sit in one static device array per GPU of about 56.4 MiB, held until the
 * process ends, looked up by mf_init.  Limits: at most 2,097,152 padded
 * rows (tilem*B) in the system solved, at most 1024 workgroups, Krylov
 * vectors resident in registers/LDS (CG up to 2,097,152 rows and BiCGSTAB up
 * to 1,048,576 on 256 CUs at B = 16; MF_E_UNSUPPORTED beyond), no outer
 * refinement (ir_max_outer = 0), BiCGSTAB publishes through grid barriers
 * (the same iterates as the neighbour waits), no phase profile.  What the
 * solve reads must exist (else MF_E_ARG): the FP64 layout and, for
 * MF_MIXED, the mixed layout of opts->tau/allow_mask (mf_matrix_layout, or
 * an earlier mf_solve*), and, where opts->reorder picks the row-reordered
 * copy as mf_solve_device does, that copy (an earlier mf_solve_device with
 * the same opts->grid builds it).  relres_true is the kernel's final FP64
 * true-residual pass.  Calls that overlap on one GPU race on its array. */
int mf_solve_device_noalloc(mf_matrix m, const double *d_b, double *d_x, const mf_opts *o,
                            mf_report *rep);

/* Vendor-library baseline on the same matrix: CG or BiCGSTAB (opts->method)
 * in FP64 as the reference's hipSPARSE drivers run it
 * (Performance-Comparison/hipSPARSE_CG.cu:182, hipSPARSE_BiCGSTAB.cu:282):
 * host loop, rocSPARSE CSR SpMV (adaptive), rocBLAS dots, ||r|| read by the
 * host every iteration; to opts->tol in true relative residual (opts->maxit
 * iterations at most).  rep->ms_solve covers the loop and the final check.
 * Device buffers of rowA entries (mf_vendor_solve_device) or host buffers. */
int mf_vendor_solve_device(mf_matrix m, const double *d_b, double *d_x, const mf_opts *o, mf_report *rep);
int mf_vendor_solve(mf_matrix m, const double *b, double *x, const mf_opts *o, mf_report *rep);

/* Tiled SpMV y = A x.  mixed: 0 = FP64 layout, 1 = mixed layout of the last
 * mf_solve/mf_matrix_layout call.  Host arrays: x has tilen*B entries (zero
 * padded by the call from colA), y receives tilem*B. */
int mf_matrix_layout(mf_matrix m, double tau, int allow_mask);
int mf_spmv(mf_matrix m, int mixed, const double *x, double *y);
/* Times reps launches of the standalone SpMV kernel (inputs resident) and
 * returns the mean kernel time in ms (device events). */
int mf_spmv_bench(mf_matrix m, int mixed, int reps, double *ms);

/* Batched solver object: nsys independent square SPD systems (CSR), each
 * with nrhs (1..8) right-hand sides, solved by CG in one persistent launch
 * captured in a hipGraph (the vectors in HBM, or, with opts->vectors =
 * MF_VECTORS_ONCHIP and 2 to 8 right-hand sides, in the LDS of gangs of
 * workgroups).  Tiles are built on the GPU per system and joined
 * into one block-diagonal layout; opts->block picks the tile size (8, 16 or
 * 32; 0 = 16), opts->grid the target number of workgroups (0 = as many as
 * can be resident: one per CU with 8 columns, where the kernel needs every
 * register), opts->precision/tau/allow_mask the per-tile classes (every system reaches
 * opts->tol in true relative residual through FP64 true-residual passes). */
typedef struct mf_batch_s *mf_batch;
typedef struct {
    int nsys, nrhs, B, G;          /* systems, columns, tile size, workgroups */
    long long rows, tilenum;       /* padded rows and tiles of the joined layout */
    double bytes_alg;              /* per iteration: matrix once + 16 n per column */
    double ms_setup;
    long long nnz_per_code[5];
    int mixed;                     /* inner layout has per-tile classes */
    double tau;                    /* their rounding threshold */
    char tuned[512];               /* the autotuner's trials and choice */
    int index_width;               /* slab index bytes shared by the systems (1, 2, 4) */
    int on_chip;                   /* 1: every system runs on a gang of workgroups that keeps
                                      its vectors in registers and LDS (8 columns, systems up
                                      to ~1024 rows per CU); 0: vectors in HBM */
    int fell_back;                 /* systems of the last solve that left lossy tiles */
} mf_batch_info_t;
int mf_batch_create(int nsys, const int *n, const int *const *rowptr, const int *const *colidx,
                    const double *const *val, int nrhs, const mf_opts *o, mf_batch *out);
/* b[s][k*n_s + i] in, x[s][k*n_s + i] out (host); per-system iterations and
 * max-over-columns true relative residual; rep->ms_solve = graph launch time
 * (right-hand sides already in HBM). */
int mf_batch_solve(mf_batch bt, const double *const *b, double *const *x, int *iters, double *relres,
                   mf_report *rep);
int mf_batch_info(mf_batch bt, mf_batch_info_t *info);
/* Autotuner: trial solves of b pick the workgroups per system (weighted by
 * measured iterations), the tile size (8/16/32) and, when o->precision is
 * MF_MIXED, the tile threshold; the fastest converged choice stays in the
 * batch.  rep describes the last trial. */
int mf_batch_tune(mf_batch bt, const double *const *b, mf_report *rep);
void mf_batch_destroy(mf_batch bt);

/* Batched: nsys independent square systems, each with nrhs right-hand sides
 * (column-major: b[s][k*n_s + i]), solved by one persistent launch captured
 * in a hipGraph.  Matrices are given as CSR.  x likewise.  Per-system
 * iterations and true relative residuals (max over RHS) are returned. */
int mf_solve_batched(int nsys, const int *n, const int *const *rowptr, const int *const *colidx,
                     const double *const *val, int nrhs, const double *const *b, double *const *x,
                     const mf_opts *o, int *iters, double *relres, mf_report *rep);

/* Memory-Cost classifier (MC/memory_Mille_feillue.cu:866-918) over the built
 * tiles: counts[0..3] = fp8, fp16, fp32, fp64 nonzeros; returns memory_MB of
 * the values. */
double mf_mc_classify(mf_matrix m, long long counts[4]);

/* Memory footprint: the reference's Memory-Cost figure for its tile format
 * (MC/memory_Mille_feillue.cu:659-922: tile metadata, compact row lists,
 * values at the width of their magnitude bin; memory_CP = 8 B per row
 * segment) beside the CSR FP64 bytes and this library's FP64 and per-tile
 * mixed layouts at (tau, allow_mask) (slabs with their padding + headers).
 * The mixed layout is built (and kept) by the call. */
typedef struct {
    double ref_memory_MB, ref_memory_CP_MB;
    double csr_fp64_MB;
    double layout_fp64_MB, layout_mixed_MB;
    int index_width;            /* slab index bytes of the mixed layout */
    long long nnz_per_code[5];  /* nonzeros per code in the mixed layout */
} mf_memory;
int mf_memory_report(mf_matrix m, double tau, int allow_mask, mf_memory *out);

/* Host preprocessing of the reference schedule, exported for the CSV rows:
 * compact row lists (PC/Mille-feuille_CG_AMD.cu:1421-1480) and the two-pointer
 * warp partition (:1483-1735; mode 1 = AMD autotune, mode 0 = fixed
 * each_block_nnz as PC/Mille-feuille_CG_NVIDIA.cu:4328-4420).  Arrays are
 * malloc'ed; release with mf_free. */
typedef struct {
    int cnt_non_new;
    int *nonzero_row_new;
    unsigned char *blockrowid_new;
    unsigned char *blockcsr_ptr_new;
    int *block_signal;
    int *tile_rowidx;
} mf_compact;
int mf_compact_rows(const mf_ref_tiles *t, mf_compact *out);
typedef struct {
    int index, each_block_nnz, block_per_warp, vector_each_warp_32, vector_total_32;
    int *balance_tile_ptr, *index_each_block, *row_each_block, *non_each_block_offset,
        *balance_tile_ptr_shared_end;
} mf_partition;
int mf_partition_build(const mf_ref_tiles *t, int nnzR, int block_nnz, int sm_num, int mode,
                       mf_partition *out);
void mf_free(void *p);

/* Sizes of the structs of this header as compiled into the library, in the
 * order mf_ref_tiles, mf_matrix_info, mf_opts, mf_report, mf_batch_info_t,
 * mf_memory, mf_compact, mf_partition, Tile_matrix (at most n written);
 * returns the number of structs.  Lets a binding check its mirrors of them (no counterpart in
 * the reference, which has no binding). */
int mf_abi_sizes(size_t *sizes, int n);

/* Synthetic matrices of the benchmark configurations, as host CSR
 * (malloc'ed; mf_free).  Rows in lexicographic order, x fastest. */
int mf_gen_laplace2d(int nx, int ny, int *n, int **rowptr, int **colidx, double **val);
int mf_gen_poisson3d(int nx, int ny, int nz, int *n, int **rowptr, int **colidx, double **val);
int mf_gen_convdiff3d(int m, double peclet, double sigma, unsigned long long seed, int *n,
                      int **rowptr, int **colidx, double **val);
int mf_gen_random_spd(int n, double zipf_a, int lmin, int lmax, unsigned long long seed, int *nnz,
                      int **rowptr, int **colidx, double **val);
int mf_gen_stencil_lognormal(int dim, int n_target, double sigma, unsigned long long seed, int *n,
                             int **rowptr, int **colidx, double **val);
/* Portable generator (mt19937_64 with explicit transforms). */
int mf_gen_uniform(unsigned long long seed, int count, double lo, double hi, double *out);
int mf_gen_normal(unsigned long long seed, int count, double *out);

#ifdef __cplusplus
}
#endif
#endif
