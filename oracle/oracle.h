/*
 * oracle.h -- CPU restatement of the Mille-feuille reference path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or, in bench.py, as the timed CPU baseline).  The product library
 * (mille-feillue_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates
 * (PC/ = Performance-Comparison/, MP/ = Mixed-Precision/, MC/ = Memory-Cost/).
 *
 * Parity pinning (what checks each function, beyond this restatement):
 *   orc_mmio_*            oracle/_ref build of PC/mmio_highlevel.h on shipped
 *                         .mtx files; nnzR = 4054 known answer for 1138_bus.
 *   orc_tile_create       oracle/_ref build of PC/csr2block.h (bit-exact) on
 *                         shipped + synthetic matrices; brute-force dense
 *                         16x16 reconstruction.
 *   orc_encode/decode     oracle/_ref build of PC/encode.h; round trip.
 *   orc_blockspmv_offsets oracle/_ref build of PC/blockspmv_cpu.h.
 *   orc_partition         known answer index=149 for 1138_bus at each_nnz=16
 *                         (MP/data/cg_syncfree_reduce_2757.csv); invariants.
 *   orc_mc_classify       closed form on hand-built values at the thresholds.
 *   orc_e4m3/f16          brute force over all codes (nearest, ties-to-even).
 *   orc_tile_classes      hand-built values whose exact formats are known.
 *   orc_tile_spmv         scipy CSR product.
 *   orc_tile_spmv_mixed   all-FP64 codes reduce to orc_tile_spmv; lossless
 *                         low codes agree with scipy's product to FP32
 *                         accumulation error.
 *   orc_cg / orc_bicgstab scipy.sparse.linalg.spsolve solution; textbook
 *                         properties (CG converges in <= n steps on tiny SPD,
 *                         exact after k steps with k distinct eigenvalues).
 *   orc_petsc_cg/_bcgs    PETSc itself cannot be built here; pinned by what the
 *                         mathematics fixes: KSPCG / KSPBCGS and the textbook
 *                         Hestenes-Stiefel CG / van der Vorst BiCGSTAB have the
 *                         same iterates in exact arithmetic, so their residual
 *                         histories agree to rounding with orc_cg / orc_bicgstab
 *                         over the first iterations; finite termination on
 *                         matrices with two or three distinct eigenvalues.
 *   orc_compact_rows      stops and listed rows brute-forced from the reference
 *                         tile arrays (golden Blockcsr_Ptr, uint8 wrap included)
 *                         and checked against the CSR row counts mod 256.
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int B; /* BLOCK_SIZE */
    int rowA, colA;
    int tilem, tilen, tilenum;
    int csrsize, csrptrlen;
    int *tile_ptr;                    /* tilem+1 */
    int *tile_columnidx;              /* tilenum */
    int *tile_nnz;                    /* tilenum+1, scanned */
    char *Format;                     /* tilenum */
    int *blknnz;                      /* tilenum+1, scanned */
    unsigned char *blknnznnz;         /* tilenum+1, unscanned mod 256 */
    int *csr_offset;                  /* tilenum+1 */
    int *csrptr_offset;               /* tilenum+1 */
    double *Blockcsr_Val;             /* csrsize */
    float *Blockcsr_Val_Low;          /* csrsize */
    unsigned char *Blockcsr_Ptr;      /* csrptrlen */
    unsigned char *Tile_csr_Col;      /* csrsize */
    unsigned char *csr_compressedIdx; /* ceil(csrsize/2), zero as in reference */
} orc_tiles;

/* ---- A1 Matrix Market (PC/mmio_highlevel.h:17, :275) ---- */
int orc_mmio_info(int *m, int *n, int *nnz, int *isSymmetric, const char *filename);
int orc_mmio_data(int *rowptr, int *colidx, double *val, const char *filename);

/* ---- A2-A4 tile layout (PC/csr2block.h:267, PC/utils.h:34-65) ---- */
int orc_tile_create(int B, int rowA, int colA, int nnzA, const int *rowptr,
                    const int *colidx, const double *val, const float *val_low,
                    orc_tiles *out);
void orc_tiles_free(orc_tiles *t);
void orc_exclusive_scan(int *input, int length);
void orc_exclusive_scan_char(unsigned char *input, int length);

/* ---- A5 nibble codec (PC/encode.h:29, :52) ---- */
void orc_encode(const unsigned char *str, unsigned char *res, int str_len, int offset);
void orc_decode(const unsigned char *str, unsigned char *res, int res_len);

/* ---- A6 per-tile offsets and row-block split (PC/blockspmv_cpu.h:2-112) ---- */
int orc_blockspmv_offsets(const orc_tiles *t, int *ptroffset1, int *ptroffset2,
                          unsigned int **blkcoostylerowidx, int **colstart, int **colstop);

/* ---- A7 compact row lists (PC/Mille-feuille_CG_AMD.cu:1421-1480) ---- */
typedef struct {
    int cnt_non_new;
    int *nonzero_row_new;            /* tilenum+1 scanned */
    unsigned char *blockrowid_new;   /* cnt+1 */
    unsigned char *blockcsr_ptr_new; /* cnt+1 */
    int *block_signal;               /* tilem+1 */
    int *tile_rowidx;                /* tilenum */
} orc_compact;
int orc_compact_rows(const orc_tiles *t, orc_compact *out);
void orc_compact_free(orc_compact *c);

/* ---- A8 two-pointer warp partition (PC/Mille-feuille_CG_AMD.cu:1483-1735;
 *      fixed-threshold form PC/Mille-feuille_CG_NVIDIA.cu:4328-4420) ---- */
typedef struct {
    int index;          /* number of warp groups */
    int each_block_nnz; /* chosen nnz threshold */
    int block_per_warp;
    int vector_each_warp_32, vector_total_32; /* valid when index < tilem */
    int *balance_tile_ptr;          /* index+1 */
    int *index_each_block;          /* tilenum+1 (permuted tile ids) */
    int *row_each_block;            /* tilenum+1 */
    int *non_each_block_offset;     /* tilenum+1 */
    int *balance_tile_ptr_shared_end; /* index+1 */
} orc_partition_t;
/* mode 0: fixed each_block_nnz = block_nnz (NVIDIA); mode 1: AMD autotune. */
int orc_partition(const orc_tiles *t, int nnzR, int block_nnz, int sm_num, int mode,
                  orc_partition_t *out);
void orc_partition_free(orc_partition_t *p);

/* ---- A14 Memory-Cost magnitude classifier (MC/memory_Mille_feillue.cu:866-918) ---- */
/* counts[0..3] = fp8, fp16, fp32, fp64 nonzeros; returns memory_MB of the values. */
double orc_mc_classify(const orc_tiles *t, const orc_compact *c, long long counts[4]);

/* ---- SpMV restatements ---- */
void orc_csr_spmv(int n, const int *rowptr, const int *colidx, const double *val,
                  const double *x, double *y);
/* FP64 tile SpMV over Blockcsr_Ptr/Tile_csr_Col, per row: acc = fma(v, x, acc)
 * in tile order then element order (PC/Mille-feuille_CG_AMD.cu:1966-1993). */
void orc_tile_spmv(const orc_tiles *t, const double *x, double *y);

/* ---- Low-precision formats (OCP FP8 E4M3; IEEE binary16) ---- */
unsigned char orc_e4m3_from_double(double v); /* RNE, saturate to +-448 */
double orc_e4m3_to_double(unsigned char c);
unsigned short orc_f16_from_double(double v); /* RNE, overflow -> inf */
double orc_f16_to_double(unsigned short h);

/* Per-tile storage class (codes follow the reference enumeration of
 * MP/Mille-feuille_CG.cu:257-352: 0=FP64, 4=FP32, 3=FP16, 2=FP8).  A tile takes
 * the narrowest format whose normal range covers every value of the tile and
 * whose rounding error is <= tau*|v| for every value (tau=0: lossless only).
 * allow_mask bit k enables code k. */
void orc_tile_classes(const orc_tiles *t, double tau, int allow_mask, unsigned char *codes);
/* Mixed SpMV (the order the MI355X layout computes in): FP64 tiles accumulate
 * fma in FP64; then FP32, FP16 and FP8 tiles, class by class in tile order,
 * accumulate fmaf in FP32 over (float)x; y = acc64 + (double)acc32. */
void orc_tile_spmv_mixed(const orc_tiles *t, const unsigned char *codes, const double *x,
                         double *y);

/* ---- Solvers (FP64) ---- */
typedef struct {
    int iterations;
    double relres_true;    /* ||b - A x|| / ||b|| at exit */
    double relres_updated; /* recursively updated residual at exit */
    int converged;
    /* input, optional: when hist is set, hist[k] receives the updated
     * residual norm ||r_k|| after iteration k (hist[0] = ||r_0||) for
     * k < hist_cap; hist_len = entries written */
    double *hist;
    int hist_cap, hist_len;
} orc_report;
/* Hestenes-Stiefel CG, x0 = 0 (PC/Mille-feuille_CG_AMD.cu:607-912 algebra).
 * Stop: updated ||r|| <= tol||b|| and then true ||b-Ax|| <= tol||b||.
 * use_tiles: SpMV through orc_tile_spmv (else CSR).  nthreads > 1: OpenMP. */
int orc_cg(int n, const int *rowptr, const int *colidx, const double *val,
           const orc_tiles *tiles, const double *b, double *x, double tol, int maxit,
           int nthreads, orc_report *rep);
/* van der Vorst BiCGSTAB, x0 = 0 (PC/Mille-feuille_BiCGSTAB_AMD.cu:61-425 algebra,
 * with the omega reduction of :735 done as intended). */
int orc_bicgstab(int n, const int *rowptr, const int *colidx, const double *val,
                 const orc_tiles *tiles, const double *b, double *x, double tol, int maxit,
                 int nthreads, orc_report *rep);
/* Unpreconditioned KSPSolve_CG / KSPSolve_BCGS loops (PC/PETSc/src/ksp/ksp/impls/
 * cg/cg.c:112, bcgs/bcgs.c:19) over a CSR MatMult; convergence test
 * rnorm < rtol * rnorm0 as KSPConvergedDefault. */
int orc_petsc_cg(int n, const int *rowptr, const int *colidx, const double *val,
                 const double *b, double *x, double rtol, int maxit, int nthreads,
                 orc_report *rep);
int orc_petsc_bcgs(int n, const int *rowptr, const int *colidx, const double *val,
                   const double *b, double *x, double rtol, int maxit, int nthreads,
                   orc_report *rep);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
