/*
 * oracle.c -- CPU restatement of the Mille-feuille reference path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.  Each function follows the reference file:line it names.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------- */
/* A1: Matrix Market reader.  Restates PC/mmio_highlevel.h:17-269 (info) and
 * :275-593 (data) with the banner/size parsing of PC/mmio.h:398-590:
 * entries read with "%d %d %lg" (pattern -> 1.0, integer -> double, complex ->
 * real part), 1-based -> 0-based, symmetric/hermitian off-diagonals mirrored,
 * CSR by counting sort in file order (no column sort, duplicates kept). */

typedef struct {
    int m, n, nz_report;
    int is_real, is_complex, is_integer, is_pattern, is_sym;
} mm_header;

static int orc_mm_header(FILE *f, mm_header *h)
{
    char line[1030], banner[64], mtx[64], crd[64], dtype[64], storage[64];
    memset(h, 0, sizeof(*h));
    if (!fgets(line, sizeof line, f)) return -2;
    if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dtype, storage) != 5) return -2;
    for (char *p = mtx; *p; ++p) *p = (char)tolower((unsigned char)*p);
    for (char *p = crd; *p; ++p) *p = (char)tolower((unsigned char)*p);
    for (char *p = dtype; *p; ++p) *p = (char)tolower((unsigned char)*p);
    for (char *p = storage; *p; ++p) *p = (char)tolower((unsigned char)*p);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0) return -2;
    if (strcmp(mtx, "matrix") != 0) return -2;
    if (strcmp(crd, "coordinate") != 0 && strcmp(crd, "array") != 0) return -2;
    if (!strcmp(dtype, "real")) h->is_real = 1;
    else if (!strcmp(dtype, "complex")) h->is_complex = 1;
    else if (!strcmp(dtype, "pattern")) h->is_pattern = 1;
    else if (!strcmp(dtype, "integer")) h->is_integer = 1;
    else return -2;
    if (!strcmp(storage, "symmetric") || !strcmp(storage, "hermitian")) h->is_sym = 1;
    else if (strcmp(storage, "general") != 0 && strcmp(storage, "skew-symmetric") != 0) return -2;
    /* size line: skip '%' comment lines (PC/mmio.h:568-590) */
    do {
        if (!fgets(line, sizeof line, f)) return -4;
    } while (line[0] == '%');
    if (sscanf(line, "%d %d %d", &h->m, &h->n, &h->nz_report) == 3) return 0;
    int got;
    do {
        got = fscanf(f, "%d %d %d", &h->m, &h->n, &h->nz_report);
        if (got == EOF) return -4;
    } while (got != 3);
    return 0;
}

static int orc_mm_read(const char *filename, mm_header *h, int **ri, int **ci, double **vv)
{
    FILE *f = fopen(filename, "r");
    if (!f) return -1;
    int rc = orc_mm_header(f, h);
    if (rc) { fclose(f); return rc; }
    int nz = h->nz_report;
    *ri = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    *ci = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    *vv = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
    for (int i = 0; i < nz; i++) {
        int a = 0, b = 0, iv = 0;
        double fv = 0.0, fim = 0.0;
        if (h->is_real) {
            if (fscanf(f, "%d %d %lg\n", &a, &b, &fv) != 3) {}
        } else if (h->is_complex) {
            if (fscanf(f, "%d %d %lg %lg\n", &a, &b, &fv, &fim) != 4) {}
        } else if (h->is_integer) {
            if (fscanf(f, "%d %d %d\n", &a, &b, &iv) != 3) {}
            fv = iv;
        } else {
            if (fscanf(f, "%d %d\n", &a, &b) != 2) {}
            fv = 1.0;
        }
        (*ri)[i] = a - 1;
        (*ci)[i] = b - 1;
        (*vv)[i] = fv;
    }
    fclose(f);
    return 0;
}

static int *orc_mm_rowcounts(const mm_header *h, const int *ri, const int *ci)
{
    int *cnt = (int *)calloc((size_t)h->m + 1, sizeof(int));
    for (int i = 0; i < h->nz_report; i++) cnt[ri[i]]++;
    if (h->is_sym)
        for (int i = 0; i < h->nz_report; i++)
            if (ri[i] != ci[i]) cnt[ci[i]]++;
    orc_exclusive_scan(cnt, h->m + 1);
    return cnt;
}

int orc_mmio_info(int *m, int *n, int *nnz, int *isSymmetric, const char *filename)
{
    mm_header h;
    int *ri, *ci;
    double *vv;
    int rc = orc_mm_read(filename, &h, &ri, &ci, &vv);
    if (rc) return rc;
    int *cnt = orc_mm_rowcounts(&h, ri, ci);
    *m = h.m;
    *n = h.n;
    *nnz = cnt[h.m];
    *isSymmetric = h.is_sym;
    free(cnt); free(ri); free(ci); free(vv);
    return 0;
}

int orc_mmio_data(int *rowptr, int *colidx, double *val, const char *filename)
{
    mm_header h;
    int *ri, *ci;
    double *vv;
    int rc = orc_mm_read(filename, &h, &ri, &ci, &vv);
    if (rc) return rc;
    int *cnt = orc_mm_rowcounts(&h, ri, ci);
    memcpy(rowptr, cnt, sizeof(int) * ((size_t)h.m + 1));
    int *fill = (int *)calloc((size_t)h.m + 1, sizeof(int));
    /* file-order scatter; a symmetric off-diagonal goes to row i, then to row j
     * (PC/mmio_highlevel.h:533-566) */
    for (int i = 0; i < h.nz_report; i++) {
        int r = ri[i], c = ci[i];
        int off = rowptr[r] + fill[r]++;
        colidx[off] = c;
        val[off] = vv[i];
        if (h.is_sym && r != c) {
            off = rowptr[c] + fill[c]++;
            colidx[off] = r;
            val[off] = vv[i];
        }
    }
    free(fill); free(cnt); free(ri); free(ci); free(vv);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* A4: in-place exclusive scans (PC/utils.h:34-65); the char form wraps mod 256 */
void orc_exclusive_scan(int *input, int length)
{
    if (length <= 1) return;
    int run = 0;
    for (int i = 0; i < length; i++) {
        int v = input[i];
        input[i] = run;
        run += v;
    }
}

void orc_exclusive_scan_char(unsigned char *input, int length)
{
    if (length <= 1) return;
    unsigned char run = 0;
    for (int i = 0; i < length; i++) {
        unsigned char v = input[i];
        input[i] = run;
        run = (unsigned char)(run + v);
    }
}

/* ------------------------------------------------------------------------- */
/* A2/A3: Tile_create (PC/csr2block.h:267-422) restated.  The four OpenMP passes
 * of the reference are folded into one pass per tile row; the linear tile
 * search of convert_step4 (:204-221) is replaced by a tile-column lookup table,
 * which selects the same tile.  uint8 wrap of the per-row counters
 * (convert_step2 :85) and of the row-start scan (:237) is kept, and the copy
 * loop of :239-253 is executed as written so that wrapped cases match too. */
static int orc_cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int orc_tile_create(int B, int rowA, int colA, int nnzA, const int *rowptr,
                    const int *colidx, const double *val, const float *val_low,
                    orc_tiles *t)
{
    (void)nnzA;
    memset(t, 0, sizeof(*t));
    t->B = B;
    t->rowA = rowA;
    t->colA = colA;
    t->tilem = rowA % B == 0 ? rowA / B : rowA / B + 1;
    t->tilen = colA % B == 0 ? colA / B : colA / B + 1;
    int tilem = t->tilem, tilen = t->tilen;
    t->tile_ptr = (int *)calloc((size_t)tilem + 1, sizeof(int));
    int *slot = (int *)malloc(sizeof(int) * (size_t)(tilen > 0 ? tilen : 1));
    for (int j = 0; j < tilen; j++) slot[j] = -1;

    /* step1 (:6-41): distinct tile columns per tile row */
    for (int bi = 0; bi < tilem; bi++) {
        int start = bi * B, end = bi == tilem - 1 ? rowA : (bi + 1) * B;
        int cnt = 0;
        for (int j = rowptr[start]; j < rowptr[end]; j++) {
            int jc = colidx[j] / B;
            if (slot[jc] < 0) { slot[jc] = 1; cnt++; }
        }
        for (int j = rowptr[start]; j < rowptr[end]; j++) slot[colidx[j] / B] = -1;
        t->tile_ptr[bi] = cnt;
    }
    orc_exclusive_scan(t->tile_ptr, tilem + 1);
    int tilenum = t->tilenum = t->tile_ptr[tilem];

    t->tile_columnidx = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    t->tile_nnz = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    unsigned char *tile_csr_ptr = (unsigned char *)calloc((size_t)tilenum * B + 1, 1);

    /* step2 (:43-107): ascending tile columns, nnz per tile, uint8 per-row counts */
    unsigned char *flag = (unsigned char *)calloc((size_t)(tilen > 0 ? tilen : 1), 1);
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = bi == tilem - 1 ? rowA - (tilem - 1) * B : B;
        int start = bi * B;
        int end = bi == tilem - 1 ? rowA : (bi + 1) * B;
        int pre = t->tile_ptr[bi], count = 0;
        for (int j = rowptr[start]; j < rowptr[end]; j++) {
            int jc = colidx[j] / B;
            if (!flag[jc]) { flag[jc] = 1; t->tile_columnidx[pre + count++] = jc; }
        }
        qsort(t->tile_columnidx + pre, (size_t)count, sizeof(int), orc_cmp_int);
        for (int c = 0; c < count; c++) { slot[t->tile_columnidx[pre + c]] = pre + c; flag[t->tile_columnidx[pre + c]] = 0; }
        for (int ri = 0; ri < rowlen; ri++)
            for (int j = rowptr[start + ri]; j < rowptr[start + ri + 1]; j++) {
                int tid = slot[colidx[j] / B];
                t->tile_nnz[tid]++;
                tile_csr_ptr[(size_t)tid * B + ri]++; /* uint8 wrap */
            }
        for (int c = 0; c < count; c++) slot[t->tile_columnidx[pre + c]] = -1;
    }
    orc_exclusive_scan(t->tile_nnz, tilenum + 1);

    /* step3 (:109-150): Format = 0, blknnz, csr_offset, csrptr_offset */
    t->Format = (char *)calloc((size_t)tilenum + 1, 1);
    t->blknnz = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    t->blknnznnz = (unsigned char *)calloc((size_t)tilenum + 1, 1);
    t->csr_offset = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    t->csrptr_offset = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = bi == tilem - 1 ? rowA - (tilem - 1) * B : B;
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            int nnztmp = t->tile_nnz[tid + 1] - t->tile_nnz[tid];
            t->blknnz[tid] = nnztmp;
            t->csr_offset[tid] = nnztmp;
            t->csrptr_offset[tid] = rowlen;
        }
    }
    orc_exclusive_scan(t->csr_offset, tilenum + 1);
    orc_exclusive_scan(t->csrptr_offset, tilenum + 1);
    /* sizes (:349-371) */
    long long csrsize = 0, csrptrlen = 0;
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = bi == tilem - 1 ? rowA - (tilem - 1) * B : B;
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            csrsize += t->blknnz[tid];
            csrptrlen += rowlen;
        }
    }
    t->csrsize = (int)csrsize;
    t->csrptrlen = (int)csrptrlen;
    for (int i = 0; i < tilenum + 1; i++) t->blknnznnz[i] = (unsigned char)t->blknnz[i];
    orc_exclusive_scan(t->blknnz, tilenum + 1);

    t->Blockcsr_Val = (double *)calloc((size_t)csrsize + 1, sizeof(double));
    t->Blockcsr_Val_Low = (float *)calloc((size_t)csrsize + 1, sizeof(float));
    t->Blockcsr_Ptr = (unsigned char *)calloc((size_t)csrptrlen + 1, 1);
    t->Tile_csr_Col = (unsigned char *)calloc((size_t)csrsize + 1, 1);
    int csize = t->csrsize % 2 == 0 ? t->csrsize / 2 : t->csrsize / 2 + 1;
    t->csr_compressedIdx = (unsigned char *)calloc((size_t)csize + 1, 1);

    /* step4 (:152-265): scatter each tile row's CSR entries into tiles in CSR
     * order, then copy row by row using the uint8 row starts */
    int nnz_temp = 0;
    for (int bi = 0; bi < tilem; bi++) {
        int start = bi * B, end = bi == tilem - 1 ? rowA : (bi + 1) * B;
        if (rowptr[end] - rowptr[start] > nnz_temp) nnz_temp = rowptr[end] - rowptr[start];
    }
    size_t tmpn = (size_t)nnz_temp + 512;
    unsigned char *col_tmp = (unsigned char *)malloc(tmpn);
    double *val_tmp = (double *)malloc(tmpn * sizeof(double));
    float *low_tmp = (float *)malloc(tmpn * sizeof(float));
    int *tcount = (int *)malloc(sizeof(int) * (size_t)(tilen > 0 ? tilen : 1) + 4);
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = bi == tilem - 1 ? rowA - (tilem - 1) * B : B;
        int start = bi * B, end = bi == tilem - 1 ? rowA : (bi + 1) * B;
        int t0 = t->tile_ptr[bi], t1 = t->tile_ptr[bi + 1];
        int seg = rowptr[end] - rowptr[start];
        memset(col_tmp, 0, (size_t)seg + 512);
        memset(val_tmp, 0, ((size_t)seg + 512) * sizeof(double));
        memset(low_tmp, 0, ((size_t)seg + 512) * sizeof(float));
        for (int tid = t0; tid < t1; tid++) { slot[t->tile_columnidx[tid]] = tid; tcount[tid - t0] = 0; }
        for (int j = rowptr[start]; j < rowptr[end]; j++) {
            int jc = colidx[j] / B;
            int tid = slot[jc];
            int bi2 = tid - t0;
            int pre = t->tile_nnz[tid] - t->tile_nnz[t0];
            val_tmp[pre + tcount[bi2]] = val[j];
            low_tmp[pre + tcount[bi2]] = val_low ? val_low[j] : (float)val[j];
            col_tmp[pre + tcount[bi2]] = (unsigned char)(colidx[j] - jc * B);
            tcount[bi2]++;
        }
        for (int tid = t0; tid < t1; tid++) {
            slot[t->tile_columnidx[tid]] = -1;
            int pre = t->tile_nnz[tid] - t->tile_nnz[t0];
            int nnztmp = t->tile_nnz[tid + 1] - t->tile_nnz[tid];
            int offset = t->csr_offset[tid];
            int ptr_offset = t->csrptr_offset[tid];
            unsigned char *ptr = tile_csr_ptr + (size_t)tid * B;
            orc_exclusive_scan_char(ptr, rowlen);
            for (int ri = 0; ri < rowlen; ri++) {
                int s = ptr[ri];
                int e = ri == rowlen - 1 ? nnztmp : ptr[ri + 1];
                for (int k = s; k < e; k++) {
                    t->Blockcsr_Val[offset + k] = val_tmp[pre + k];
                    t->Blockcsr_Val_Low[offset + k] = low_tmp[pre + k];
                    t->Tile_csr_Col[offset + k] = col_tmp[pre + k];
                }
                t->Blockcsr_Ptr[ptr_offset + ri] = ptr[ri];
            }
        }
    }
    free(col_tmp); free(val_tmp); free(low_tmp); free(tcount);
    free(flag); free(slot); free(tile_csr_ptr);
    return 0;
}

void orc_tiles_free(orc_tiles *t)
{
    free(t->tile_ptr); free(t->tile_columnidx); free(t->tile_nnz); free(t->Format);
    free(t->blknnz); free(t->blknnznnz); free(t->csr_offset); free(t->csrptr_offset);
    free(t->Blockcsr_Val); free(t->Blockcsr_Val_Low); free(t->Blockcsr_Ptr);
    free(t->Tile_csr_Col); free(t->csr_compressedIdx);
    memset(t, 0, sizeof(*t));
}

/* ------------------------------------------------------------------------- */
/* A5: nibble codec (PC/encode.h:29-72): byte = (c[2k] << 4) + c[2k+1]; an odd
 * tail stores c << 4; decode masks with 240 and 15. */
void orc_encode(const unsigned char *str, unsigned char *res, int str_len, int offset)
{
    int count = offset;
    int pairs = str_len / 2;
    for (int k = 0; k < pairs; k++) res[count++] = (unsigned char)((str[2 * k] << 4) + str[2 * k + 1]);
    if (str_len % 2) res[count] = (unsigned char)(str[str_len - 1] << 4);
}

void orc_decode(const unsigned char *str, unsigned char *res, int res_len)
{
    int pairs = res_len / 2;
    for (int k = 0; k < pairs; k++) {
        res[2 * k] = (unsigned char)((str[k] & 240) >> 4);
        res[2 * k + 1] = (unsigned char)(str[k] & 15);
    }
    if (res_len % 2) res[res_len - 1] = (unsigned char)(str[res_len / 2] >> 4);
}

/* ------------------------------------------------------------------------- */
/* A6: PC/blockspmv_cpu.h:2-112 -- row-block split at PREFETCH_SMEM_TH = 4 with
 * bit 31 marking split tile rows, and the per-tile offsets ptroffset1
 * (nonzeros before the tile) and ptroffset2 (row slots before the tile).
 * Returns rowblkblock. */
int orc_blockspmv_offsets(const orc_tiles *t, int *ptroffset1, int *ptroffset2,
                          unsigned int **blkrow, int **colstart, int **colstop)
{
    const int TH = 4;
    int rbb = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int nb = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
        rbb += nb <= TH ? 1 : (int)ceil((double)nb / (double)TH);
    }
    unsigned int *br = (unsigned int *)calloc((size_t)rbb + 1, sizeof(unsigned int));
    int *cs = (int *)calloc((size_t)rbb + 1, sizeof(int));
    int *ce = (int *)calloc((size_t)rbb + 1, sizeof(int));
    int k = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int nb = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
        if (nb <= TH) {
            br[k++] = (unsigned int)bi;
        } else {
            int nloc = (int)ceil((double)nb / (double)TH);
            int len = (int)ceil((double)nb / (double)nloc);
            for (int q = 0; q < nloc; q++) {
                br[k] = (unsigned int)bi | 0x80000000u;
                cs[k] = t->tile_ptr[bi] + q * len;
                ce[k] = q == nloc - 1 ? t->tile_ptr[bi] + nb : t->tile_ptr[bi] + (q + 1) * len;
                k++;
            }
        }
    }
    int off1 = 0, off2 = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rowlen = bi == t->tilem - 1 ? t->rowA - (t->tilem - 1) * t->B : t->B;
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            ptroffset1[tid] = off1;
            ptroffset2[tid] = off2;
            off1 += t->blknnz[tid + 1] - t->blknnz[tid];
            off2 += rowlen;
        }
    }
    *blkrow = br;
    *colstart = cs;
    *colstop = ce;
    return rbb;
}

/* ------------------------------------------------------------------------- */
/* A7: compact non-empty-row lists (PC/Mille-feuille_CG_AMD.cu:1421-1480).
 * nonzero_row_new[t] = non-empty rows + 1 (scanned); per tile a sentinel
 * (row 0, stop 0) then (row, stop) for each non-empty row; stops are stored
 * as uint8 (a 256-entry tile wraps to 0). */
int orc_compact_rows(const orc_tiles *t, orc_compact *c)
{
    memset(c, 0, sizeof(*c));
    int tilenum = t->tilenum, tilem = t->tilem, B = t->B;
    c->nonzero_row_new = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    c->tile_rowidx = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    c->block_signal = (int *)calloc((size_t)tilem + 1, sizeof(int));
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = bi == tilem - 1 ? t->rowA - (tilem - 1) * B : B;
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            int cc = t->csrptr_offset[tid];
            c->tile_rowidx[tid] = bi;
            for (int ri = 0; ri < rowlen; ri++) {
                int stop = ri == rowlen - 1 ? t->blknnz[tid + 1] - t->blknnz[tid] : t->Blockcsr_Ptr[ri + 1 + cc];
                if (stop != t->Blockcsr_Ptr[cc + ri]) c->nonzero_row_new[tid] += 1;
            }
            c->nonzero_row_new[tid] += 1;
        }
    }
    orc_exclusive_scan(c->nonzero_row_new, tilenum + 1);
    int cnt = c->cnt_non_new = c->nonzero_row_new[tilenum];
    c->blockrowid_new = (unsigned char *)calloc((size_t)cnt + 1, 1);
    c->blockcsr_ptr_new = (unsigned char *)calloc((size_t)cnt + 1, 1);
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = bi == tilem - 1 ? t->rowA - (tilem - 1) * B : B;
        c->block_signal[bi] = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            int cc = t->csrptr_offset[tid];
            int base = c->nonzero_row_new[tid];
            int fl = 0;
            for (int ri = 0; ri < rowlen; ri++) {
                int stop = ri == rowlen - 1 ? t->blknnz[tid + 1] - t->blknnz[tid] : t->Blockcsr_Ptr[ri + 1 + cc];
                if (ri == 0) {
                    c->blockrowid_new[base + fl] = 0;
                    c->blockcsr_ptr_new[base + fl] = 0;
                    fl++;
                }
                if (stop != t->Blockcsr_Ptr[cc + ri]) {
                    c->blockrowid_new[base + fl] = (unsigned char)ri;
                    c->blockcsr_ptr_new[base + fl] = (unsigned char)stop;
                    fl++;
                }
            }
        }
    }
    return 0;
}

void orc_compact_free(orc_compact *c)
{
    free(c->nonzero_row_new); free(c->blockrowid_new); free(c->blockcsr_ptr_new);
    free(c->block_signal); free(c->tile_rowidx);
    memset(c, 0, sizeof(*c));
}

/* ------------------------------------------------------------------------- */
/* A8: two-pointer greedy warp partition.  Tiles stay in natural order; the
 * greedy alternately takes from the front (i) and back (j); a group closes
 * when adding a tile reaches each_block_nnz or the group holds block_per_warp
 * tiles (PC/Mille-feuille_CG_AMD.cu:1556-1590).  mode 1 walks the AMD
 * autotune list {32..4096, nnzR/sm_num, 0} doubling block_per_warp (:1548-1596;
 * the list's 12th entry is read past its 11 initialisers in the reference and
 * is taken as 0 here); mode 0 uses block_nnz directly
 * (PC/Mille-feuille_CG_NVIDIA.cu:4353-4376). */
static int orc_count_groups(const int *non, int tilenum, int ebn, int bpw)
{
    int i = 0, j = tilenum - 1, cnt = 0, index = 1, cb = 0;
    while (i < j) {
        if (non[i] + cnt < ebn && cb + 1 < bpw) { cnt += non[i]; i++; cb++; }
        else { i++; index++; cnt = 0; cb = 0; }
        if (non[j] + cnt < ebn && cb + 1 < bpw) { cnt += non[j]; j--; cb++; }
        else { j--; index++; cnt = 0; cb = 0; }
    }
    return index;
}

int orc_partition(const orc_tiles *t, int nnzR, int block_nnz, int sm_num, int mode,
                  orc_partition_t *p)
{
    memset(p, 0, sizeof(*p));
    int tilenum = t->tilenum, tilem = t->tilem;
    int *non = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    int *row = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    int *idx = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    for (int bi = 0; bi < tilem; bi++)
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            non[tid] = t->blknnz[tid + 1] - t->blknnz[tid];
            row[tid] = bi;
            idx[tid] = tid;
        }
    int bpw = 180, ebn = block_nnz, index = 1;
    if (mode == 1) {
        int list[13] = {16, 32, 64, 96, 128, 256, 512, 1024, 2048, 4096, sm_num ? nnzR / sm_num : 0, 0, 0};
        int done = 0;
        while (!done) {
            for (int k = 0; k < 12; k++) {
                ebn = list[k + 1];
                index = orc_count_groups(non, tilenum, ebn, bpw);
                if (index < sm_num && index != tilem - 1 && index != tilem + 1 && index != tilem) { done = 1; break; }
            }
            if (done) break;
            bpw *= 2;
            if (bpw > (1 << 28)) break; /* the reference loops forever here */
        }
    } else {
        index = orc_count_groups(non, tilenum, ebn, bpw);
    }
    p->each_block_nnz = ebn;
    p->block_per_warp = bpw;
    if (index < tilem && index > 0) {
        int vew16 = (int)ceil((double)tilem / (double)index);
        int tilem_32 = (int)ceil((double)tilem / 2);
        p->vector_each_warp_32 = vew16 * 2;
        int vt = tilem_32 / p->vector_each_warp_32;
        p->vector_total_32 = (vt / 8 + 1) * 8;
    }
    /* second pass: record the order (:1620-1685) */
    int *bt = (int *)calloc((size_t)index + 2, sizeof(int));
    int *idx_new = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    int *row_new = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    int *non_new = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    int i = 0, j = tilenum - 1, cnt = 0, step = 0, cb = 0, ix = 1;
#define TAKE(q) do { idx_new[step] = idx[q]; row_new[step] = row[q]; non_new[step] = non[q]; } while (0)
    while (i < j) {
        if (non[i] + cnt < ebn && cb + 1 < bpw) { cnt += non[i]; TAKE(i); i++; step++; cb++; }
        else { TAKE(i); i++; step++; if (ix <= index) bt[ix] = step; ix++; cnt = 0; cb = 0; }
        if (non[j] + cnt < ebn && cb + 1 < bpw) { cnt += non[j]; TAKE(j); j--; step++; cb++; }
        else { TAKE(j); j--; step++; if (ix <= index) bt[ix] = step; ix++; cnt = 0; cb = 0; }
        if (i == j) { TAKE(j); step++; if (ix <= index) bt[ix] = step; }
        if (i > j) { TAKE(j); if (ix <= index) bt[ix] = step; }
    }
#undef TAKE
    p->index = index;
    /* shared-memory prefix within each group, 256 nonzeros (:1701-1735) */
    int *off = (int *)calloc((size_t)tilenum + 1, sizeof(int));
    int *send = (int *)calloc((size_t)index + 2, sizeof(int));
    for (int g = 0; g <= index; g++) send[g] = bt[g];
    for (int g = 0; g < index; g++) {
        int cn = 0;
        for (int q = bt[g]; q < bt[g + 1]; q++) {
            if (q == bt[g]) off[q] = 0;
            cn += non_new[q];
            if (q != bt[g] && cn <= 256) off[q] = non_new[q - 1] + off[q - 1];
            if (cn > 256) { send[g + 1] = q; break; }
        }
    }
    p->balance_tile_ptr = bt;
    p->index_each_block = idx_new;
    p->row_each_block = row_new;
    p->non_each_block_offset = off;
    p->balance_tile_ptr_shared_end = send;
    free(non_new); free(non); free(row); free(idx);
    return 0;
}

void orc_partition_free(orc_partition_t *p)
{
    free(p->balance_tile_ptr); free(p->index_each_block); free(p->row_each_block);
    free(p->non_each_block_offset); free(p->balance_tile_ptr_shared_end);
    memset(p, 0, sizeof(*p));
}

/* ------------------------------------------------------------------------- */
/* A14: MC/memory_Mille_feillue.cu:866-918 -- per nonzero |v| <= 3.9375 -> fp8,
 * <= 60 -> fp16, <= 6000 -> fp32, else fp64; memory of values
 * 8*fp64 + 2*fp16 + 4*fp32 + 1*fp8 bytes, in MB (/1048576). */
double orc_mc_classify(const orc_tiles *t, const orc_compact *c, long long counts[4])
{
    long long f8 = 0, f16 = 0, f32 = 0, f64 = 0;
    for (int bi = 0; bi < t->tilem; bi++)
        for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
            int off = t->csr_offset[tid];
            for (int ri = c->nonzero_row_new[tid]; ri < c->nonzero_row_new[tid + 1]; ri++)
                for (int rj = c->blockcsr_ptr_new[ri]; rj < c->blockcsr_ptr_new[ri + 1]; rj++) {
                    double v = t->Blockcsr_Val[off + rj];
                    if (v >= -3.9375 && v <= 3.9375) f8++;
                    else if (v >= -60.0 && v <= 60.0) f16++;
                    else if (v >= -6000.0 && v <= 6000.0) f32++;
                    else f64++;
                }
        }
    counts[0] = f8; counts[1] = f16; counts[2] = f32; counts[3] = f64;
    return (double)(8 * f64 + 2 * f16 + 4 * f32 + f8) / 1048576.0;
}

/* ------------------------------------------------------------------------- */
void orc_csr_spmv(int n, const int *rowptr, const int *colidx, const double *val,
                  const double *x, double *y)
{
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s = fma(val[j], x[colidx[j]], s);
        y[i] = s;
    }
}

void orc_tile_spmv(const orc_tiles *t, const double *x, double *y)
{
    int B = t->B;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rowlen = bi == t->tilem - 1 ? t->rowA - (t->tilem - 1) * B : B;
        for (int ri = 0; ri < rowlen; ri++) {
            double acc = 0.0;
            for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
                int off = t->csr_offset[tid], po = t->csrptr_offset[tid];
                int nz = t->blknnz[tid + 1] - t->blknnz[tid];
                int s = t->Blockcsr_Ptr[po + ri];
                int e = ri == rowlen - 1 ? nz : t->Blockcsr_Ptr[po + ri + 1];
                int xo = t->tile_columnidx[tid] * B;
                for (int k = s; k < e; k++) acc = fma(t->Blockcsr_Val[off + k], x[xo + t->Tile_csr_Col[off + k]], acc);
            }
            y[bi * B + ri] = acc;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Low-precision encodings written from the format definitions:
 * OCP 8-bit FP E4M3 (bias 7, no infinities, S.1111.111 = NaN, max 448,
 * min normal 2^-6, subnormal quantum 2^-9) and IEEE 754 binary16.  Both round
 * the double once, to nearest with ties to even. */
static double orc_quantize(double a, int min_exp, int mbits, int *out_exp, double *out_r)
{
    int e2;
    frexp(a, &e2);
    int ue = e2 - 1;
    if (ue < min_exp) ue = min_exp - 1; /* subnormal band */
    double q = ldexp(1.0, (ue < min_exp ? min_exp : ue) - mbits);
    double r = rint(a / q);
    *out_exp = ue;
    *out_r = r;
    return r * q;
}

unsigned char orc_e4m3_from_double(double v)
{
    if (isnan(v)) return 0x7F;
    unsigned char s = signbit(v) ? 0x80 : 0;
    double a = fabs(v);
    if (a >= 448.0) return (unsigned char)(s | 0x7E);
    if (a == 0.0) return s;
    int ue;
    double r;
    orc_quantize(a, -6, 3, &ue, &r);
    if (ue < -6) { /* subnormal: r in [0, 8] */
        if (r >= 8.0) return (unsigned char)(s | (1 << 3));
        return (unsigned char)(s | (int)r);
    }
    int mant = (int)r - 8; /* r in [8, 16] */
    if (mant == 8) { ue += 1; mant = 0; }
    int code = ((ue + 7) << 3) | mant;
    if (code > 0x7E) code = 0x7E;
    return (unsigned char)(s | code);
}

double orc_e4m3_to_double(unsigned char c)
{
    int e = (c >> 3) & 15, m = c & 7;
    double sgn = (c & 0x80) ? -1.0 : 1.0;
    if (e == 15 && m == 7) return NAN;
    if (e == 0) return sgn * ldexp((double)m, -9);
    return sgn * ldexp(1.0 + m / 8.0, e - 7);
}

unsigned short orc_f16_from_double(double v)
{
    if (isnan(v)) return 0x7E00;
    unsigned short s = signbit(v) ? 0x8000 : 0;
    double a = fabs(v);
    if (a >= 65520.0) return (unsigned short)(s | 0x7C00);
    if (a == 0.0) return s;
    int ue;
    double r;
    orc_quantize(a, -14, 10, &ue, &r);
    if (ue < -14) {
        if (r >= 1024.0) return (unsigned short)(s | (1 << 10));
        return (unsigned short)(s | (int)r);
    }
    int mant = (int)r - 1024;
    if (mant == 1024) { ue += 1; mant = 0; }
    if (ue > 15) return (unsigned short)(s | 0x7C00);
    return (unsigned short)(s | ((ue + 15) << 10) | mant);
}

double orc_f16_to_double(unsigned short h)
{
    int e = (h >> 10) & 31, m = h & 1023;
    double sgn = (h & 0x8000) ? -1.0 : 1.0;
    if (e == 31) return m ? NAN : sgn * INFINITY;
    if (e == 0) return sgn * ldexp((double)m, -24);
    return sgn * ldexp(1.0 + m / 1024.0, e - 15);
}

static double orc_round_to(int code, double v)
{
    switch (code) {
    case 2: return orc_e4m3_to_double(orc_e4m3_from_double(v));
    case 3: return orc_f16_to_double(orc_f16_from_double(v));
    case 4: return (double)(float)v;
    default: return v;
    }
}

static int orc_fits(int code, double v, double tau)
{
    static const double maxv[5] = {0, 0, 448.0, 65504.0, 3.4028234663852886e38};
    static const double minn[5] = {0, 0, 0.015625, 6.103515625e-05, 1.1754943508222875e-38};
    if (!isfinite(v)) return 0;
    double a = fabs(v);
    if (a == 0.0) return 1;
    if (a > maxv[code] || a < minn[code]) return 0;
    double r = orc_round_to(code, v);
    return fabs(r - v) <= tau * a;
}

void orc_tile_classes(const orc_tiles *t, double tau, int allow_mask, unsigned char *codes)
{
    static const int order[3] = {2, 3, 4}; /* narrowest first */
    for (int tid = 0; tid < t->tilenum; tid++) {
        int off = t->csr_offset[tid], nz = t->blknnz[tid + 1] - t->blknnz[tid];
        unsigned char code = 0;
        for (int o = 0; o < 3 && !code; o++) {
            int c = order[o];
            if (!(allow_mask & (1 << c))) continue;
            int ok = 1;
            for (int k = 0; k < nz && ok; k++) ok = orc_fits(c, t->Blockcsr_Val[off + k], tau);
            if (ok) code = (unsigned char)c;
        }
        codes[tid] = code;
    }
}

void orc_tile_spmv_mixed(const orc_tiles *t, const unsigned char *codes, const double *x,
                         double *y)
{
    /* per row: FP64 tiles with fma in FP64, then the FP32, FP16 and FP8 tiles
     * (that class order, reference tile order inside a class) with fmaf in FP32
     * over (float)x; y = acc64 + (double)acc32 */
    static const int low_order[3] = {4, 3, 2};
    int B = t->B;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rowlen = bi == t->tilem - 1 ? t->rowA - (t->tilem - 1) * B : B;
        for (int ri = 0; ri < rowlen; ri++) {
            double acc = 0.0;
            float acc32 = 0.0f;
            for (int pass = 0; pass < 4; pass++) {
                int want = pass == 0 ? 0 : low_order[pass - 1];
                for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
                    if (codes[tid] != want) continue;
                    int off = t->csr_offset[tid], po = t->csrptr_offset[tid];
                    int nz = t->blknnz[tid + 1] - t->blknnz[tid];
                    int s = t->Blockcsr_Ptr[po + ri];
                    int e = ri == rowlen - 1 ? nz : t->Blockcsr_Ptr[po + ri + 1];
                    int xo = t->tile_columnidx[tid] * B;
                    for (int k = s; k < e; k++) {
                        double v = t->Blockcsr_Val[off + k];
                        double xv = x[xo + t->Tile_csr_Col[off + k]];
                        if (want == 0) acc = fma(v, xv, acc);
                        else acc32 = fmaf((float)orc_round_to(want, v), (float)xv, acc32);
                    }
                }
            }
            y[bi * B + ri] = acc + (double)acc32;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Solvers.  Vector kernels are plain loops (OpenMP when nthreads > 1). */
typedef struct {
    int n;
    const int *rowptr, *colidx;
    const double *val;
    const orc_tiles *tiles;
    int nthreads;
} orc_op;

static void op_apply(const orc_op *A, const double *x, double *y)
{
    if (A->tiles) {
        const orc_tiles *t = A->tiles;
        int B = t->B;
#pragma omp parallel for num_threads(A->nthreads) schedule(static) if (A->nthreads > 1)
        for (int bi = 0; bi < t->tilem; bi++) {
            int rowlen = bi == t->tilem - 1 ? t->rowA - (t->tilem - 1) * B : B;
            for (int ri = 0; ri < rowlen; ri++) {
                double acc = 0.0;
                for (int tid = t->tile_ptr[bi]; tid < t->tile_ptr[bi + 1]; tid++) {
                    int off = t->csr_offset[tid], po = t->csrptr_offset[tid];
                    int nz = t->blknnz[tid + 1] - t->blknnz[tid];
                    int s = t->Blockcsr_Ptr[po + ri];
                    int e = ri == rowlen - 1 ? nz : t->Blockcsr_Ptr[po + ri + 1];
                    int xo = t->tile_columnidx[tid] * B;
                    for (int k = s; k < e; k++) acc = fma(t->Blockcsr_Val[off + k], x[xo + t->Tile_csr_Col[off + k]], acc);
                }
                y[bi * B + ri] = acc;
            }
        }
        for (int i = t->rowA; i < A->n; i++) y[i] = 0.0;
    } else {
#pragma omp parallel for num_threads(A->nthreads) schedule(static) if (A->nthreads > 1)
        for (int i = 0; i < A->n; i++) {
            double s = 0.0;
            for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) s = fma(A->val[j], x[A->colidx[j]], s);
            y[i] = s;
        }
    }
}

static double vdot(int n, const double *a, const double *b, int nt)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) num_threads(nt) schedule(static) if (nt > 1)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* residual history (orc_report.hist): entry k = updated ||r_k|| */
static void hist_put(orc_report *rep, int k, double v)
{
    if (rep->hist && k < rep->hist_cap) {
        rep->hist[k] = v;
        if (k + 1 > rep->hist_len) rep->hist_len = k + 1;
    }
}

static double true_relres(const orc_op *A, const double *b, const double *x, double *tmp, double bnorm)
{
    op_apply(A, x, tmp);
    double s = 0.0;
    for (int i = 0; i < A->n; i++) { double d = b[i] - tmp[i]; s += d * d; }
    return bnorm > 0 ? sqrt(s) / bnorm : sqrt(s);
}

int orc_cg(int n, const int *rowptr, const int *colidx, const double *val,
           const orc_tiles *tiles, const double *b, double *x, double tol, int maxit,
           int nthreads, orc_report *rep)
{
    orc_op A = {n, rowptr, colidx, val, tiles, nthreads < 1 ? 1 : nthreads};
    int nt = A.nthreads;
    double *r = (double *)malloc(sizeof(double) * n), *d = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n), *tmp = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; d[i] = b[i]; }
    double bnorm = sqrt(vdot(n, b, b, nt));
    double snew = vdot(n, r, r, nt);
    int it = 0, conv = 0;
    double tr = 1.0;
    rep->hist_len = 0;
    hist_put(rep, 0, sqrt(snew));
    if (bnorm == 0.0) { conv = 1; tr = 0.0; }
    while (!conv && it < maxit) {
        op_apply(&A, d, q);                      /* q = A d */
        double dq = vdot(n, d, q, nt);
        if (dq == 0.0) break;
        double alpha = snew / dq;                 /* alpha = snew / (d.q) */
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
        for (int i = 0; i < n; i++) { x[i] += alpha * d[i]; r[i] -= alpha * q[i]; }
        double sold = snew;
        snew = vdot(n, r, r, nt);
        double beta = snew / sold;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
        for (int i = 0; i < n; i++) d[i] = r[i] + beta * d[i];
        it++;
        hist_put(rep, it, sqrt(snew));
        if (sqrt(snew) <= tol * bnorm) {
            tr = true_relres(&A, b, x, tmp, bnorm);
            if (tr <= tol) conv = 1;
        }
    }
    if (!conv) tr = true_relres(&A, b, x, tmp, bnorm);
    rep->iterations = it;
    rep->relres_true = tr;
    rep->relres_updated = bnorm > 0 ? sqrt(snew) / bnorm : 0.0;
    rep->converged = conv;
    free(r); free(d); free(q); free(tmp);
    return conv ? 0 : 1;
}

int orc_bicgstab(int n, const int *rowptr, const int *colidx, const double *val,
                 const orc_tiles *tiles, const double *b, double *x, double tol, int maxit,
                 int nthreads, orc_report *rep)
{
    orc_op A = {n, rowptr, colidx, val, tiles, nthreads < 1 ? 1 : nthreads};
    int nt = A.nthreads;
    size_t sz = sizeof(double) * (size_t)n;
    double *r = malloc(sz), *rh = malloc(sz), *p = malloc(sz), *v = malloc(sz);
    double *s = malloc(sz), *t = malloc(sz), *tmp = malloc(sz);
    for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; rh[i] = b[i]; p[i] = b[i]; }
    double bnorm = sqrt(vdot(n, b, b, nt));
    double rho = vdot(n, r, rh, nt), rr = rho;
    int it = 0, conv = 0;
    double tr = 1.0;
    rep->hist_len = 0;
    hist_put(rep, 0, sqrt(rr));
    if (bnorm == 0.0) { conv = 1; tr = 0.0; }
    while (!conv && it < maxit) {
        op_apply(&A, p, v);                       /* v = A p */
        double rv = vdot(n, rh, v, nt);
        if (rv == 0.0) break;
        double alpha = rho / rv;                   /* alpha = rho / (rh.v) */
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        op_apply(&A, s, t);                       /* t = A s */
        double ts = vdot(n, t, s, nt), tt = vdot(n, t, t, nt);
        double omega = tt != 0.0 ? ts / tt : 0.0;  /* omega = (t.s)/(t.t) */
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i] + omega * s[i]; r[i] = s[i] - omega * t[i]; }
        double rho_new = vdot(n, r, rh, nt);
        rr = vdot(n, r, r, nt);
        it++;
        hist_put(rep, it, sqrt(rr));
        if (sqrt(rr) <= tol * bnorm) {
            tr = true_relres(&A, b, x, tmp, bnorm);
            if (tr <= tol) { conv = 1; break; }
        }
        if (omega == 0.0 || rho == 0.0) break;
        double beta = (rho_new / rho) * (alpha / omega);
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        rho = rho_new;
    }
    if (!conv) tr = true_relres(&A, b, x, tmp, bnorm);
    rep->iterations = it;
    rep->relres_true = tr;
    rep->relres_updated = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
    rep->converged = conv;
    free(r); free(rh); free(p); free(v); free(s); free(t); free(tmp);
    return conv ? 0 : 1;
}

int orc_petsc_cg(int n, const int *rowptr, const int *colidx, const double *val,
                 const double *b, double *x, double rtol, int maxit, int nthreads,
                 orc_report *rep)
{
    orc_op A = {n, rowptr, colidx, val, NULL, nthreads < 1 ? 1 : nthreads};
    int nt = A.nthreads;
    size_t sz = sizeof(double) * (size_t)n;
    double *R = malloc(sz), *P = malloc(sz), *W = malloc(sz), *tmp = malloc(sz);
    for (int i = 0; i < n; i++) { x[i] = 0.0; R[i] = b[i]; }
    double dp = sqrt(vdot(n, R, R, nt)), rnorm0 = dp;
    double beta = vdot(n, R, R, nt), betaold = 1.0, dpi = 0.0, dpiold;
    int its = 0, conv = dp == 0.0, i = 0;
    rep->hist_len = 0;
    hist_put(rep, 0, dp);
    while (!conv && i < maxit) {
        its = i + 1;
        if (beta == 0.0) { conv = 1; break; }
        if (i > 0 && beta * betaold < 0.0) break;
        if (!i) memcpy(P, R, sz);
        else {
            double bb = beta / betaold;
            for (int k = 0; k < n; k++) P[k] = R[k] + bb * P[k];
        }
        dpiold = dpi;
        op_apply(&A, P, W);
        dpi = vdot(n, P, W, nt);
        betaold = beta;
        if (dpi == 0.0 || (i > 0 && dpi * dpiold < 0.0)) break;
        double a = beta / dpi;
        for (int k = 0; k < n; k++) { x[k] += a * P[k]; R[k] -= a * W[k]; }
        beta = vdot(n, R, R, nt); /* beta <- z'r with z = r (PCNONE), cg.c:363 */
        dp = sqrt(beta);          /* dp <- ||r||, cg.c:316 */
        hist_put(rep, i + 1, dp);
        if (dp < rtol * rnorm0) { conv = 1; break; }
        i++;
    }
    rep->iterations = its;
    rep->relres_updated = rnorm0 > 0 ? dp / rnorm0 : 0.0;
    rep->relres_true = true_relres(&A, b, x, tmp, rnorm0);
    rep->converged = conv;
    free(R); free(P); free(W); free(tmp);
    return conv ? 0 : 1;
}

int orc_petsc_bcgs(int n, const int *rowptr, const int *colidx, const double *val,
                   const double *b, double *x, double rtol, int maxit, int nthreads,
                   orc_report *rep)
{
    orc_op A = {n, rowptr, colidx, val, NULL, nthreads < 1 ? 1 : nthreads};
    int nt = A.nthreads;
    size_t sz = sizeof(double) * (size_t)n;
    double *R = malloc(sz), *RP = malloc(sz), *V = calloc((size_t)n, sizeof(double));
    double *T = malloc(sz), *S = malloc(sz), *P = calloc((size_t)n, sizeof(double)), *tmp = malloc(sz);
    for (int i = 0; i < n; i++) { x[i] = 0.0; R[i] = b[i]; RP[i] = b[i]; }
    double dp = sqrt(vdot(n, R, R, nt)), rnorm0 = dp;
    double rhoold = 1.0, alpha = 1.0, omegaold = 1.0;
    int its = 0, conv = dp == 0.0, i = 0;
    rep->hist_len = 0;
    hist_put(rep, 0, dp);
    while (!conv && i < maxit) {
        double rho = vdot(n, R, RP, nt);
        double beta = (rho / rhoold) * (alpha / omegaold);
        for (int k = 0; k < n; k++) P[k] = R[k] - omegaold * beta * V[k] + beta * P[k];
        op_apply(&A, P, V);
        double d1 = vdot(n, V, RP, nt);
        if (d1 == 0.0) break;
        alpha = rho / d1;
        for (int k = 0; k < n; k++) S[k] = R[k] - alpha * V[k];
        op_apply(&A, S, T);
        d1 = vdot(n, S, T, nt);
        double d2 = vdot(n, T, T, nt);
        if (d2 == 0.0) {
            /* t = 0: converged only when s = 0 too, else a singular operator (bcgs.c:96-110) */
            if (vdot(n, S, S, nt) != 0.0) break;
            for (int k = 0; k < n; k++) x[k] += alpha * P[k];
            its++; dp = 0.0; conv = 1;
            hist_put(rep, its, dp);
            break;
        }
        double omega = d1 / d2;
        for (int k = 0; k < n; k++) { x[k] += alpha * P[k] + omega * S[k]; R[k] = S[k] - omega * T[k]; }
        dp = sqrt(vdot(n, R, R, nt));
        rhoold = rho; omegaold = omega;
        its++;
        hist_put(rep, its, dp);
        if (dp < rtol * rnorm0) { conv = 1; break; }
        if (rho == 0.0) break;
        i++;
    }
    rep->iterations = its;
    rep->relres_updated = rnorm0 > 0 ? dp / rnorm0 : 0.0;
    rep->relres_true = true_relres(&A, b, x, tmp, rnorm0);
    rep->converged = conv;
    free(R); free(RP); free(V); free(T); free(S); free(P); free(tmp);
    return conv ? 0 : 1;
}
