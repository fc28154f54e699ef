/*
 * ref_shim.c -- exposes the reference's own host C (compiled from the files
 * where they lie under the reference tree, never copied) through a few
 * extern functions, so the tests can check the oracle against the reference
 * on the shipped and synthetic matrices.  TEST INFRASTRUCTURE ONLY; built into
 * oracle/_ref/ by oracle/Makefile, with BLOCK_SIZE set on the command line.
 *
 * Headers used: PC/mmio_highlevel.h (+ PC/mmio.h), PC/csr2block.h (+ common.h,
 * encode.h, format.h, utils.h), PC/blockspmv_cpu.h.
 */
#include "mmio_highlevel.h"
#include "csr2block.h"
#include "blockspmv_cpu.h"

/* Same field order as orc_tiles in oracle.h. */
typedef struct {
    int B, rowA, colA, tilem, tilen, tilenum, csrsize, csrptrlen;
    int *tile_ptr, *tile_columnidx, *tile_nnz;
    char *Format;
    int *blknnz;
    unsigned char *blknnznnz;
    int *csr_offset, *csrptr_offset;
    double *Blockcsr_Val;
    float *Blockcsr_Val_Low;
    unsigned char *Blockcsr_Ptr, *Tile_csr_Col, *csr_compressedIdx;
} ref_tiles;

int ref_block_size(void) { return BLOCK_SIZE; }

int ref_mmio_info(int *m, int *n, int *nnz, int *isSym, char *filename)
{
    return mmio_info(m, n, nnz, isSym, filename);
}

int ref_mmio_data(int *rowptr, int *colidx, double *val, char *filename)
{
    return mmio_data(rowptr, colidx, val, filename);
}

void *ref_tile_create(int rowA, int colA, int nnzA, int *rowptr, int *colidx, double *val,
                      float *val_low, ref_tiles *out)
{
    Tile_matrix *m = (Tile_matrix *)calloc(1, sizeof(Tile_matrix));
    Tile_create(m, rowA, colA, nnzA, rowptr, colidx, val, val_low);
    out->B = BLOCK_SIZE;
    out->rowA = rowA;
    out->colA = colA;
    out->tilem = m->tilem;
    out->tilen = m->tilen;
    out->tilenum = m->tilenum;
    out->csrsize = m->csrsize;
    out->csrptrlen = m->csrptrlen;
    out->tile_ptr = m->tile_ptr;
    out->tile_columnidx = m->tile_columnidx;
    out->tile_nnz = m->tile_nnz;
    out->Format = m->Format;
    out->blknnz = m->blknnz;
    out->blknnznnz = m->blknnznnz;
    out->csr_offset = m->csr_offset;
    out->csrptr_offset = m->csrptr_offset;
    out->Blockcsr_Val = m->Blockcsr_Val;
    out->Blockcsr_Val_Low = m->Blockcsr_Val_Low;
    out->Blockcsr_Ptr = m->Blockcsr_Ptr;
    out->Tile_csr_Col = m->Tile_csr_Col;
    out->csr_compressedIdx = m->csr_compressedIdx;
    return m;
}

void ref_tile_free(void *handle)
{
    Tile_matrix *m = (Tile_matrix *)handle;
    free(m->tile_ptr); free(m->tile_columnidx); free(m->tile_nnz); free(m->Format);
    free(m->blknnz); free(m->blknnznnz); free(m->tilewidth); free(m->csr_offset);
    free(m->csrptr_offset); free(m->Blockcsr_Val); free(m->Blockcsr_Val_Low);
    free(m->Blockcsr_Ptr); free(m->Tile_csr_Col); free(m->csr_compressedIdx);
    free(m);
}

void ref_encode(unsigned char *str, unsigned char *res, int len, int offset)
{
    encode(str, res, len, offset);
}

void ref_decode(unsigned char *str, unsigned char *res, int len) { decode(str, res, len); }

void ref_exclusive_scan(int *v, int len) { exclusive_scan(v, len); }

void ref_exclusive_scan_char(unsigned char *v, int len) { exclusive_scan_char(v, len); }

/* blockspmv_cpu's side outputs (ptroffset1/2 and the row-block split); its y
 * is not used because it decodes the zeroed csr_compressedIdx. */
int ref_blockspmv_offsets(void *handle, int rowA, int colA, int nnzA, int *rowptr, int *colidx,
                          double *val, int *ptroffset1, int *ptroffset2, unsigned int *blkrow,
                          int *colstart, int *colstop, int cap)
{
    Tile_matrix *m = (Tile_matrix *)handle;
    int rbb = 0;
    unsigned int *br;
    int *cs, *ce;
    double *x = (double *)calloc((size_t)colA + 64, sizeof(double));
    double *y = (double *)calloc((size_t)m->tilem * BLOCK_SIZE + 64, sizeof(double));
    blockspmv_cpu(m, ptroffset1, ptroffset2, &rbb, &br, &cs, &ce, rowA, colA, nnzA, rowptr,
                  colidx, val, x, y, NULL);
    for (int i = 0; i < rbb && i < cap; i++) {
        blkrow[i] = br[i];
        colstart[i] = cs[i];
        colstop[i] = ce[i];
    }
    free(br); free(cs); free(ce); free(x); free(y);
    return rbb;
}
