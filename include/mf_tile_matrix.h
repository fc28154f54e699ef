/*
 * mf_tile_matrix.h -- the reference's Tile_matrix record (PC/format.h:14-86),
 * declared with the same field order and types so that code written against
 * the reference can pass its struct to Tile_create/Tile_destroy of this
 * library unchanged.  Only the CSR-tile fields are filled; the ELL/HYB/COO/
 * dense fields are unused in the reference too and stay NULL/0.
 */
#ifndef MF_TILE_MATRIX_H
#define MF_TILE_MATRIX_H

#define MF_TILE_BLOCK_SIZE 16

typedef struct {
    double csrval_high;
    float csrval_low;
} mixval;

typedef struct {
    int tilem;
    int tilen;
    int tilenum;
    int *tile_ptr;
    int *tile_columnidx;
    int *tile_nnz;
    char *Format;
    int *blknnz;
    unsigned char *blknnznnz;
    int *dnsrowptr;
    int *dnscolptr;
    char *tilewidth;
    int *csr_offset;
    int *csrptr_offset;
    int *coo_offset;
    int *ell_offset;
    int *hyb_offset;
    int *hyb_coocount;
    int *dns_offset;
    int *dnsrow_offset;
    int *dnscol_offset;
    int *new_coocount;
    double *Blockcsr_Val;
    float *Blockcsr_Val_Low;
    unsigned char *Tile_csr_Col;
    mixval *Blockcsr_Val_mix;
    unsigned char *Blockcsr_Ptr;
    unsigned char *csr_compressedIdx;
    int csrsize;
    int csrptrlen;
    double *Blockcoo_Val;
    float *Blockcoo_Val_Low;
    unsigned char *coo_compressed_Idx;
    int coosize;
    double *Blockell_Val;
    float *Blockell_Val_Low;
    unsigned char *ell_compressedIdx;
    int ellsize;
    double *Blockhyb_Val;
    float *Blockhyb_Val_Low;
    unsigned char *hybIdx;
    int hybsize;
    int hybellsize;
    int hybcoosize;
    double *Blockdense_Val;
    float *Blockdense_Val_Low;
    int dnssize;
    double *Blockdenserow_Val;
    float *Blockdenserow_Val_Low;
    char *denserowid;
    int dnsrowsize;
    double *Blockdensecol_Val;
    float *Blockdensecol_Val_Low;
    char *densecolid;
    int dnscolsize;
    int coototal;
    double *deferredcoo_val;
    float *deferredcoo_val_Low;
    int *deferredcoo_colidx;
    int *deferredcoo_ptr;
    int rowblkblock;
    unsigned int *flag_bal_tile_rowidx;
    int *tile_bal_rowidx_colstart;
    int *tile_bal_rowidx_colstop;
    int *flag_tilerow_start;
    int *flag_tilerow_stop;
    int *tile_bal_rowidx_colstart_v2;
    int *tile_bal_rowidx_colstop_v2;
    int *map;
} Tile_matrix;

#endif
