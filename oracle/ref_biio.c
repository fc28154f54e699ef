/*
 * ref_biio.c -- exposes the reference's binary CSR reader and writer
 * (biio2.0/src/biio.h with its own mmio_highlevel.h, compiled from the files
 * where they lie under the reference tree, never copied) so the tests can
 * check the library's drop-ins against them.  TEST INFRASTRUCTURE ONLY; built
 * into oracle/_ref/ by oracle/Makefile.
 */
#include "biio.h"

int ref_binary_read_csr(int *row, int *col, int64_t *nnz, int64_t **row_ptr, int **col_idx, double **val,
                        char *filename)
{
    return binary_read_csr(row, col, nnz, row_ptr, col_idx, val, filename);
}

void ref_binary_write_csr(int row, int col, int64_t nnz, int64_t *row_ptr, int *col_idx, double *val,
                          char *filename)
{
    binary_write_csr(row, col, nnz, row_ptr, col_idx, val, filename);
}

void ref_read_Dmatrix_32(int *row, int *col, int *nnz, int **row_ptr, int **col_idx, double **val, int *isSym,
                         char *filename)
{
    read_Dmatrix_32(row, col, nnz, row_ptr, col_idx, val, isSym, filename);
}

void ref_read_Dmatrix_convert(int *row, int *col, int64_t *nnz, int64_t **row_ptr, int **col_idx, double **val,
                              int *isSym, char *filename)
{
    read_Dmatrix_convert(row, col, nnz, row_ptr, col_idx, val, isSym, filename);
}

void ref_biio_free(void *p) { free(p); }
