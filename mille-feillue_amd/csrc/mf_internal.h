// Internal declarations shared by the host (.cpp) and device (.hip) sources of
// libmfamd.so.  Not installed; include/mf.h is the public C-ABI.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/mf.h"

namespace mf {

// Error plumbing: every C-ABI entry point returns an mf_status and leaves a
// message in a thread-local buffer.
int fail(int status, const std::string &msg);
void clear_error();

// Matrix Market parse result (file order, 0-based, symmetric entries mirrored
// by the CSR build).
struct MMData {
    int m = 0, n = 0, nz_report = 0;
    bool is_sym = false;
    int *ri = nullptr, *ci = nullptr;
    double *vv = nullptr;
};
int mm_parse(const char *path, MMData *out);
void mm_free(MMData *d);

// Host CSR build from a parse (PC/mmio_highlevel.h:507-575 semantics).
int mm_to_csr(const MMData &d, int *rowptr, int *colidx, double *val);
long long mm_nnz_expanded(const MMData &d);

// CSV rows of the reference drivers.
int csv_append_cg(const char *path, const char *matrix, int iterations, double time_spmv,
                  int nnzR, double l2_norm, double time_format, int index, int each_block_nnz,
                  double norm);
int csv_append_bicg(const char *path, const char *matrix, int itr, double time_spmv, int nnzR,
                    double l2_norm, double resid_ratio, int each_block_nnz);

double now_ms();

// Called before each OpenMP loop of the library (host.cpp, built with
// -fopenmp): the calling thread's workers sleep as soon as a loop ends
// instead of spinning for the runtime's default 200 ms.  Spinning workers,
// with another busy thread pool in the process (numpy's BLAS), were measured
// delaying the completion of the next GPU solves by ~2.6 ms each.
void omp_quiet();

}  // namespace mf
