// Batched persistent solver (many independent systems x several RHS).
#include <hip/hip_runtime.h>

#include "mf_matrix.h"

using namespace mf;

extern "C" int mf_solve_batched(int nsys, const int *n, const int *const *rowptr,
                                const int *const *colidx, const double *const *val, int nrhs,
                                const double *const *b, double *const *x, const mf_opts *o,
                                int *iters, double *relres, mf_report *rep)
{
    (void)nsys; (void)n; (void)rowptr; (void)colidx; (void)val; (void)nrhs; (void)b; (void)x;
    (void)o; (void)iters; (void)relres; (void)rep;
    return fail(MF_E_UNSUPPORTED, "batched solver not built yet");
}
