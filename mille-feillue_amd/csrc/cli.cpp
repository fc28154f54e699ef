// Command-line drivers with the reference's interface:
//   mf_cg       <matrix.mtx> <block_nnz>   (PC/Mille-feuille_CG_AMD.cu:2102-2166)
//   mf_bicgstab <matrix.mtx> <block_nnz>   (PC/Mille-feuille_BiCGSTAB_AMD.cu:1630-1661)
// Each appends one row to the reference's CSV file in the working directory.
// Extra flags (after the two positional arguments):
//   --mixed [tau]  per-tile mixed precision with FP64 refinement; prints an
//                  extended result line (iterations, true residual, times).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mf.h"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <matrix.mtx> [block_nnz] [--mixed [tau]]\n", argv[0]);
        return 2;
    }
    char *filename = argv[1];
    int block_nnz = argc > 2 ? std::atoi(argv[2]) : 16;
    bool mixed = false;
    double tau = 0.0;
    for (int i = 3; i < argc; i++)
        if (!std::strcmp(argv[i], "--mixed")) {
            mixed = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') tau = std::atof(argv[++i]);
        }
    int m = 0, n = 0, nnzR = 0, isSym = 0;
    int rc = mmio_info(&m, &n, &nnzR, &isSym, filename);
    if (rc) {
        std::fprintf(stderr, "mmio_info: %d (%s)\n", rc, mf_last_error());
        return 1;
    }
    std::vector<int> RowPtr(m + 1), ColIdx(std::max(nnzR, 1));
    std::vector<double> Val(std::max(nnzR, 1));
    rc = mmio_data(RowPtr.data(), ColIdx.data(), Val.data(), filename);
    if (rc) {
        std::fprintf(stderr, "mmio_data: %d (%s)\n", rc, mf_last_error());
        return 1;
    }
#ifdef MF_CLI_CG
    const int method = MF_CG;
#else
    const int method = MF_BICGSTAB;
#endif
    if (!mixed) {
#ifdef MF_CLI_CG
        // b = A * 1 over the truncated rows (:2122-2157)
        std::vector<float> low(std::max(nnzR, 1));
        for (int i = 0; i < nnzR; i++) low[i] = (float)Val[i];
        int nt = (n / 16) * 16;
        std::vector<double> X(std::max(n, 1), 0.0), Y(std::max(m, 1), 0.0);
        for (int i = 0; i < nt; i++)
            for (int j = RowPtr[i]; j < RowPtr[i + 1]; j++) Y[i] += Val[j];
        int iter = 0;
        cg_solve_inc(RowPtr.data(), ColIdx.data(), Val.data(), low.data(), X.data(), Y.data(), nt, &iter, 10,
                     1e-5, filename, nnzR, n, block_nnz);
        std::printf("iterations=%d\n", iter);
        return 0;
#else
        std::vector<double> x(m, 0.0), rhs(m, 1.0);
        return bicgstab(RowPtr.data(), ColIdx.data(), Val.data(), rhs.data(), x.data(), nnzR, filename, m,
                        block_nnz);
#endif
    }
    // mixed precision through the library API, b = A * 1
    if (m != n) {
        std::fprintf(stderr, "square matrix required\n");
        return 1;
    }
    mf_matrix A = nullptr;
    rc = mf_matrix_create(m, n, nnzR, RowPtr.data(), ColIdx.data(), Val.data(), nullptr, 16, &A);
    if (rc) {
        std::fprintf(stderr, "mf_matrix_create: %s (%s)\n", mf_status_str(rc), mf_last_error());
        return 1;
    }
    std::vector<double> b(m, 0.0), x(m, 0.0);
    for (int i = 0; i < m; i++)
        for (int j = RowPtr[i]; j < RowPtr[i + 1]; j++) b[i] += Val[j];
    mf_opts o;
    mf_opts_default(&o);
    o.method = method;
    o.precision = MF_MIXED;
    o.tau = tau;
    o.rr_factor = 1e-2;
    mf_report rep;
    rc = mf_solve(A, b.data(), x.data(), &o, &rep);
    std::printf("%s,method=%s,precision=mixed,tau=%g,status=%d,iterations=%d,outer=%d,relres=%e,"
                "time_ms=%.3f,setup_ms=%.3f\n",
                filename, method == MF_CG ? "cg" : "bicgstab", tau, rc ? rc : rep.status, rep.iterations,
                rep.outer, rep.relres_true, rep.ms_solve, rep.ms_setup);
    mf_matrix_destroy(A);
    return rc || rep.status ? 1 : 0;
}
