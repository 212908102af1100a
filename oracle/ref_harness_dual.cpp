// ref_harness_dual.cpp -- exposes the reference's own xdual primal/dual
// primitives (src/xdual_main.cpp) and its dual inverse (invert_to,
// src/mat_io.hpp) through a C ABI so tests/golden/gen_golden.py can record
// golden vectors from them.
//
// Built by `make -C oracle ref` into oracle/_ref/ (git-ignored) directly from
// the reference sources under REF; nothing of the reference is copied here. The
// reference's main() is renamed so this file can be a shared library.
#define main xdual_reference_main
#include "xdual_main.cpp"
#undef main

#include <cstring>

namespace {
Matrix to_matrix(const int64_t* p, int n) {
    Matrix M((size_t)n);
    std::memcpy(M.data.data(), p, sizeof(int64_t) * (size_t)n * (size_t)n);
    return M;
}
Matrixd to_matrixd(const double* p, int n) {
    Matrixd M((size_t)n);
    std::memcpy(M.data.data(), p, sizeof(double) * (size_t)n * (size_t)n);
    return M;
}
}  // namespace

extern "C" {

// eval_add (primal, Add(i, j, s)) and eval_add_d (dual, Add(j, i, -s)), as
// DualObjective::evaluate combines them; the long double deltas as doubles.
int refd_eval(const int64_t* P, const double* Q, int n, int i, int j, int64_t s,
              int64_t* new_p_row, double* new_q_row, double* d_off_p, double* d_off_q,
              int64_t* dnz) {
    const Matrix MP = to_matrix(P, n);
    const Matrixd MQ = to_matrixd(Q, n);
    std::vector<int64_t> prow((size_t)n, 0);
    std::vector<double> qrow((size_t)n, 0.0);
    long double dp = 0.0L, dq = 0.0L;
    int64_t nz = 0;
    if (!eval_add(MP, i, j, s, prow, dp, nz)) return 0;
    eval_add_d(MQ, j, i, -s, qrow, dq);
    std::memcpy(new_p_row, prow.data(), sizeof(int64_t) * (size_t)n);
    std::memcpy(new_q_row, qrow.data(), sizeof(double) * (size_t)n);
    *d_off_p = (double)dp;
    *d_off_q = (double)dq;
    *dnz = nz;
    return 1;
}

int64_t refd_best_shear(const int64_t* P, const double* Q, int n, double c, int i, int j) {
    Matrix X((size_t)n);
    X.fill_identity();
    const DualObjective obj(to_matrix(P, n), to_matrixd(Q, n), X, (long double)c);
    return obj.best_shear(i, j);
}

int refd_invert(const int64_t* P, int n, double* Q) {
    Matrixd MQ;
    if (!invert_to(to_matrix(P, n), MQ)) return 0;
    std::memcpy(Q, MQ.data.data(), sizeof(double) * (size_t)n * (size_t)n);
    return 1;
}

double refd_frob2_off_p(const int64_t* P, int n) { return (double)to_matrix(P, n).offdiag_frob2(); }

double refd_frob2_off_q(const double* Q, int n) { return (double)to_matrixd(Q, n).offdiag_frob2(); }

}  // extern "C"
