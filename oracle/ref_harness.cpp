// ref_harness.cpp -- exposes the reference's own xtax move primitives through a
// C ABI so tests/golden/gen_golden.py can record golden vectors from them.
//
// Built by `make -C oracle ref` into oracle/_ref/ (git-ignored) directly from
// the reference sources under REF (default /root/reference); nothing of the
// reference is copied here. The reference's main() is renamed so this file can
// be a shared library.
#define main xtax_reference_main
#include "main.cpp"
#undef main

#include <cstring>

namespace {
Matrix to_matrix(const int64_t* p, int n) {
    Matrix M((size_t)n);
    std::memcpy(M.data.data(), p, sizeof(int64_t) * (size_t)n * (size_t)n);
    return M;
}
void from_matrix(const Matrix& M, int64_t* p) {
    std::memcpy(p, M.data.data(), sizeof(int64_t) * M.data.size());
}
}  // namespace

extern "C" {

int64_t ref_rounded_div(int64_t a, int64_t b) { return rounded_div(a, b); }

int64_t ref_score(const int64_t* A, int n) { return to_matrix(A, n).score(); }

int64_t ref_offdiag_nonzero(const int64_t* A, int n) {
    return to_matrix(A, n).count_offdiag_nonzero();
}

int ref_try_add(const int64_t* A, const int64_t* Xt, int n, int i, int j, int64_t s,
                int64_t* new_a_row, int64_t* new_xt_row, int64_t* delta, int64_t* dnz) {
    const Matrix MA = to_matrix(A, n);
    const Matrix MX = to_matrix(Xt, n);
    std::vector<int64_t> row((size_t)n, 0), xrow((size_t)n, 0);
    const bool ok = try_add(MA, MX, i, j, s, row, new_xt_row ? &xrow : nullptr, *delta, *dnz);
    std::memcpy(new_a_row, row.data(), sizeof(int64_t) * (size_t)n);
    if (new_xt_row) std::memcpy(new_xt_row, xrow.data(), sizeof(int64_t) * (size_t)n);
    return ok ? 1 : 0;
}

int ref_add_transform_row(const int64_t* Xt, int n, int i, int j, int64_t s, int64_t* out) {
    const Matrix MX = to_matrix(Xt, n);
    std::vector<int64_t> xrow((size_t)n, 0);
    const bool ok = add_transform_row(MX, i, j, s, xrow);
    std::memcpy(out, xrow.data(), sizeof(int64_t) * (size_t)n);
    return ok ? 1 : 0;
}

void ref_commit_add(int64_t* A, int64_t* Xt, int n, int j, const int64_t* new_a_row,
                    const int64_t* new_xt_row, int64_t* row_mass) {
    Matrix MA = to_matrix(A, n);
    Matrix MX = to_matrix(Xt, n);
    std::vector<int64_t> row(new_a_row, new_a_row + n), xrow(new_xt_row, new_xt_row + n);
    std::vector<int64_t> mass(row_mass, row_mass + n);
    commit_add(MA, MX, j, row, xrow, mass);
    from_matrix(MA, A);
    from_matrix(MX, Xt);
    std::memcpy(row_mass, mass.data(), sizeof(int64_t) * (size_t)n);
}

void ref_compute_row_mass(const int64_t* A, int n, int64_t* row_mass) {
    std::vector<int64_t> mass((size_t)n, 0);
    compute_row_mass(to_matrix(A, n), mass);
    std::memcpy(row_mass, mass.data(), sizeof(int64_t) * (size_t)n);
}

int64_t ref_score_delta_l1(const int64_t* A, int n, int i, int j, int64_t s) {
    return score_delta_l1(to_matrix(A, n), i, j, s);
}

int64_t ref_best_shear_l1(const int64_t* A, int n, int i, int j) {
    return best_shear_l1(to_matrix(A, n), i, j);
}

void ref_congruence_of(const int64_t* A, const int64_t* X, int n, int64_t* out) {
    from_matrix(congruence_of(to_matrix(A, n), to_matrix(X, n)), out);
}

void ref_gram_of(const int64_t* L, int m, int d, int64_t* out) {
    Lattice lat;
    lat.m = m;
    lat.d = d;
    lat.data.assign(L, L + (size_t)m * d);
    from_matrix(gram_of(lat), out);
}

void ref_xt_times(const int64_t* X, const int64_t* L, int m, int d, int64_t* out) {
    Lattice lat;
    lat.m = m;
    lat.d = d;
    lat.data.assign(L, L + (size_t)m * d);
    const Lattice R = xt_times(to_matrix(X, m), lat);
    std::memcpy(out, R.data.data(), sizeof(int64_t) * R.data.size());
}

// ---- publish-time reordering and deflation (main.cpp:244-350, :435-455;
// mat_io.hpp:597-644) ----

void ref_rcm_order(const int64_t* A, int n, int* perm) {
    const std::vector<int> p = rcm_order(to_matrix(A, n));
    std::memcpy(perm, p.data(), sizeof(int) * p.size());
}

void ref_centroid_order(const int64_t* A, int n, int* perm) {
    const std::vector<int> p = centroid_order(to_matrix(A, n));
    std::memcpy(perm, p.data(), sizeof(int) * p.size());
}

// L1Objective::reorder_for_publish(true) on (A, Xt) with the given flags.
void ref_reorder_publish(int64_t* A, int64_t* Xt, int n, int band, int centroid) {
    Matrix MA = to_matrix(A, n), MX = to_matrix(Xt, n);
    std::vector<int64_t> mass((size_t)n, 0);
    compute_row_mass(MA, mass);
    if (band) permute_state(MA, MX, rcm_order(MA), mass);
    if (centroid) permute_state(MA, MX, centroid_order(MA), mass);
    from_matrix(MA, A);
    from_matrix(MX, Xt);
}

int ref_is_unimodular(const int64_t* A, int n) { return is_unimodular(to_matrix(A, n)) ? 1 : 0; }

int64_t ref_det_mod_prime(const int64_t* A, int n, int64_t p) {
    return det_mod_prime(to_matrix(A, n), p);
}

int ref_deflate_index(int64_t* M, int64_t* Xt, int n, int i, const int* active, int n_active) {
    Matrix MM = to_matrix(M, n), MX = to_matrix(Xt, n);
    std::vector<int> act(active, active + n_active);
    std::vector<int64_t> row((size_t)n), xrow((size_t)n), mass((size_t)n, 0);
    const bool ok = deflate_index(MM, MX, i, act, row, xrow, mass);
    from_matrix(MM, M);
    from_matrix(MX, Xt);
    return ok ? 1 : 0;
}

}  // extern "C"
