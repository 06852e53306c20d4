// Host-only checks built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_sanitizers.py): the product's host logic (the eigensolver's
// tridiagonal / arrowhead eigensolve, the omega recurrence of the partial
// reorthogonalisation, the group closure and inverses, the colex ranks of
// the spread of x) and the oracle's C restatement on
// small inputs, so that an out-of-bounds access or undefined behaviour on
// these paths aborts the run.  Returns 0 when every check holds.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <cmath>
#include <random>
#include <vector>

#include "../../exactdiagonalization.jl_amd/csrc/host_logic.hpp"

extern "C" {
int64_t orc_symmetry_reduce_x(int64_t n_basis, const uint64_t* basis, int n_sites, const int32_t* bitoffsets,
                              const int32_t* bitwidths, int n_group, const int32_t* maps, const uint64_t* flips,
                              const double* chi_re, const double* chi_im, double tol, uint64_t* reduced_out,
                              int64_t* mapping_index, double* amp_re, double* amp_im);
void orc_apply(int64_t n_parent, const uint64_t* parent, int reduced, int64_t n_red, const uint64_t* red,
               const int64_t* map_idx, const double* map_re, const double* map_im, int64_t n_terms,
               const uint64_t* mask, const uint64_t* row, const uint64_t* col, const double* a_re, const double* a_im,
               int transpose, const double* x, double* y);
int64_t orc_sparse(int64_t n_parent, const uint64_t* parent, int reduced, int64_t n_red, const uint64_t* red,
                   const int64_t* map_idx, const double* map_re, const double* map_im, int64_t n_terms,
                   const uint64_t* mask, const uint64_t* row, const uint64_t* col, const double* a_re,
                   const double* a_im, double tol, int64_t* colptr, int64_t* rowval, double* nz_re, double* nz_im,
                   int64_t cap);
int64_t orc_sparse_parallel(int64_t n_parent, const uint64_t* parent, int reduced, int64_t n_red, const uint64_t* red,
                            const int64_t* map_idx, const double* map_re, const double* map_im, int64_t n_terms,
                            const uint64_t* mask, const uint64_t* row, const uint64_t* col, const double* a_re,
                            const double* a_im, double tol, int n_threads, int64_t** colptr_out, int64_t** rowval_out,
                            double** nz_re_out, double** nz_im_out);
void orc_free(void* p);
}

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                  \
    }                                                              \
  } while (0)

// eigen-decomposition through the product's host routines: A = Z diag(w) Z^T
static void check_eig(int n, const std::vector<double>& A, bool tridiagonal) {
  std::vector<double> d(n), e(n, 0.0), Z;
  if (tridiagonal) {
    for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; ++i) e[i] = A[(size_t)i * n + i + 1];
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
  } else {
    std::vector<double> B = A;
    edhip::householder_tridiag(n, B, Z, d, e);
  }
  CHECK(edhip::tridiag_ql(n, d, e, Z));
  double scale = 1.0, err = 0.0, orth = 0.0;
  for (double v : A) scale = std::max(scale, std::fabs(v));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0, o = 0.0;
      for (int k = 0; k < n; ++k) {
        s += Z[(size_t)i * n + k] * d[k] * Z[(size_t)j * n + k];
        o += Z[(size_t)k * n + i] * Z[(size_t)k * n + j];
      }
      err = std::max(err, std::fabs(s - A[(size_t)i * n + j]));
      orth = std::max(orth, std::fabs(o - (i == j ? 1.0 : 0.0)));
    }
  CHECK(err < 1e-10 * scale * n);
  CHECK(orth < 1e-10 * n);
  for (int i = 0; i + 1 < n; ++i) CHECK(d[i] <= d[i + 1]);
}

static void test_linear_algebra() {
  std::mt19937_64 rng(7);
  std::normal_distribution<double> nd;
  for (int n : {1, 2, 3, 7, 24, 61}) {
    std::vector<double> T((size_t)n * n, 0.0), A((size_t)n * n, 0.0), W((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
      T[(size_t)i * n + i] = nd(rng);
      if (i + 1 < n) T[(size_t)i * n + i + 1] = T[(size_t)(i + 1) * n + i] = (i % 5 == 3) ? 0.0 : nd(rng);
      for (int j = 0; j <= i; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = nd(rng);
    }
    // thick-restart arrowhead: a diagonal block of kept Ritz values coupled to
    // one vector, then the tridiagonal continuation
    const int k = n / 3;
    for (int i = 0; i < n; ++i) {
      W[(size_t)i * n + i] = i < k ? -1.0 + 0.1 * i : nd(rng);
      if (i < k && k < n) W[(size_t)i * n + k] = W[(size_t)k * n + i] = nd(rng);
      if (i >= k && i + 1 < n) W[(size_t)i * n + i + 1] = W[(size_t)(i + 1) * n + i] = nd(rng);
    }
    check_eig(n, T, true);
    check_eig(n, A, false);
    check_eig(n, W, false);
  }
  // degenerate spectrum (identity-like) and a zero matrix
  for (int n : {4, 9}) {
    std::vector<double> I((size_t)n * n, 0.0), Zr((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 2.0;
    check_eig(n, I, true);
    check_eig(n, Zr, false);
  }
}

static void test_group_closure() {
  // translations of a ring of L sites, with and without the global flip
  for (int L : {1, 2, 5, 8}) {
    for (int flip = 0; flip < 2; ++flip) {
      const int G = L * (flip ? 2 : 1);
      std::vector<int32_t> perm((size_t)G * L);
      std::vector<uint64_t> xr(G, 0ull);
      for (int f = 0; f <= flip; ++f)
        for (int a = 0; a < L; ++a) {
          const int g = f * L + a;
          for (int i = 0; i < L; ++i) perm[(size_t)g * L + i] = (i + a) % L;
          xr[g] = f ? ((1ull << L) - 1ull) : 0ull;
        }
      std::vector<int32_t> inv;
      CHECK(edhip::group_inverses(G, L, perm, xr, inv).empty());
      for (int g = 0; g < G; ++g) {
        CHECK(inv[g] >= 0 && inv[g] < G);
        const int a = g % L, f = g / L;
        CHECK(inv[g] == f * L + (L - a) % L);
      }
    }
  }
  // not closed: T(1) of a 5-ring without T(2..4); repeated element
  std::vector<int32_t> perm = {0, 1, 2, 3, 4, 1, 2, 3, 4, 0};
  std::vector<uint64_t> xr = {0, 0};
  std::vector<int32_t> inv;
  CHECK(!edhip::group_inverses(2, 5, perm, xr, inv).empty());
  std::vector<int32_t> rep = {0, 1, 0, 1};
  CHECK(!edhip::group_inverses(2, 2, rep, std::vector<uint64_t>{0, 0}, inv).empty());
}

static void test_colex() {
  // the colex rank of a 12-bit word is a bijection of the words of each
  // popcount j onto [0, C(12, j)), below the stride 924
  const std::vector<uint16_t> t = edhip::colex12_tables();
  CHECK(t.size() == 64 + 64 * 7);
  for (int j = 0; j <= 12; ++j) {
    std::vector<int> seen(edhip::kColexStride, 0);
    int n = 0;
    for (uint32_t l = 0; l < 4096; ++l) {
      if (__builtin_popcount(l) != j) continue;
      ++n;
      const uint32_t r = t[l & 63u] + t[64u + (l >> 6) * 7u + (uint32_t)__builtin_popcount(l & 63u)];
      CHECK(r < (uint32_t)edhip::kColexStride);
      if (r < (uint32_t)edhip::kColexStride) ++seen[r];
    }
    for (int r = 0; r < edhip::kColexStride; ++r) CHECK(seen[r] == (r < n ? 1 : 0));
  }
  // the device's one-read table holds the same ranks
  const std::vector<uint16_t> f = edhip::colex12_full();
  CHECK(f.size() == 4096);
  for (uint32_t l = 0; l < 4096; ++l)
    CHECK(f[l] == t[l & 63u] + t[64u + (l >> 6) * 7u + (uint32_t)__builtin_popcount(l & 63u)]);
}

static void test_oracle() {
  // 8-site Heisenberg ring, Sz = 0, reduced by translations at k = 1 and by
  // translations x flip; apply (both forms) and sparse (serial and parallel)
  const int L = 8;
  std::vector<uint64_t> parent;
  for (uint64_t b = 0; b < (1ull << L); ++b)
    if (__builtin_popcountll(b) == L / 2) parent.push_back(b);
  const int64_t np = (int64_t)parent.size();
  std::vector<uint64_t> mask, row, col;
  std::vector<double> are, aim;
  for (int i = 0; i < L; ++i) {
    const uint64_t bi = 1ull << i, bj = 1ull << ((i + 1) % L), m = bi | bj;
    const uint64_t rc[6][2] = {{0, 0}, {bi, bi}, {bj, bj}, {m, m}, {bi, bj}, {bj, bi}};
    const double a[6] = {1, -1, -1, 1, 2, 2};
    for (int t = 0; t < 6; ++t) {
      mask.push_back(m);
      row.push_back(rc[t][0]);
      col.push_back(rc[t][1]);
      are.push_back(a[t]);
      aim.push_back(0.0);
    }
  }
  const int64_t nt = (int64_t)mask.size();
  std::vector<int32_t> offs(L + 1), widths(L, 1);
  for (int i = 0; i <= L; ++i) offs[i] = i;
  for (int flip = 0; flip < 2; ++flip) {
    const int G = L * (flip ? 2 : 1);
    std::vector<int32_t> maps((size_t)G * L);
    std::vector<uint64_t> fl(G, 0ull);
    std::vector<double> cre(G), cim(G);
    for (int f = 0; f <= flip; ++f)
      for (int a = 0; a < L; ++a) {
        const int g = f * L + a;
        for (int i = 0; i < L; ++i) maps[(size_t)g * L + i] = (i + a) % L;
        fl[g] = f ? ((1ull << L) - 1ull) : 0ull;
        const double ph = -2.0 * M_PI * a / L;
        cre[g] = std::cos(ph);
        cim[g] = std::sin(ph);
      }
    std::vector<uint64_t> red(np);
    std::vector<int64_t> midx(np);
    std::vector<double> mre(np), mim(np);
    const int64_t nr = orc_symmetry_reduce_x(np, parent.data(), L, offs.data(), widths.data(), G, maps.data(),
                                             flip ? fl.data() : nullptr, cre.data(), cim.data(), 1e-8, red.data(),
                                             midx.data(), mre.data(), mim.data());
    CHECK(nr > 0 && nr < np);
    if (nr <= 0) continue;
    std::vector<double> x(2 * (size_t)nr), y1(2 * (size_t)nr, 0.0), y2(2 * (size_t)nr, 0.0);
    for (int64_t r = 0; r < nr; ++r) {
      x[2 * r] = 0.1 * (double)(r % 7) - 0.3;
      x[2 * r + 1] = 0.05 * (double)(r % 3);
    }
    orc_apply(np, parent.data(), 1, nr, red.data(), midx.data(), mre.data(), mim.data(), nt, mask.data(), row.data(),
              col.data(), are.data(), aim.data(), 0, x.data(), y1.data());
    orc_apply(np, parent.data(), 1, nr, red.data(), midx.data(), mre.data(), mim.data(), nt, mask.data(), row.data(),
              col.data(), are.data(), aim.data(), 1, x.data(), y2.data());
    const int64_t cap = nr * nt;
    std::vector<int64_t> colptr(nr + 1), rowval(cap);
    std::vector<double> nre(cap), nim(cap);
    const int64_t nnz = orc_sparse(np, parent.data(), 1, nr, red.data(), midx.data(), mre.data(), mim.data(), nt,
                                   mask.data(), row.data(), col.data(), are.data(), aim.data(), 1.5e-8, colptr.data(),
                                   rowval.data(), nre.data(), nim.data(), cap);
    CHECK(nnz > 0);
    for (int threads : {1, 3, 16}) {
      int64_t *cp = nullptr, *rv = nullptr;
      double *pr = nullptr, *pi = nullptr;
      const int64_t nnz2 = orc_sparse_parallel(np, parent.data(), 1, nr, red.data(), midx.data(), mre.data(),
                                               mim.data(), nt, mask.data(), row.data(), col.data(), are.data(),
                                               aim.data(), 1.5e-8, threads, &cp, &rv, &pr, &pi);
      CHECK(nnz2 == nnz);
      if (nnz2 == nnz) {
        for (int64_t c = 0; c <= nr; ++c) CHECK(cp[c] == colptr[c]);
        for (int64_t k = 0; k < nnz; ++k) CHECK(rv[k] == rowval[k] && pr[k] == nre[k] && pi[k] == nim[k]);
      }
      orc_free(cp);
      orc_free(rv);
      orc_free(pr);
      orc_free(pi);
    }
    // a Hermitian operator: the column form is the conjugate transpose product
    double dmax = 0.0;
    for (int64_t r = 0; r < 2 * nr; ++r) dmax = std::max(dmax, std::fabs(y1[r]));
    CHECK(dmax > 0.0);
  }
}

// Lanczos on a dense random symmetric matrix (n = 300) with the three-term
// recurrence only, beside the omega recurrence of the partial
// reorthogonalisation.  The recurrence without its rounding term is an
// identity for the actual overlaps up to local rounding: fed the measured
// overlaps of v_j and v_{j-1}, it predicts those of v_{j+1} to 1e-3 where
// they exceed 1e-9 (a dropped or wrong term misses by far more).  Its
// estimates (rounding term included) do not underestimate the overlaps by
// more than a factor 10 while they stay below 1e-4.  With the partial
// reorthogonalisation (the whole basis when an estimate exceeds sqrt(eps),
// and the next vector too) the basis stays semi-orthogonal and the lowest
// Ritz value matches the dense eigensolve.
static void test_omega_recurrence() {
  const int n = 300, m = 120;
  std::mt19937_64 rng(11);
  std::normal_distribution<double> nd;
  std::vector<double> A((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = nd(rng) / std::sqrt((double)n);
  for (int i = 0; i < n; ++i) A[(size_t)i * n + i] += 0.05 * i;  // spread spectrum, a clear lowest eigenvalue
  std::vector<double> w_ref, Z;
  {
    std::vector<double> B = A, d(n), e(n, 0.0);
    edhip::householder_tridiag(n, B, Z, d, e);
    CHECK(edhip::tridiag_ql(n, d, e, Z));
    w_ref = d;
  }
  const double eps = 0x1p-53, thr = std::sqrt(eps);
  for (int partial = 0; partial < 2; ++partial) {
    std::vector<std::vector<double>> V(1, std::vector<double>(n));
    double nrm = 0.0;
    for (double& x : V[0]) nrm += (x = nd(rng)) * x;
    for (double& x : V[0]) x /= std::sqrt(nrm);
    std::vector<double> alpha, beta, om_prev, om_cur(1, 1.0), om_new;
    std::vector<double> W_prev, W_cur(1, 1.0), W_pred;  // measured overlaps, and their prediction
    int predicted = 0;
    bool force = false;
    int full = 0;
    double worst_actual = 0.0;
    for (int j = 0; j < m; ++j) {
      std::vector<double> w(n, 0.0);
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) w[(size_t)r] += A[(size_t)r * n + c] * V[(size_t)j][(size_t)c];
      double a = 0.0;
      for (int r = 0; r < n; ++r) a += w[(size_t)r] * V[(size_t)j][(size_t)r];
      for (int r = 0; r < n; ++r) w[(size_t)r] -= a * V[(size_t)j][(size_t)r] + (j ? beta[(size_t)j - 1] * V[(size_t)j - 1][(size_t)r] : 0.0);
      alpha.push_back(a);
      double b = 0.0;
      for (double x : w) b += x * x;
      b = std::sqrt(b);
      const double est = edhip::omega_step(
          j, [&](int q) { return alpha[(size_t)q]; }, [&](int q) { return beta[(size_t)q]; }, b, om_cur, om_prev, eps,
          om_new);
      if (!partial)
        (void)edhip::omega_step(
            j, [&](int q) { return alpha[(size_t)q]; }, [&](int q) { return beta[(size_t)q]; }, b, W_cur, W_prev, 0.0,
            W_pred);
      if (partial && (force || est > thr)) {
        for (int pass = 0; pass < 2; ++pass)
          for (int q = 0; q <= j; ++q) {
            double c = 0.0;
            for (int r = 0; r < n; ++r) c += V[(size_t)q][(size_t)r] * w[(size_t)r];
            for (int r = 0; r < n; ++r) w[(size_t)r] -= c * V[(size_t)q][(size_t)r];
          }
        b = 0.0;
        for (double x : w) b += x * x;
        b = std::sqrt(b);
        for (int q = 0; q <= j; ++q) om_new[(size_t)q] = eps;
        force = !force;
        ++full;
      }
      beta.push_back(b);
      std::vector<double> v(n);
      for (int r = 0; r < n; ++r) v[(size_t)r] = w[(size_t)r] / b;
      double actual = 0.0;  // max overlap of the new vector with the older ones
      std::vector<double> W_next((size_t)j + 2, 0.0);
      for (int q = 0; q <= j; ++q) {
        double c = 0.0;
        for (int r = 0; r < n; ++r) c += V[(size_t)q][(size_t)r] * v[(size_t)r];
        W_next[(size_t)q] = c;
        if (q < j) actual = std::max(actual, std::fabs(c));
        if (!partial && q < j && std::fabs(c) > 1e-9) {
          CHECK(std::fabs(W_pred[(size_t)q] - c) <= 1e-3 * std::fabs(c));
          ++predicted;
        }
      }
      W_next[(size_t)j + 1] = 1.0;
      W_prev.swap(W_cur);
      W_cur.swap(W_next);
      double est_now = 0.0;
      for (int q = 0; q < j; ++q) est_now = std::max(est_now, std::fabs(om_new[(size_t)q]));
      if (!partial && est_now < 1e-4) CHECK(actual <= 10.0 * est_now + 1e-14);
      worst_actual = std::max(worst_actual, actual);
      V.push_back(v);
      om_prev.swap(om_cur);
      om_cur.swap(om_new);
    }
    if (partial) {
      CHECK(worst_actual < 10.0 * thr);
      CHECK(full > 0 && full < m / 2);
      std::vector<double> d(alpha), e(m, 0.0), Zt((size_t)m * m, 0.0);
      for (int i = 0; i + 1 < m; ++i) e[(size_t)i] = beta[(size_t)i];
      for (int i = 0; i < m; ++i) Zt[(size_t)i * m + i] = 1.0;
      CHECK(edhip::tridiag_ql(m, d, e, Zt));
      CHECK(std::fabs(d[0] - w_ref[0]) < 1e-10 * std::max(1.0, std::fabs(w_ref[0])));
    } else {
      CHECK(worst_actual > thr);  // without it the basis does lose orthogonality here
      CHECK(predicted > 100);     // overlaps above 1e-9 were predicted and checked
    }
  }
}

int main() {
  test_linear_algebra();
  test_omega_recurrence();
  test_group_closure();
  test_colex();
  test_oracle();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("sanitized host checks passed\n");
  return 0;
}
