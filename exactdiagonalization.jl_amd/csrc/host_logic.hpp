// Host-side logic with no device code, so that a host-only build (with
// sanitizers) can test it: the linear algebra of the eigensolver's small
// projected (tridiagonal or arrowhead) matrices, and the closure and
// inverses of a symmetry group given as bit permutations with XOR words.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace edhip {

// Householder reduction of a dense symmetric matrix (row-major, n x n) to
// tridiagonal form A = Q T Q^T; returns d (diagonal), e (e[i] = T[i][i+1]).
inline void householder_tridiag(int n, std::vector<double>& A, std::vector<double>& Q, std::vector<double>& d,
                                std::vector<double>& e) {
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  std::vector<double> v(n), p(n), w(n);
  for (int k = 0; k + 2 < n; ++k) {
    double nx = 0.0;
    for (int i = k + 1; i < n; ++i) nx += A[(size_t)i * n + k] * A[(size_t)i * n + k];
    nx = std::sqrt(nx);
    if (nx == 0.0) continue;
    double x0 = A[(size_t)(k + 1) * n + k];
    double alpha = (x0 >= 0) ? -nx : nx;
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    for (int i = k + 1; i < n; ++i) v[i] = A[(size_t)i * n + k];
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = k + 1; i < n; ++i) v[i] /= vn;
    // p = A v ; K = v^T p ; w = p - K v ; A -= 2 (v w^T + w v^T)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * v[j];
      p[i] = s;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; ++i) K += v[i] * p[i];
    for (int i = 0; i < n; ++i) w[i] = p[i] - K * v[i];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) A[(size_t)i * n + j] -= 2.0 * (v[i] * w[j] + w[i] * v[j]);
    // Q = Q H
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += Q[(size_t)i * n + j] * v[j];
      for (int j = k + 1; j < n; ++j) Q[(size_t)i * n + j] -= 2.0 * s * v[j];
    }
  }
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = A[(size_t)i * n + i + 1];
}

// Implicit QL with Wilkinson-type shifts on a symmetric tridiagonal matrix
// (d diagonal, e[i] coupling i and i+1): the published tql2 procedure of
// EISPACK (Bowdler, Martin, Reinsch and Wilkinson, "The QR and QL
// algorithms for symmetric matrices", Numer. Math. 11 (1968) 293-306).  Z (row-major n x n) accumulates the
// rotations: on entry the basis the tridiagonal matrix is expressed in, on
// exit the eigenvectors as columns.  Eigenvalues are sorted ascending.
inline bool tridiag_ql(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& Z) {
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) return false;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool early = false;
      for (; i >= l; --i) {
        double f = s * e[i];
        double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          double zf = Z[(size_t)k * n + i + 1];
          Z[(size_t)k * n + i + 1] = s * Z[(size_t)k * n + i] + c * zf;
          Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * zf;
        }
      }
      if (early) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // sort ascending
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> d2(n), Z2((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    d2[j] = d[ord[j]];
    for (int k = 0; k < n; ++k) Z2[(size_t)k * n + j] = Z[(size_t)k * n + ord[j]];
  }
  d = d2;
  Z = Z2;
  return true;
}

// Closure of a group of G elements (bit permutation perm[g*nb + i]: bit i of
// b moves to bit perm[g*nb + i]; then XOR with xr[g]) and the index of each
// element's inverse.  An element (p, x) acts as b -> p(b) ^ x; (p1, x1)(p2, x2)
// = (p1 p2, p1(x2) ^ x1) = (p1 p2, x2 ^ x1) since every x is 0 or the full
// word, which every permutation fixes.  Returns an empty string, or the
// reason the elements are not a group: a repeated element, or a product
// outside the set.
inline std::string group_inverses(int G, int nb, const std::vector<int32_t>& perm, const std::vector<uint64_t>& xr,
                                  std::vector<int32_t>& inv) {
  std::map<std::pair<std::vector<int32_t>, uint64_t>, int> index;
  for (int g = 0; g < G; ++g) {
    std::vector<int32_t> p(perm.begin() + (size_t)g * nb, perm.begin() + (size_t)(g + 1) * nb);
    auto key = std::make_pair(p, xr[g]);
    if (index.count(key)) return "repeated symmetry operation";
    index[key] = g;
  }
  inv.assign(G, -1);
  std::vector<int32_t> comp(nb);
  for (int g1 = 0; g1 < G; ++g1)
    for (int g2 = 0; g2 < G; ++g2) {
      for (int i = 0; i < nb; ++i) comp[i] = perm[(size_t)g1 * nb + perm[(size_t)g2 * nb + i]];
      auto it = index.find(std::make_pair(comp, xr[g1] ^ xr[g2]));
      if (it == index.end()) return "symmetry operations must form a group";
      if (it->second == 0) inv[g1] = g2;
    }
  return std::string();
}

// Colex (combinatorial number system) rank of a 12-bit word among the 12-bit
// words of its popcount j, in [0, C(12, j)) with C(12, j) <= 924, from two
// 6-bit halves: rank(l) = lo[l & 63] + hi[(l >> 6) * 7 + popcount(l & 63)],
// lo[v] = sum_i C(p_i, i) over the set bits p_1 < p_2 < ... of v, and
// hi[v * 7 + m] = sum_i C(6 + q_i, m + i) over the set bits q_1 < ... of v.
// Returns the 64 + 64 * 7 entries (lo, then hi).
constexpr int kColexStride = 924;  // C(12, 6): room for every popcount class
// One step of the omega recurrence of partial reorthogonalisation (Simon,
// "The Lanczos algorithm with partial reorthogonalization", Math. Comp. 42
// (1984) 115-142).  Given the estimates om_cur[q] ~ <v_j, v_q> (q <= j) and
// om_prev[q] ~ <v_{j-1}, v_q> (q < j), the Lanczos coefficients alpha(q)
// (q <= j) and beta(q) (q < j: the coupling of v_q and v_{q+1}) and the new
// beta_j > 0, writes om_new[q] ~ <v_{j+1}, v_q> for q <= j + 1:
//   beta_j om_new[q] = beta(q) om_cur[q+1] + (alpha(q) - alpha(j)) om_cur[q]
//                      + beta(q-1) om_cur[q-1] - beta(j-1) om_prev[q]
// plus a rounding term eps (beta(q) + beta_j) of the sum's sign, with
// om_new[j] = eps and om_new[j+1] = 1.  Returns max |om_new[q]| over q < j.
template <typename Alpha, typename Beta>
inline double omega_step(int j, Alpha alpha, Beta beta, double beta_j, const std::vector<double>& om_cur,
                         const std::vector<double>& om_prev, double eps, std::vector<double>& om_new) {
  om_new.assign((size_t)j + 2, 0.0);
  const double a_j = alpha(j);
  const double b_jm1 = j > 0 ? beta(j - 1) : 0.0;
  double worst = 0.0;
  for (int q = 0; q < j; ++q) {
    const double b_q = beta(q);
    double v = b_q * om_cur[(size_t)q + 1] + (alpha(q) - a_j) * om_cur[(size_t)q] - b_jm1 * om_prev[(size_t)q];
    if (q > 0) v += beta(q - 1) * om_cur[(size_t)q - 1];
    v += std::copysign(eps * (b_q + beta_j), v);
    om_new[(size_t)q] = v / beta_j;
    worst = std::max(worst, std::fabs(om_new[(size_t)q]));
  }
  om_new[(size_t)j] = eps;
  om_new[(size_t)j + 1] = 1.0;
  return worst;
}

inline std::vector<uint16_t> colex12_tables() {
  auto binom = [](int n, int k) -> uint32_t {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return (uint32_t)r;
  };
  std::vector<uint16_t> t(64 + 64 * 7, 0);
  for (int v = 0; v < 64; ++v) {
    uint32_t r = 0;
    int i = 0;
    for (int p = 0; p < 6; ++p)
      if ((v >> p) & 1) r += binom(p, ++i);
    t[v] = (uint16_t)r;
    for (int m = 0; m <= 6; ++m) {
      uint32_t h = 0;
      int q = 0;
      for (int p = 0; p < 6; ++p)
        if ((v >> p) & 1) h += binom(6 + p, m + (++q));
      t[64 + v * 7 + m] = (uint16_t)h;
    }
  }
  return t;
}

// The rank of every 12-bit word (4096 entries) from the two halves above:
// one table read per word on the device.
inline std::vector<uint16_t> colex12_full() {
  const std::vector<uint16_t> t = colex12_tables();
  std::vector<uint16_t> f(4096);
  for (uint32_t l = 0; l < 4096; ++l)
    f[l] = (uint16_t)(t[l & 63u] + t[64u + (l >> 6) * 7u + (uint32_t)__builtin_popcount(l & 63u)]);
  return f;
}

}  // namespace edhip
