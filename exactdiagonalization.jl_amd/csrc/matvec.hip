// Matrix-free application of an operator on a (possibly symmetry-reduced)
// basis, in gather form with no atomics: one lane owns one output entry.
//
// Row form (apply!(out, opr, state), abstract_operator_representation.jl:296-316):
//   y[r] += sum_t [ (b_r & m_t) == R_t ] w * x[s],  image b' = (b_r & ~m_t) | C_t
// Column form (apply!(out, state, opr), :327-347) swaps R_t and C_t and uses
// the column element of reduced_operator_representation.jl:88-116.
//
// For a reduced basis, b' is canonicalised to its representative s = h(b'),
// and the element is a * phase[h] * sqrt(N_r / N_s) (row form) with
// phase[h] = conj(chi(h^-1)); this equals a * U[b'] / U[b_r] of
// reduced_operator_representation.jl:57-85 with U[g(s)] = conj(chi(g))/sqrt(N_s).
#include <type_traits>

#include "kernels.hpp"

namespace edhip {

namespace {

constexpr int kBlock = 256;

template <int MODE>
struct Canon {
  // NEED_H = false: h is left 0 where the canonicaliser can skip tracking it
  template <bool NEED_H = true>
  __device__ __forceinline__ static void run(const GroupDev& g, const GroupLds& L, uint64_t b, uint64_t& s, int& h,
                                             int& cnt) {
    if constexpr (MODE == kGroupNone) {
      s = b;
      h = 0;
      cnt = 1;
    } else if constexpr (MODE == kGroupTorusPairs) {
      canonicalize_torus_pairs<0, 0, NEED_H>(g, L, b, s, h, cnt);
    } else if constexpr (MODE == kTorusPairs6x6) {
      canonicalize_torus_pairs<6, 6, NEED_H>(g, L, b, s, h, cnt);
    } else if constexpr (MODE == kTorusPairs4x6) {
      canonicalize_torus_pairs<4, 6, NEED_H>(g, L, b, s, h, cnt);
    } else if constexpr (MODE == kTorusPairs6x4) {
      canonicalize_torus_pairs<6, 4, NEED_H>(g, L, b, s, h, cnt);
    } else if constexpr (MODE == kTorusPairs4x4) {
      canonicalize_torus_pairs<4, 4, NEED_H>(g, L, b, s, h, cnt);
    } else if constexpr (MODE >= kTransPairs6x6 && MODE <= kTransPairs4x4) {
      // translations only: the identity coset, whose element table maps the
      // translation found to the group element
      constexpr int LX = (MODE == kTransPairs6x6 || MODE == kTransPairs6x4) ? 6 : 4;
      constexpr int LY = (MODE == kTransPairs6x6 || MODE == kTransPairs4x6) ? 6 : 4;
      int pos;
      canonicalize_torus_pairs_t<LX, LY, NEED_H>(g, L, L.coset_inv, b, s, pos, cnt);
      h = NEED_H ? L.coset_elem[pos] : 0;
    } else {
      canonicalize(g, L, b, s, h, cnt);
    }
  }
};

struct Staged {
  GroupLds L;
  const OffTerm* off;
  const uint64_t* W;
};

template <int MODE>
__device__ __forceinline__ Staged stage_op(const OpDev& op, unsigned char* lds) {
  Staged S;
  if constexpr (MODE != kGroupNone) {
    S.L = stage_group(op.g, op.lay, lds);
  }
  lds_copy(lds + op.off_lds, op.off, (uint32_t)(op.n_off * sizeof(OffTerm)));
  S.off = (const OffTerm*)(lds + op.off_lds);
  lds_copy(lds + op.w_lds, op.ix.W, op.w_bytes);
  S.W = (const uint64_t*)(lds + op.w_lds);
  return S;
}

// The matvec's copy of the off-diagonal terms in LDS, split by use: (mask,
// match) pairs read by every lane alike while matching a row, and (flip,
// amplitude) pairs read per matched term, where flip = match ^ image gives the
// image of a matching row b as b ^ flip; imaginary parts apart.
struct MvTerms {
  const uint64_t* mm;  // term by term: mask, match; paired (n_mv): mask, match 2q, match 2q+1
  const uint64_t* fa;  // 2 per term: flip, bits of Re(amplitude)
  const double* ai;    // Im(amplitude)
};

// paired: the padded copy (op.n_mv > 0) with one mask per pair of terms;
// else the terms of op.off one by one, indexed as in op.off.  amp32: the real
// amplitude is kept as an FP32 value in the low half of its slot.
__device__ __forceinline__ MvTerms stage_mv_terms(const OpDev& op, unsigned char* lds, bool paired,
                                                  bool amp32 = false) {
  const int n = paired ? op.n_mv : op.n_off;
  const OffTerm* src = paired ? op.mv_off : op.off;
  uint64_t* mm = (uint64_t*)(lds + op.off_lds);  // at most 40 bytes per term, as reserved
  uint64_t* fa = mm + (paired ? 3 * (n / 2) : 2 * n);
  double* ai = (double*)(fa + 2 * n);
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const OffTerm o = src[t];
    if (paired) {
      if ((t & 1) == 0) {
        mm[3 * (t >> 1)] = o.mask;
        mm[3 * (t >> 1) + 1] = o.match;
      } else {
        mm[3 * (t >> 1) + 2] = o.match;
      }
    } else {
      mm[2 * t] = o.mask;
      mm[2 * t + 1] = o.match;
    }
    fa[2 * t] = o.match ^ o.image;
    fa[2 * t + 1] = amp32 ? (uint64_t)__float_as_uint((float)o.are) : (uint64_t)__double_as_longlong(o.are);
    ai[t] = o.aim;
  }
  return MvTerms{mm, fa, ai};
}

// Weight of an off-diagonal element: a * phase^(conj if TRANS) * sqrt(N_own) / sqrt(N_img)
// (SCALED: without 1 / sqrt(N_img), which the spread x / sqrt(N) carries)
template <int MODE, bool TRANS, bool UNIT = false, bool SCALED = false>
__device__ __forceinline__ void element_weight_a(const OpDev& op, const GroupLds& L, double are, double aim, int h,
                                                 int cnt, double sqrt_own, double& wr, double& wi) {
  if constexpr (MODE == kGroupNone) {
    wr = are;
    wi = aim;
  } else {
    const double sc = SCALED ? sqrt_own : sqrt_own * L.sqcnt[cnt];
    if (UNIT || op.g.unit_phases) {
      wr = are * sc;
      wi = aim * sc;
    } else {
      const double pr = L.ph_re[h];
      const double pi = TRANS ? -L.ph_im[h] : L.ph_im[h];
      wr = (are * pr - aim * pi) * sc;
      wi = (are * pi + aim * pr) * sc;
    }
  }
}

template <int MODE, bool TRANS>
__device__ __forceinline__ void element_weight(const OpDev& op, const Staged& S, const OffTerm& t, int h, int cnt,
                                               double sqrt_own, double& wr, double& wi) {
  if constexpr (MODE == kGroupNone) {
    wr = t.are;
    wi = t.aim;
  } else {
    double pr = S.L.ph_re[h];
    double pi = TRANS ? -S.L.ph_im[h] : S.L.ph_im[h];
    double sc = sqrt_own * S.L.sqcnt[cnt];
    wr = (t.are * pr - t.aim * pi) * sc;
    wi = (t.are * pi + t.aim * pr) * sc;
  }
}

__device__ __forceinline__ void diag_sum(const OpDev& op, uint64_t b, double& dre, double& dim_) {
  for (int t = 0; t < op.n_diag; ++t) {
    const DiagTerm d = op.diag[t];
    if ((b & d.mask) == d.match) {
      dre += d.are;
      dim_ += d.aim;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_diag_values(OpDev op, double* __restrict__ re, double* __restrict__ im) {
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (row >= op.dim) return;
  double dre = 0.0, dim_ = 0.0;
  diag_sum(op, op.rows[row], dre, dim_);
  re[row] = dre;
  if (im) im[row] = dim_;
}

// The matvec runs 512-lane blocks, one row per lane, over a grid that covers
// the rows in order (blocks that run together work on neighbouring rows, whose
// images share cache lines of the index and of x).
constexpr int kMvBlock = 512;

// UNIT: every character is 1 (op.g.unit_phases): the weight needs no phase,
// so the canonicaliser does not track the element reaching the image.
// SCALED: x is read from its spread holding x / sqrt(N) (op.xm.scaled), so the
// weight needs no orbit size of the image and the stabiliser goes uncounted.
template <typename V, int MODE, bool TRANS, bool UNIT, bool SCALED>
__global__ __launch_bounds__(kMvBlock) void k_matvec(OpDev op, const V* __restrict__ x, V* __restrict__ y, int accumulate) {
  extern __shared__ unsigned char lds[];
  GroupLds L{};
  if constexpr (MODE != kGroupNone) L = stage_group(op.g, op.lay, lds);
  // FP32 weights: a real FP32 product whose images need no phase and no orbit
  // size (unit characters, scaled spread) weighs each image a * sqrt(N_r) in
  // FP32, from FP32 amplitudes (no FP64 multiply and conversion per image)
  constexpr bool kW32 = UNIT && SCALED && std::is_same<V, float>::value;
  const MvTerms T = stage_mv_terms(op, lds, op.n_mv != 0, kW32);
  lds_copy(lds + op.w_lds, op.ix.W, op.w_bytes);
  const uint64_t* W = (const uint64_t*)(lds + op.w_lds);
  // code table of the spread x (the host sets xm only with room for it)
  uint16_t* codes = (uint16_t*)(lds + op.codes_lds);
  if (op.xm.values) xmap_stage(op.xm, codes);
  __syncthreads();
  const int64_t row = op.row_begin + xcd_chunked_block(blockIdx.x, gridDim.x, op.xcd_chunk) * kMvBlock + threadIdx.x;
  if (row >= op.dim) return;
  using R = typename VT<V>::R;
  const uint64_t b = op.rows[row];

  // diagonal terms (C == R): the element is the summed amplitude (U[b]/U[b] = 1)
  double dre = 0.0, dim_ = 0.0;
  if (op.diag_re) {
    dre = op.diag_re[row];
    if (op.diag_im) dim_ = op.diag_im[row];
  } else {
    diag_sum(op, b, dre, dim_);
  }
  V acc = vzero<V>();
  vfma(acc, (R)dre, (R)dim_, x[row]);

  double sqrt_own = 1.0;
  if constexpr (MODE != kGroupNone) sqrt_own = sqrt((double)op.orbit[row]);
  const float so32 = (float)sqrt_own;

  const int n_terms = op.n_mv ? op.n_mv : op.n_off;
  for (int seg = 0; seg < n_terms; seg += 64) {
    const int nseg = (n_terms - seg) < 64 ? (n_terms - seg) : 64;
    uint64_t bits = 0;
    if (op.n_mv) {  // two terms with one mask per test (nseg is even)
      for (int j = 0; j < nseg; j += 2) {
        const uint64_t* m = T.mm + 3 * ((seg + j) >> 1);
        const uint64_t xm = b & m[0];
        bits |= (uint64_t)((uint32_t)(xm == m[1]) | ((uint32_t)(xm == m[2]) << 1)) << j;
      }
    } else {
      for (int j = 0; j < nseg; ++j) {
        const uint64_t* m = T.mm + 2 * (seg + j);
        bits |= (uint64_t)((b & m[0]) == m[1]) << j;
      }
    }
    // two images per step: both canonical forms first, then both lookups and
    // both x gathers issued before either is used (twice the loads in flight;
    // measured 5% faster at 6x6 than one image per step, three or four per
    // step are slower: more steps end with an empty slot)
    constexpr int NW = 2;
    while (bits) {
      int t[NW];
      bool ok[NW];
      t[0] = seg + __builtin_ctzll(bits);  // bits != 0 here
      ok[0] = true;
      bits &= bits - 1;
#pragma unroll
      for (int k = 1; k < NW; ++k) {  // an empty slot repeats the first image and is dropped
        ok[k] = bits != 0;
        t[k] = ok[k] ? seg + __builtin_ctzll(bits) : t[0];
        bits &= ok[k] ? bits - 1 : bits;
      }
      ulonglong2 fa[NW];  // flip = match ^ image: a matching row b has the image b ^ flip
      uint64_t sk[NW];
      int hk[NW], ck[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        fa[k] = *(const ulonglong2*)(T.fa + 2 * t[k]);
        Canon<MODE>::template run<!UNIT>(op.g, L, b ^ fa[k].x, sk[k], hk[k], ck[k]);
      }
      int64_t ik[NW];
      V xv[NW];
      if (op.xm.values) {  // x of each image with one load at its word position
        bool f0, f1;
        xmap_fetch2(op.xm, codes, sk[0], sk[1], xv[0], xv[1], f0, f1);
        ik[0] = f0 ? 0 : -1;
        ik[1] = (ok[1] && f1) ? 0 : -1;
      } else {
        index_of2(op.ix, W, sk[0], sk[1], ik[0], ik[1]);
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          ik[k] = ok[k] ? ik[k] : -1;
          xv[k] = x[ik[k] >= 0 ? ik[k] : 0];
        }
      }
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        if (ik[k] >= 0) {
          if constexpr (kW32) {
            vfma(acc, __uint_as_float((uint32_t)fa[k].y) * so32, 0.f, xv[k]);
          } else {
            double wr, wi;
            element_weight_a<MODE, TRANS, UNIT, SCALED>(op, L, __longlong_as_double((long long)fa[k].y),
                                                  VT<V>::cx ? T.ai[t[k]] : 0.0, hk[k], ck[k], sqrt_own, wr, wi);
            vfma(acc, (R)wr, (R)wi, xv[k]);
          }
        }
      }
    }
  }
  V* yr = y + (row - op.row_begin);
  *yr = accumulate ? vadd(*yr, acc) : acc;
}

template <typename V, int MODE>
hipError_t launch_mv_t(const OpDev& op, int transpose, int accumulate, const void* x, void* y, hipStream_t st) {
  dim3 grid((unsigned)((op.dim - op.row_begin + kMvBlock - 1) / kMvBlock));
  // unit-phase kernels only where the canonicaliser can drop the element (row-pair tori)
  const bool unit = MODE >= kGroupTorusPairs && op.g.unit_phases;
  const uint32_t lds = op.xm.values ? op.codes_lds + xmap_lds_bytes(op.xm) : op.lds_total;
#define EDHIP_MV(T, U, S)                                                                                              \
  hipLaunchKernelGGL((k_matvec<V, MODE, T, U, S>), grid, dim3(kMvBlock), lds, st, op, (const V*)x, (V*)y, accumulate)
  if constexpr (MODE != kGroupNone) {
    if (op.xm.values && op.xm.scaled) {
      if (transpose) {
        if (unit) EDHIP_MV(true, MODE >= kGroupTorusPairs, true); else EDHIP_MV(true, false, true);
      } else {
        if (unit) EDHIP_MV(false, MODE >= kGroupTorusPairs, true); else EDHIP_MV(false, false, true);
      }
      return hipGetLastError();
    }
  }
  if (transpose) {
    if (unit) EDHIP_MV(true, MODE >= kGroupTorusPairs, false); else EDHIP_MV(true, false, false);
  } else {
    if (unit) EDHIP_MV(false, MODE >= kGroupTorusPairs, false); else EDHIP_MV(false, false, false);
  }
#undef EDHIP_MV
  return hipGetLastError();
}

template <typename V>
hipError_t launch_mv_v(const OpDev& op, int transpose, int accumulate, const void* x, void* y, hipStream_t st) {
  switch (op.g.mode) {
    case kGroupNone:
      return launch_mv_t<V, kGroupNone>(op, transpose, accumulate, x, y, st);
    case kGroupGeneric:
      return launch_mv_t<V, kGroupGeneric>(op, transpose, accumulate, x, y, st);
    default:
      if (op.g.pair_table) {
        const int Lx = op.g.Lx, Ly = op.g.Ly;
        if (op.g.n_cosets == 1) {
          if (Lx == 6 && Ly == 6) return launch_mv_t<V, kTransPairs6x6>(op, transpose, accumulate, x, y, st);
          if (Lx == 4 && Ly == 6) return launch_mv_t<V, kTransPairs4x6>(op, transpose, accumulate, x, y, st);
          if (Lx == 6 && Ly == 4) return launch_mv_t<V, kTransPairs6x4>(op, transpose, accumulate, x, y, st);
          if (Lx == 4 && Ly == 4) return launch_mv_t<V, kTransPairs4x4>(op, transpose, accumulate, x, y, st);
        }
        if (Lx == 6 && Ly == 6) return launch_mv_t<V, kTorusPairs6x6>(op, transpose, accumulate, x, y, st);
        if (Lx == 4 && Ly == 6) return launch_mv_t<V, kTorusPairs4x6>(op, transpose, accumulate, x, y, st);
        if (Lx == 6 && Ly == 4) return launch_mv_t<V, kTorusPairs6x4>(op, transpose, accumulate, x, y, st);
        if (Lx == 4 && Ly == 4) return launch_mv_t<V, kTorusPairs4x4>(op, transpose, accumulate, x, y, st);
        return launch_mv_t<V, kGroupTorusPairs>(op, transpose, accumulate, x, y, st);
      }
      return launch_mv_t<V, kGroupTorus>(op, transpose, accumulate, x, y, st);
  }
}

// ---------------------------------------------------------------------------
// Raw iterators (index or -1, value) in term order, for a list of rows.
// Reduced invalid images carry the bare amplitude, as at
// reduced_operator_representation.jl:71-74.
// ---------------------------------------------------------------------------
template <int MODE, bool TRANS>
__global__ __launch_bounds__(kBlock) void k_iterate(OpDev op, const OffTerm* __restrict__ terms, int n_terms,
                                                    int64_t n_rows, const int64_t* __restrict__ rows,
                                                    int64_t* __restrict__ idx_out, double* __restrict__ re_out,
                                                    double* __restrict__ im_out) {
  extern __shared__ unsigned char lds[];
  Staged S = stage_op<MODE>(op, lds);
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n_rows) return;
  const int64_t row = rows[k];
  const uint64_t b = op.rows[row];
  double sqrt_own = 1.0;
  if constexpr (MODE != kGroupNone) sqrt_own = sqrt((double)op.orbit[row]);
  int64_t n = 0;
  const int64_t base = k * (int64_t)n_terms;
  for (int ti = 0; ti < n_terms; ++ti) {
    const OffTerm t = terms[ti];
    if ((b & t.mask) != t.match) continue;
    const uint64_t img = (b & ~t.mask) | t.image;
    uint64_t s;
    int h, cnt;
    Canon<MODE>::run(op.g, S.L, img, s, h, cnt);
    const int64_t idx = index_of(op.ix, S.W, s);
    double wr = t.are, wi = t.aim;
    if (idx >= 0) element_weight<MODE, TRANS>(op, S, t, h, cnt, sqrt_own, wr, wi);
    idx_out[base + n] = idx;
    re_out[base + n] = wr;
    im_out[base + n] = wi;
    ++n;
  }
  for (int64_t j = n; j < n_terms; ++j) idx_out[base + j] = -2;  // end marker
}

// ---------------------------------------------------------------------------
// Sparse export: per column, collect (row, value), merge duplicates, drop
// |v| < tol, sort rows (sparse_serial, abstract_operator_representation.jl:145-169).
// The same walk over selected rows (row-form operator, TRANS = false) or
// columns (column form, TRANS = true) gives get_row / get_column (:217-243).
// ---------------------------------------------------------------------------
template <int MODE, bool TRANS>
__global__ __launch_bounds__(kBlock) void k_sparse_columns(OpDev op, double tol, int slot, int64_t n_sel,
                                                           const int64_t* __restrict__ sel, int64_t* __restrict__ sidx,
                                                           double* __restrict__ sre, double* __restrict__ sim,
                                                           int64_t* __restrict__ count) {
  extern __shared__ unsigned char lds[];
  Staged S = stage_op<MODE>(op, lds);
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n_sel) return;
  const int64_t col = sel ? sel[k] : k;
  const uint64_t b = op.rows[col];
  int64_t* I = sidx + k * (int64_t)slot;
  double* Re = sre + k * (int64_t)slot;
  double* Im = sim + k * (int64_t)slot;
  int n = 0;
  double dre = 0.0, dim_ = 0.0;
  bool any_diag = false;
  for (int t = 0; t < op.n_diag; ++t) {
    const DiagTerm d = op.diag[t];
    if ((b & d.mask) == d.match) {
      dre += d.are;
      dim_ += d.aim;
      any_diag = true;
    }
  }
  if (any_diag) {
    I[n] = col;
    Re[n] = dre;
    Im[n] = dim_;
    ++n;
  }
  double sqrt_own = 1.0;
  if constexpr (MODE != kGroupNone) sqrt_own = sqrt((double)op.orbit[col]);
  for (int ti = 0; ti < op.n_off; ++ti) {
    const OffTerm t = S.off[ti];
    if ((b & t.mask) != t.match) continue;
    const uint64_t img = (b & ~t.mask) | t.image;
    uint64_t s;
    int h, cnt;
    Canon<MODE>::run(op.g, S.L, img, s, h, cnt);
    const int64_t idx = index_of(op.ix, S.W, s);
    if (idx < 0) continue;
    double wr, wi;
    element_weight<MODE, TRANS>(op, S, t, h, cnt, sqrt_own, wr, wi);
    I[n] = idx;
    Re[n] = wr;
    Im[n] = wi;
    ++n;
  }
  // stable insertion sort by row index
  for (int i = 1; i < n; ++i) {
    int64_t ki = I[i];
    double ri = Re[i], ii = Im[i];
    int j = i - 1;
    while (j >= 0 && I[j] > ki) {
      I[j + 1] = I[j];
      Re[j + 1] = Re[j];
      Im[j + 1] = Im[j];
      --j;
    }
    I[j + 1] = ki;
    Re[j + 1] = ri;
    Im[j + 1] = ii;
  }
  // merge duplicates, chop
  int m = 0;
  int i = 0;
  while (i < n) {
    int64_t r = I[i];
    double sr = 0.0, si = 0.0;
    while (i < n && I[i] == r) {
      sr += Re[i];
      si += Im[i];
      ++i;
    }
    if (hypot(sr, si) < tol) continue;
    I[m] = r;
    Re[m] = sr;
    Im[m] = si;
    ++m;
  }
  count[k] = m;
}

// Matrix(opr) (abstract_operator_representation.jl:121-132): column icol of the
// column-major dense matrix is the sum of column icol's iterator entries,
// valid rows only, in iterator order (diagonal terms, then off-diagonal terms).
// One lane per column: no two lanes write the same element.  out is cleared.
template <int MODE, bool CX>
__global__ __launch_bounds__(kBlock) void k_dense_columns(OpDev op, double* __restrict__ out) {
  extern __shared__ unsigned char lds[];
  Staged S = stage_op<MODE>(op, lds);
  __syncthreads();
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= op.dim) return;
  const uint64_t b = op.rows[col];
  const int64_t base = col * op.dim;
  for (int t = 0; t < op.n_diag; ++t) {
    const DiagTerm d = op.diag[t];
    if ((b & d.mask) != d.match) continue;
    if (CX) {
      out[2 * (base + col)] += d.are;
      out[2 * (base + col) + 1] += d.aim;
    } else {
      out[base + col] += d.are;
    }
  }
  double sqrt_own = 1.0;
  if constexpr (MODE != kGroupNone) sqrt_own = sqrt((double)op.orbit[col]);
  for (int ti = 0; ti < op.n_off; ++ti) {
    const OffTerm t = S.off[ti];
    if ((b & t.mask) != t.match) continue;
    const uint64_t img = (b & ~t.mask) | t.image;
    uint64_t s;
    int h, cnt;
    Canon<MODE>::run(op.g, S.L, img, s, h, cnt);
    const int64_t idx = index_of(op.ix, S.W, s);
    if (idx < 0) continue;
    double wr, wi;
    element_weight<MODE, true>(op, S, t, h, cnt, sqrt_own, wr, wi);
    if (CX) {
      out[2 * (base + idx)] += wr;
      out[2 * (base + idx) + 1] += wi;
    } else {
      out[base + idx] += wr;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_sparse_compact(int64_t dim, int slot, const int64_t* __restrict__ colptr,
                                                           const int64_t* __restrict__ sidx, const double* __restrict__ sre,
                                                           const double* __restrict__ sim, int64_t* __restrict__ rowval,
                                                           double* __restrict__ re, double* __restrict__ im) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= dim) return;
  const int64_t o = colptr[col];
  const int64_t n = colptr[col + 1] - o;
  for (int64_t k = 0; k < n; ++k) {
    rowval[o + k] = sidx[col * (int64_t)slot + k];
    re[o + k] = sre[col * (int64_t)slot + k];
    im[o + k] = sim[col * (int64_t)slot + k];
  }
}

// ---------------------------------------------------------------------------
// Per-term bilinear forms (backward pass):
//   g_t = sum_r conj(phi_r) * sum_{images of term t from b_r} w_unit * psi_s
// where w_unit is the element weight with unit amplitude.  Output is
// interleaved (re, im) per term, diagonal terms first, then off-diagonal ones.
// ---------------------------------------------------------------------------

template <typename V>
__device__ __forceinline__ void load_c(const V* p, int64_t i, double& re, double& im) {
  if constexpr (VT<V>::cx) {
    re = (double)p[i].re;
    im = (double)p[i].im;
  } else {
    re = (double)p[i];
    im = 0.0;
  }
}

// LDS offset of the term pass's per-wave bins: after the operator's tables,
// and after the spread's code table when psi is read from the spread.
__host__ __device__ __forceinline__ uint32_t expect_bins_offset(const OpDev& op) {
  return op.xm.values ? op.codes_lds + xmap_lds_bytes(op.xm) : op.lds_total;
}

// blocks of the term pass (rows grid-stride; one row of partial sums each):
// 1024 measured 6% faster than 2048 on the 4x6 XXZ pass (half the partial
// sums to reduce, the same occupancy)
constexpr int kExpectBlocksMax = 1024;
constexpr int kExpectWaves = kBlock / 64;

// Per-term bilinear forms g_t = sum_r conj(phi_r) sum_{images of t} w psi_s
// (unit amplitudes), for the off-diagonal terms [t0, t0 + nw) of the combined
// numbering (diagonal terms first, t0 >= n_diag; the diagonal ones are
// k_diag_expect's).  One lane per row, as in the
// matvec: a lane walks its row's matched terms and adds each contribution to
// its wave's bin of that term in LDS (float64 LDS atomics; a wave only adds
// to its own bins, in program order).  Blocks walk the rows grid-stride; the
// block sums its waves' bins in a fixed order into partial[block][t][2].
template <typename V, int MODE>
__global__ __launch_bounds__(kBlock) void k_term_expect(OpDev op, const V* __restrict__ phi, const V* __restrict__ psi,
                                                        int t0, int nw, double* __restrict__ partial) {
  constexpr bool CX = VT<V>::cx;  // real vectors come with real weights
  constexpr int NC = CX ? 2 : 1;
  extern __shared__ unsigned char lds[];
  GroupLds L{};
  if constexpr (MODE != kGroupNone) L = stage_group(op.g, op.lay, lds);
  const MvTerms T = stage_mv_terms(op, lds, false);
  lds_copy(lds + op.w_lds, op.ix.W, op.w_bytes);
  const uint64_t* W = (const uint64_t*)(lds + op.w_lds);
  // psi of the images from its spread over the word positions (op.xm set):
  // the code table after the operator's tables, then the bins
  uint16_t* codes = (uint16_t*)(lds + op.codes_lds);
  if (op.xm.values) xmap_stage(op.xm, codes);
  double* bins = (double*)(lds + expect_bins_offset(op));  // [wave][nw][NC]
  for (int i = threadIdx.x; i < kExpectWaves * nw * NC; i += kBlock) bins[i] = 0.0;
  __syncthreads();
  double* mine = bins + (threadIdx.x >> 6) * nw * NC;
  const int n_diag = op.n_diag, n_terms = op.n_diag + op.n_off;
  const int o_lo = t0 - n_diag;                                        // off-diagonal terms [o_lo, o_hi)
  const int o_hi = (t0 + nw < n_terms ? t0 + nw : n_terms) - n_diag;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < op.dim; row += (int64_t)gridDim.x * kBlock) {
    const uint64_t b = op.rows[row];
    double fr, fi;
    load_c(phi, row, fr, fi);
    if (o_lo < o_hi) {
      double sqrt_own = 1.0;
      if constexpr (MODE != kGroupNone) sqrt_own = sqrt((double)op.orbit[row]);
      for (int seg = o_lo; seg < o_hi; seg += 64) {
        const int n = o_hi - seg < 64 ? o_hi - seg : 64;
        uint64_t bits = 0;
        for (int j = 0; j < n; ++j) {
          const uint64_t* m = T.mm + 2 * (seg + j);
          bits |= (uint64_t)((b & m[0]) == m[1]) << j;
        }
        while (bits) {
          const int j = seg + __builtin_ctzll(bits);
          bits &= bits - 1;
          const uint64_t img = b ^ T.fa[2 * j];
          uint64_t s;
          int h, cnt;
          Canon<MODE>::run(op.g, L, img, s, h, cnt);
          double sr, si;
          if (op.xm.values) {
            V xv;
            if (!xmap_fetch1(op.xm, codes, s, xv)) continue;
            load_c(&xv, 0, sr, si);
          } else {
            const int64_t idx = index_of(op.ix, W, s);
            if (idx < 0) continue;
            load_c(psi, idx, sr, si);
          }
          double wr = 1.0, wi = 0.0;
          if constexpr (MODE != kGroupNone) {
            const double sc = sqrt_own * L.sqcnt[cnt];
            wr = L.ph_re[h] * sc;
            if constexpr (CX) wi = L.ph_im[h] * sc;
          }
          const double qr = wr * sr - wi * si;  // w * psi_s
          const int t = n_diag + j - t0;
          if constexpr (CX) {
            const double qi = wr * si + wi * sr;
            atomicAdd(mine + t * 2, fr * qr + fi * qi);  // conj(phi_r) * q
            atomicAdd(mine + t * 2 + 1, fr * qi - fi * qr);
          } else {
            atomicAdd(mine + t, fr * qr);
          }
        }
      }
    }
  }
  __syncthreads();
  double* out = partial + (int64_t)blockIdx.x * n_terms * 2;
  for (int i = threadIdx.x; i < nw * 2; i += kBlock) {
    double v = 0.0;
    if (CX || (i & 1) == 0)
      for (int w = 0; w < kExpectWaves; ++w) v += bins[w * nw * NC + (CX ? i : (i >> 1))];
    out[2 * t0 + i] = v;
  }
}

// The off-diagonal terms of the term pass without a symmetry group (unit
// weights), with the lanes in lock-step over a window of KT terms
// (blockIdx.y): one lane per row (rows grid-stride over blockIdx.x), each
// term's sum held in the lane's registers; at the end the window's sums are
// reduced over the wave (shuffles) and over the block's waves in a fixed
// order into partial[block][t][2].  A lane whose row does not match a term
// idles through that term's steps, which costs less than the LDS atomics of
// k_term_expect, whose lanes add to the bins of different terms at once.
// Windows of 16 terms for real vectors, 8 for complex ones (4x6 XXZ pass:
// 0.44 ms, against 0.46 with 8 and 0.54 with 32, whose registers cost waves).
constexpr int kLockWindow = 16;
template <typename V, int KT>
__global__ __launch_bounds__(kBlock) void k_term_expect_lock(OpDev op, const V* __restrict__ phi,
                                                             const V* __restrict__ psi, double* __restrict__ partial) {
  constexpr bool CX = VT<V>::cx;
  extern __shared__ unsigned char lds[];
  __shared__ double sums[kExpectWaves][KT][2];
  lds_copy(lds + op.w_lds, op.ix.W, op.w_bytes);
  uint16_t* codes = (uint16_t*)(lds + op.codes_lds);
  if (op.xm.values) xmap_stage(op.xm, codes);
  __syncthreads();
  const uint64_t* W = (const uint64_t*)(lds + op.w_lds);
  const int j0 = (int)blockIdx.y * KT;  // first off-diagonal term of the window
  double ar[KT], ai[KT];
#pragma unroll
  for (int q = 0; q < KT; ++q) ar[q] = ai[q] = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x; row < op.dim; row += (int64_t)gridDim.x * kBlock) {
    const uint64_t b = op.rows[row];
    double fr, fi;
    load_c(phi, row, fr, fi);
    // every image's lookup and load first, then the sums
    double sr[KT], si[KT];
#pragma unroll
    for (int q = 0; q < KT; ++q) {
      sr[q] = si[q] = 0.0;
      const int j = j0 + q;
      if (j < op.n_off) {  // the same for every lane
        const uint64_t mask = op.off[j].mask, match = op.off[j].match;
        if ((b & mask) == match) {
          const uint64_t s = (b & ~mask) | op.off[j].image;
          if (op.xm.values) {
            V xv;
            if (xmap_fetch1(op.xm, codes, s, xv)) load_c(&xv, 0, sr[q], si[q]);
          } else {
            const int64_t idx = index_of(op.ix, W, s);
            if (idx >= 0) load_c(psi, idx, sr[q], si[q]);
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < KT; ++q) {
      ar[q] += fr * sr[q] + fi * si[q];  // conj(phi_r) psi_s
      if constexpr (CX) ai[q] += fr * si[q] - fi * sr[q];
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < KT; ++q) {
    double a = ar[q], c = ai[q];
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      if constexpr (CX) c += __shfl_xor(c, o, 64);
    }
    if (lane == 0) {
      sums[wave][q][0] = a;
      sums[wave][q][1] = CX ? c : 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * KT) {
    const int q = threadIdx.x >> 1, c = threadIdx.x & 1;
    if (j0 + q < op.n_off) {
      double v = 0.0;
      for (int w = 0; w < kExpectWaves; ++w) v += sums[w][q][c];
      partial[(int64_t)blockIdx.x * 2 * (op.n_diag + op.n_off) + 2 * (op.n_diag + j0 + q) + c] = v;
    }
  }
}

// Sum of the rows [64 c, 64 c + 64) of partial (n_parts x width) for 64
// columns per block: inter[c * width + col], rows added in order.
__global__ __launch_bounds__(64) void k_sum_parts_chunks(const double* __restrict__ partial, int64_t n_parts, int width,
                                                         double* __restrict__ inter) {
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= width) return;
  const int64_t p0 = (int64_t)blockIdx.y * 64;
  const int64_t p1 = p0 + 64 < n_parts ? p0 + 64 : n_parts;
  double s = 0.0;
  for (int64_t p = p0; p < p1; ++p) s += partial[p * width + col];
  inter[(int64_t)blockIdx.y * width + col] = s;
}

__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ partial, int64_t n_blocks, int width,
                                                            double* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= width) return;
  double s = 0.0;
  for (int64_t b = 0; b < n_blocks; ++b) s += partial[b * width + i];
  out[i] = s;
}

// Diagonal terms of the term pass: g_t = sum_r [b_r & mask_t == match_t]
// conj(phi_r) psi_r, one thread per term (tiles of kBlock terms).  The block
// stages a chunk of its rows (word and conj(phi_r) psi_r) in LDS and every
// thread walks the chunk in order, adding into registers: no atomics, since
// each term has its own lane (with LDS atomics, one add per matched term and
// row, the diagonal terms took half of the pass).  Rows are taken
// grid-stride in chunks of kBlock; partial[block][2 t] as the term pass.
template <typename V, typename W>
__global__ __launch_bounds__(kBlock) void k_diag_expect(OpDev op, const V* __restrict__ phi,
                                                        const V* __restrict__ psi, double* __restrict__ partial,
                                                        int width) {
  // W: the word type, uint32_t when every word and mask fits in 32 bits (two
  // rows per 8 bytes of LDS, a 32-bit test per row), else uint64_t
  constexpr bool CX = VT<V>::cx;
  constexpr int RW = 16 / (int)sizeof(W);  // rows per 16-byte LDS read of the words
  __shared__ __align__(16) W rb[kBlock];
  __shared__ __align__(16) double rv[kBlock];
  __shared__ __align__(16) double ri[kBlock];
  for (int tt = 0; tt < op.n_diag; tt += kBlock) {
    const int t = tt + (int)threadIdx.x;
    W m = 0, c = 1;  // no term: (b & 0) != 1, never matches
    if (t < op.n_diag) {
      m = (W)op.diag[t].mask;
      c = (W)op.diag[t].match;
    }
    double ar = 0.0, ai = 0.0;
    for (int64_t r0 = (int64_t)blockIdx.x * kBlock; r0 < op.dim; r0 += (int64_t)gridDim.x * kBlock) {
      const int64_t r = r0 + threadIdx.x;
      __syncthreads();  // the previous chunk is done with
      if (r < op.dim) {
        double fr, fi, pr, pi;
        load_c(phi, r, fr, fi);
        load_c(psi, r, pr, pi);
        rb[threadIdx.x] = (W)op.rows[r];
        rv[threadIdx.x] = fr * pr + fi * pi;  // conj(phi_r) psi_r
        ri[threadIdx.x] = fr * pi - fi * pr;
      }
      __syncthreads();
      const int nr = op.dim - r0 < kBlock ? (int)(op.dim - r0) : kBlock;
      int j = 0;
      for (; j + RW <= nr; j += RW) {  // RW rows per 16-byte read of the words
        W bb[RW];
        *(uint4*)bb = *(const uint4*)(rb + j);
#pragma unroll
        for (int q = 0; q < RW; q += 2) {
          const double2 vv = *(const double2*)(rv + j + q);
          ar += ((bb[q] & m) == c) ? vv.x : 0.0;
          ar += ((bb[q + 1] & m) == c) ? vv.y : 0.0;
          if constexpr (CX) {
            const double2 wi = *(const double2*)(ri + j + q);
            ai += ((bb[q] & m) == c) ? wi.x : 0.0;
            ai += ((bb[q + 1] & m) == c) ? wi.y : 0.0;
          }
        }
      }
      for (; j < nr; ++j) {
        if ((rb[j] & m) == c) {
          ar += rv[j];
          if constexpr (CX) ai += ri[j];
        }
      }
    }
    if (t < op.n_diag) {
      partial[(int64_t)blockIdx.x * width + 2 * t] = ar;
      partial[(int64_t)blockIdx.x * width + 2 * t + 1] = ai;
    }
  }
}

template <typename V, int MODE>
hipError_t launch_expect_t(const OpDev& op, const void* phi, const void* psi, double* d_partial, int64_t n_blocks,
                           double* d_out, hipStream_t st) {
  const int n_terms = op.n_diag + op.n_off;
  const int nw_max = term_expect_window(op, VT<V>::cx ? 2 : 1);
  if (nw_max <= 0) return hipErrorInvalidValue;
  // diagonal terms in their own kernel, then windows of the off-diagonal ones
  if (op.n_diag > 0) {
    if (op.word_bits <= 32)
      hipLaunchKernelGGL((k_diag_expect<V, uint32_t>), dim3((unsigned)n_blocks), dim3(kBlock), 0, st, op,
                         (const V*)phi, (const V*)psi, d_partial, n_terms * 2);
    else
      hipLaunchKernelGGL((k_diag_expect<V, uint64_t>), dim3((unsigned)n_blocks), dim3(kBlock), 0, st, op,
                         (const V*)phi, (const V*)psi, d_partial, n_terms * 2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // without a symmetry group (unit weights, no canonical form): the lanes take the terms in lock-step
  const bool own = MODE == kGroupNone;
  if (own && op.n_off > 0) {
    constexpr int KT = VT<V>::cx ? kLockWindow / 2 : kLockWindow;
    const dim3 grid((unsigned)n_blocks, (unsigned)((op.n_off + KT - 1) / KT));
    hipLaunchKernelGGL((k_term_expect_lock<V, KT>), grid, dim3(kBlock), expect_bins_offset(op), st, op,
                       (const V*)phi, (const V*)psi, d_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int t0 = op.n_diag; !own && t0 < n_terms; t0 += nw_max) {
    const int nw = n_terms - t0 < nw_max ? n_terms - t0 : nw_max;
    const uint32_t lds = expect_bins_offset(op) + (uint32_t)(kExpectWaves * nw * (VT<V>::cx ? 2 : 1) * 8);
    hipLaunchKernelGGL((k_term_expect<V, MODE>), dim3((unsigned)n_blocks), dim3(kBlock), lds, st, op, (const V*)phi,
                       (const V*)psi, t0, nw, d_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // fixed-order reduction of the per-block rows: chunks of 64 rows, then the chunks
  const int width = n_terms * 2;
  const int64_t n_chunks = (n_blocks + 63) / 64;
  double* inter = d_partial + n_blocks * width;
  hipLaunchKernelGGL(k_sum_parts_chunks, dim3((unsigned)((width + 63) / 64), (unsigned)n_chunks), dim3(64), 0, st,
                     d_partial, n_blocks, width, inter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((width + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, inter,
                     n_chunks, width, d_out);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_expect_v(const OpDev& op, const void* phi, const void* psi, double* d_partial, int64_t n_blocks,
                           double* d_out, hipStream_t st) {
  switch (op.g.mode) {
    case kGroupNone:
      return launch_expect_t<V, kGroupNone>(op, phi, psi, d_partial, n_blocks, d_out, st);
    case kGroupGeneric:
      return launch_expect_t<V, kGroupGeneric>(op, phi, psi, d_partial, n_blocks, d_out, st);
    default:
      return launch_expect_t<V, kGroupTorus>(op, phi, psi, d_partial, n_blocks, d_out, st);
  }
}

// ---------------------------------------------------------------------------
// Assembled rows (row form): the off-diagonal images of every row, found once
// by the matvec's own matching and canonicalisation, stored as 32-bit entries
//   col | cls << col_bits
// in the matvec's term order, rows one after another (rowptr).  Class 0 is an
// image outside the space (weight 0); class c > 0 weighs a * phase, the term's
// amplitude times the character phase of the element reaching the image.  The
// orbit factor sqrt(N_r / N_s) is not stored: a product reads x / sqrt(N_s)
// (prepared once per product) and scales each row's sum by sqrt(N_r).
// ---------------------------------------------------------------------------

// Bits of the terms seg .. seg + nseg - 1 (nseg <= 64) that match the word b,
// as the matvec tests them (two terms per mask test on the paired copy).
__device__ __forceinline__ uint64_t match_bits(const OpDev& op, const MvTerms& T, uint64_t b, int seg, int nseg) {
  uint64_t bits = 0;
  if (op.n_mv) {
    for (int j = 0; j < nseg; j += 2) {
      const uint64_t* m = T.mm + 3 * ((seg + j) >> 1);
      const uint64_t xm = b & m[0];
      bits |= (uint64_t)((uint32_t)(xm == m[1]) | ((uint32_t)(xm == m[2]) << 1)) << j;
    }
  } else {
    for (int j = 0; j < nseg; ++j) {
      const uint64_t* m = T.mm + 2 * (seg + j);
      bits |= (uint64_t)((b & m[0]) == m[1]) << j;
    }
  }
  return bits;
}

// counts[r] = number of terms matching row r (the entries it will hold)
__global__ __launch_bounds__(kMvBlock) void k_asm_count(OpDev op, uint32_t* __restrict__ counts) {
  extern __shared__ unsigned char lds[];
  const MvTerms T = stage_mv_terms(op, lds, op.n_mv != 0);
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * kMvBlock + threadIdx.x;
  if (row >= op.dim) return;
  const uint64_t b = op.rows[row];
  const int n_terms = op.n_mv ? op.n_mv : op.n_off;
  uint32_t n = 0;
  for (int seg = 0; seg < n_terms; seg += 64)
    n += (uint32_t)__builtin_popcountll(match_bits(op, T, b, seg, (n_terms - seg) < 64 ? (n_terms - seg) : 64));
  counts[row] = n;
}

// Entries of every row: acls[t] is the amplitude class (1-based) of matvec
// term t; with phases, cls = 1 + (acls - 1) * n_ph + pcls[h].
template <int MODE, bool UNIT>
__global__ __launch_bounds__(kMvBlock) void k_asm_fill(OpDev op, AsmFillDev a) {
  extern __shared__ unsigned char lds[];
  GroupLds L{};
  if constexpr (MODE != kGroupNone) L = stage_group(op.g, op.lay, lds);
  const MvTerms T = stage_mv_terms(op, lds, op.n_mv != 0);
  lds_copy(lds + op.w_lds, op.ix.W, op.w_bytes);
  const uint64_t* W = (const uint64_t*)(lds + op.w_lds);
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * kMvBlock + threadIdx.x;
  if (row >= op.dim) return;
  const uint64_t b = op.rows[row];
  uint32_t* out = a.ent + a.rowptr[row];
  const int n_terms = op.n_mv ? op.n_mv : op.n_off;
  for (int seg = 0; seg < n_terms; seg += 64) {
    uint64_t bits = match_bits(op, T, b, seg, (n_terms - seg) < 64 ? (n_terms - seg) : 64);
    while (bits) {
      const int t = seg + __builtin_ctzll(bits);
      bits &= bits - 1;
      uint64_t s;
      int h, cnt;
      Canon<MODE>::template run<!UNIT>(op.g, L, b ^ T.fa[2 * t], s, h, cnt);
      const int64_t idx = index_of(op.ix, W, s);
      uint32_t e = 0;
      if (idx >= 0) {
        uint32_t cls = a.acls[t];
        if (!UNIT && a.pcls) cls = 1u + (cls - 1u) * (uint32_t)a.n_ph + a.pcls[h];
        e = (uint32_t)idx | (cls << a.col_bits);
      }
      *out++ = e;
    }
  }
}

// xt[i] = x[i] / sqrt(N_i)
template <typename V>
__global__ __launch_bounds__(kBlock) void k_asm_scale(int64_t dim, const uint16_t* __restrict__ orbit,
                                                      const V* __restrict__ x, V* __restrict__ xt) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= dim) return;
  using R = typename VT<V>::R;
  const double s = 1.0 / sqrt((double)orbit[i]);
  V v = x[i];
  if constexpr (VT<V>::cx) {
    v.re = (R)((double)v.re * s);
    v.im = (R)((double)v.im * s);
  } else {
    v = (R)((double)v * s);
  }
  xt[i] = v;
}

template <typename V>
__device__ __forceinline__ V vshfl_xor(const V& v, int o, int width) {
  if constexpr (VT<V>::cx) {
    return V{__shfl_xor(v.re, o, width), __shfl_xor(v.im, o, width)};
  } else {
    return __shfl_xor(v, o, width);
  }
}

// y[r] = (accumulate ? y[r] : 0) + d_r x[r] + sqrt(N_r) sum_e W[cls_e] xt[col_e]:
// GL lanes per row stride over its entries (a wave reads whole runs of
// neighbouring rows' entries) in batches of U entries per lane, every load of
// a batch issued before any is used (entries past the row's end are not
// loaded and weigh nothing), then a butterfly sum over the GL lanes.  Blocks
// are numbered over a 2D grid (a grid dimension holds fewer than 2^32 lanes).
template <typename V, int GL, int U>
__global__ __launch_bounds__(kBlock) void k_asm_spmv(OpDev op, AsmDev a, const V* __restrict__ x,
                                                     const V* __restrict__ xt, V* __restrict__ y, int accumulate) {
  using R = typename VT<V>::R;
  __shared__ R w_re[kAsmMaxClasses], w_im[kAsmMaxClasses];
  for (int i = threadIdx.x; i < a.n_cls; i += kBlock) {
    w_re[i] = (R)a.w_re[i];
    w_im[i] = (R)a.w_im[i];
  }
  __syncthreads();
  const int64_t blk = xcd_chunked_block((int64_t)blockIdx.y * gridDim.x + blockIdx.x,
                                        (int64_t)gridDim.x * gridDim.y, op.xcd_chunk);
  const int64_t row = (blk * kBlock + threadIdx.x) / GL;
  const int sub = (int)(threadIdx.x & (GL - 1));
  const uint32_t cmask = (a.col_bits >= 32) ? ~0u : ((1u << a.col_bits) - 1u);
  V acc = vzero<V>();
  if (row < op.dim) {
    const uint64_t end = a.rowptr[row + 1];
    for (uint64_t e0 = a.rowptr[row] + (uint64_t)sub; e0 < end; e0 += (uint64_t)GL * U) {
      uint32_t v[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint64_t e = e0 + (uint64_t)k * GL;
        v[k] = e < end ? __builtin_nontemporal_load(a.ent + e) : 0u;
      }
      V xv[U];
#pragma unroll
      for (int k = 0; k < U; ++k) xv[k] = xt[v[k] & cmask];
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (e0 + (uint64_t)k * GL < end) vfma(acc, w_re[v[k] >> a.col_bits], w_im[v[k] >> a.col_bits], xv[k]);
    }
  }
#pragma unroll
  for (int o = GL / 2; o > 0; o >>= 1) acc = vadd(acc, vshfl_xor(acc, o, GL));
  if (row >= op.dim || sub != 0) return;
  double dre = 0.0, dim_ = 0.0;
  if (op.diag_re) {
    dre = op.diag_re[row];
    if (op.diag_im) dim_ = op.diag_im[row];
  } else {
    diag_sum(op, op.rows[row], dre, dim_);
  }
  V out = vzero<V>();
  vfma(out, (R)dre, (R)dim_, x[row]);
  const R so = op.orbit ? (R)sqrt((double)op.orbit[row]) : (R)1;
  vfma(out, so, (R)0, acc);
  y[row] = accumulate ? vadd(y[row], out) : out;
}

// Number of off-diagonal images generated over all rows (matched terms).
__global__ __launch_bounds__(kBlock) void k_count_images(OpDev op, unsigned long long* __restrict__ total) {
  extern __shared__ unsigned char lds[];
  const OffTerm* off = (const OffTerm*)(lds + op.off_lds);
  lds_copy(lds + op.off_lds, op.off, (uint32_t)(op.n_off * sizeof(OffTerm)));
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long c = 0;
  if (row < op.dim) {
    const uint64_t b = op.rows[row];
    for (int t = 0; t < op.n_off; ++t) c += ((b & off[t].mask) == off[t].match) ? 1ull : 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}

template <int MODE>
hipError_t launch_asm_fill_t(const OpDev& op, const AsmFillDev& a, hipStream_t st) {
  dim3 grid((unsigned)((op.dim + kMvBlock - 1) / kMvBlock));
  if (MODE == kGroupNone || op.g.unit_phases)
    hipLaunchKernelGGL((k_asm_fill<MODE, true>), grid, dim3(kMvBlock), op.lds_total, st, op, a);
  else
    hipLaunchKernelGGL((k_asm_fill<MODE, false>), grid, dim3(kMvBlock), op.lds_total, st, op, a);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_asm_spmv_v(const OpDev& op, const AsmDev& a, int lanes, int batch, int accumulate, const void* x,
                             const void* xt, void* y, hipStream_t st) {
  const uint64_t blocks = ((uint64_t)op.dim * (uint64_t)lanes + kBlock - 1) / kBlock;
  const uint64_t gx = blocks < 65536 ? blocks : 65536;
  dim3 grid((unsigned)gx, (unsigned)((blocks + gx - 1) / gx));
#define EDHIP_ASM(G, B)                                                                                          \
  hipLaunchKernelGGL((k_asm_spmv<V, G, B>), grid, dim3(kBlock), 0, st, op, a, (const V*)x, (const V*)xt, (V*)y, \
                     accumulate)
#define EDHIP_ASM_B(G)                 \
  switch (batch) {                     \
    case 4: EDHIP_ASM(G, 4); break;    \
    case 8: EDHIP_ASM(G, 8); break;    \
    case 16: EDHIP_ASM(G, 16); break;  \
    default: return hipErrorInvalidValue; \
  }
  switch (lanes) {
    case 4: EDHIP_ASM_B(4); break;
    case 8: EDHIP_ASM_B(8); break;
    case 16: EDHIP_ASM_B(16); break;
    case 32: EDHIP_ASM_B(32); break;
    default: return hipErrorInvalidValue;
  }
#undef EDHIP_ASM_B
#undef EDHIP_ASM
  return hipGetLastError();
}

}  // namespace

hipError_t launch_asm_count(const OpDev& op, uint32_t* d_counts, hipStream_t st) {
  if (op.dim == 0) return hipSuccess;
  hipLaunchKernelGGL(k_asm_count, dim3((unsigned)((op.dim + kMvBlock - 1) / kMvBlock)), dim3(kMvBlock), op.lds_total,
                     st, op, d_counts);
  return hipGetLastError();
}

hipError_t launch_asm_fill(const OpDev& op, const AsmFillDev& a, hipStream_t st) {
  if (op.dim == 0) return hipSuccess;
  switch (op.g.mode) {
    case kGroupNone:
      return launch_asm_fill_t<kGroupNone>(op, a, st);
    case kGroupGeneric:
      return launch_asm_fill_t<kGroupGeneric>(op, a, st);
    default:
      if (op.g.pair_table) {
        const int Lx = op.g.Lx, Ly = op.g.Ly;
        if (op.g.n_cosets == 1) {
          if (Lx == 6 && Ly == 6) return launch_asm_fill_t<kTransPairs6x6>(op, a, st);
          if (Lx == 4 && Ly == 6) return launch_asm_fill_t<kTransPairs4x6>(op, a, st);
          if (Lx == 6 && Ly == 4) return launch_asm_fill_t<kTransPairs6x4>(op, a, st);
          if (Lx == 4 && Ly == 4) return launch_asm_fill_t<kTransPairs4x4>(op, a, st);
        }
        if (Lx == 6 && Ly == 6) return launch_asm_fill_t<kTorusPairs6x6>(op, a, st);
        if (Lx == 4 && Ly == 6) return launch_asm_fill_t<kTorusPairs4x6>(op, a, st);
        if (Lx == 6 && Ly == 4) return launch_asm_fill_t<kTorusPairs6x4>(op, a, st);
        if (Lx == 4 && Ly == 4) return launch_asm_fill_t<kTorusPairs4x4>(op, a, st);
        return launch_asm_fill_t<kGroupTorusPairs>(op, a, st);
      }
      return launch_asm_fill_t<kGroupTorus>(op, a, st);
  }
}

hipError_t launch_asm_scale(int prec, int64_t dim, const uint16_t* d_orbit, const void* x, void* xt, hipStream_t st) {
  if (dim == 0) return hipSuccess;
  dim3 grid((unsigned)((dim + kBlock - 1) / kBlock));
  switch (prec) {
    case kF64:
      hipLaunchKernelGGL(k_asm_scale<double>, grid, dim3(kBlock), 0, st, dim, d_orbit, (const double*)x, (double*)xt);
      break;
    case kC128:
      hipLaunchKernelGGL(k_asm_scale<Cx<double>>, grid, dim3(kBlock), 0, st, dim, d_orbit, (const Cx<double>*)x,
                         (Cx<double>*)xt);
      break;
    case kF32:
      hipLaunchKernelGGL(k_asm_scale<float>, grid, dim3(kBlock), 0, st, dim, d_orbit, (const float*)x, (float*)xt);
      break;
    case kC64:
      hipLaunchKernelGGL(k_asm_scale<Cx<float>>, grid, dim3(kBlock), 0, st, dim, d_orbit, (const Cx<float>*)x,
                         (Cx<float>*)xt);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_asm_spmv(const OpDev& op, const AsmDev& a, int prec, int lanes, int batch, int accumulate,
                           const void* x, const void* xt, void* y, hipStream_t st) {
  if (op.dim == 0) return hipSuccess;
  if (a.n_cls > kAsmMaxClasses) return hipErrorInvalidValue;
  switch (prec) {
    case kF64:
      return launch_asm_spmv_v<double>(op, a, lanes, batch, accumulate, x, xt, y, st);
    case kC128:
      return launch_asm_spmv_v<Cx<double>>(op, a, lanes, batch, accumulate, x, xt, y, st);
    case kF32:
      return launch_asm_spmv_v<float>(op, a, lanes, batch, accumulate, x, xt, y, st);
    case kC64:
      return launch_asm_spmv_v<Cx<float>>(op, a, lanes, batch, accumulate, x, xt, y, st);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_count_images(const OpDev& op, unsigned long long* d_total, hipStream_t st) {
  if (op.dim == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_images, dim3((unsigned)((op.dim + kBlock - 1) / kBlock)), dim3(kBlock), op.lds_total, st, op,
                     d_total);
  return hipGetLastError();
}

hipError_t launch_diag_values(const OpDev& op, double* d_re, double* d_im, hipStream_t st) {
  if (op.dim == 0) return hipSuccess;
  hipLaunchKernelGGL(k_diag_values, dim3((unsigned)((op.dim + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, op, d_re,
                     d_im);
  return hipGetLastError();
}

hipError_t launch_matvec(const OpDev& op, int prec, int transpose, int accumulate, const void* x, void* y,
                         hipStream_t st) {
  if (op.dim <= op.row_begin) return hipSuccess;
  switch (prec) {
    case kF64:
      return launch_mv_v<double>(op, transpose, accumulate, x, y, st);
    case kC128:
      return launch_mv_v<Cx<double>>(op, transpose, accumulate, x, y, st);
    case kF32:
      return launch_mv_v<float>(op, transpose, accumulate, x, y, st);
    case kC64:
      return launch_mv_v<Cx<float>>(op, transpose, accumulate, x, y, st);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_iterate(const OpDev& op, const OffTerm* all_terms, int n_all, int transpose, int64_t n_rows,
                          const int64_t* d_rows, int64_t* d_idx, double* d_re, double* d_im, hipStream_t st) {
  if (n_rows == 0) return hipSuccess;
  dim3 grid((unsigned)((n_rows + kBlock - 1) / kBlock));
#define EDHIP_IT(M, T)                                                                                             \
  hipLaunchKernelGGL((k_iterate<M, T>), grid, dim3(kBlock), op.lds_total, st, op, all_terms, n_all, n_rows, d_rows, \
                     d_idx, d_re, d_im)
  if (op.g.mode == kGroupNone) {
    if (transpose) EDHIP_IT(kGroupNone, true); else EDHIP_IT(kGroupNone, false);
  } else if (op.g.mode == kGroupGeneric) {
    if (transpose) EDHIP_IT(kGroupGeneric, true); else EDHIP_IT(kGroupGeneric, false);
  } else {
    if (transpose) EDHIP_IT(kGroupTorus, true); else EDHIP_IT(kGroupTorus, false);
  }
#undef EDHIP_IT
  return hipGetLastError();
}

hipError_t launch_sparse_columns(const OpDev& op, int transpose, double tol, int slot, int64_t n_sel,
                                 const int64_t* d_sel, int64_t* d_scratch_idx, double* d_scratch_re,
                                 double* d_scratch_im, int64_t* d_count, hipStream_t st) {
  if (n_sel == 0) return hipSuccess;
  dim3 grid((unsigned)((n_sel + kBlock - 1) / kBlock));
#define EDHIP_SC(M, T)                                                                                                 \
  hipLaunchKernelGGL((k_sparse_columns<M, T>), grid, dim3(kBlock), op.lds_total, st, op, tol, slot, n_sel, d_sel,     \
                     d_scratch_idx, d_scratch_re, d_scratch_im, d_count)
  if (op.g.mode == kGroupNone) {
    if (transpose) EDHIP_SC(kGroupNone, true); else EDHIP_SC(kGroupNone, false);
  } else if (op.g.mode == kGroupGeneric) {
    if (transpose) EDHIP_SC(kGroupGeneric, true); else EDHIP_SC(kGroupGeneric, false);
  } else {
    if (transpose) EDHIP_SC(kGroupTorus, true); else EDHIP_SC(kGroupTorus, false);
  }
#undef EDHIP_SC
  return hipGetLastError();
}

hipError_t launch_dense_columns(const OpDev& op, int complex_out, double* d_out, hipStream_t st) {
  if (op.dim == 0) return hipSuccess;
  dim3 grid((unsigned)((op.dim + kBlock - 1) / kBlock));
#define EDHIP_DC(M, C) hipLaunchKernelGGL((k_dense_columns<M, C>), grid, dim3(kBlock), op.lds_total, st, op, d_out)
  if (op.g.mode == kGroupNone) {
    if (complex_out) EDHIP_DC(kGroupNone, true); else EDHIP_DC(kGroupNone, false);
  } else if (op.g.mode == kGroupGeneric) {
    if (complex_out) EDHIP_DC(kGroupGeneric, true); else EDHIP_DC(kGroupGeneric, false);
  } else {
    if (complex_out) EDHIP_DC(kGroupTorus, true); else EDHIP_DC(kGroupTorus, false);
  }
#undef EDHIP_DC
  return hipGetLastError();
}

hipError_t launch_sparse_compact(int64_t dim, int slot, const int64_t* d_colptr, const int64_t* d_scratch_idx,
                                 const double* d_scratch_re, const double* d_scratch_im, int64_t* d_rowval,
                                 double* d_re, double* d_im, hipStream_t st) {
  if (dim == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sparse_compact, dim3((unsigned)((dim + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, dim, slot,
                     d_colptr, d_scratch_idx, d_scratch_re, d_scratch_im, d_rowval, d_re, d_im);
  return hipGetLastError();
}

// terms per pass: the per-wave bins (nc doubles per term) fill the LDS left
// over by the operator and symmetry tables (64 KiB per block)
int term_expect_window(const OpDev& op, int nc) {
  const int64_t room = 64 * 1024 - (int64_t)expect_bins_offset(op);
  const int64_t per_term = (int64_t)kExpectWaves * nc * 8;
  const int64_t nw = room / per_term;
  const int n_terms = op.n_diag + op.n_off;
  if (nw < 1) return 0;
  return nw < n_terms ? (int)nw : n_terms;
}

int64_t term_expect_blocks(const OpDev& op) {
  int64_t nb = (op.dim + kBlock - 1) / kBlock;
  if (nb > kExpectBlocksMax) nb = kExpectBlocksMax;
  if (nb < 1) nb = 1;
  return nb;
}

int64_t term_expect_scratch_rows(const OpDev& op) {
  const int64_t nb = term_expect_blocks(op);
  return nb + (nb + 63) / 64;
}

// The allocation-free term pass's scratch: (blocks + chunks) rows of partial
// sums of 2 kNoAllocMaxTerms doubles, then the 2 kNoAllocMaxTerms results.
// Every entry a call reads was stored by the same call (the per-block passes
// store each (block, term) entry once; the reductions store theirs), so no
// clearing is needed between calls.
constexpr size_t kNoAllocRows = (size_t)kExpectBlocksMax + ((size_t)kExpectBlocksMax + 63) / 64;
constexpr size_t kNoAllocDoubles = kNoAllocRows * 2 * kNoAllocMaxTerms + 2 * kNoAllocMaxTerms;
__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocMaxDevices = 64;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

hipError_t term_expect_noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  // the device address of this device's copy of the array (the name itself is
  // a host-side stand-in and never reaches a kernel)
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = (double*)p;
  return hipSuccess;
}

double* term_expect_noalloc_scratch(int device) {
  return (device >= 0 && device < kNoAllocMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

size_t term_expect_noalloc_doubles() { return kNoAllocDoubles; }

hipError_t launch_term_expect(const OpDev& op, int prec, const void* phi, const void* psi, double* d_partial,
                              int64_t n_blocks, double* d_out, hipStream_t st) {
  switch (prec) {
    case kF64:
      return launch_expect_v<double>(op, phi, psi, d_partial, n_blocks, d_out, st);
    case kC128:
      return launch_expect_v<Cx<double>>(op, phi, psi, d_partial, n_blocks, d_out, st);
    case kF32:
      return launch_expect_v<float>(op, phi, psi, d_partial, n_blocks, d_out, st);
    case kC64:
      return launch_expect_v<Cx<float>>(op, phi, psi, d_partial, n_blocks, d_out, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace edhip

// ============================================================================
// Fast path: bond-pair groups, register torus canonicaliser, hash lookup with
// the input vector embedded in the table slots.
// ============================================================================
namespace edhip {

namespace {

constexpr int kLyMax = 8;

template <typename V>
__device__ __forceinline__ bool hash_fetch(const HashDev& H, uint64_t s, V& xv) {
  uint32_t h = hash_word(s) & H.mask;
  for (;;) {
    const unsigned char* sp = H.slots + (size_t)h * (size_t)H.slot_bytes;
    const uint4 a = *(const uint4*)sp;
    const uint64_t k = (uint64_t)a.x | ((uint64_t)a.y << 32);
    if (k == s) {
      // element at byte 8 of a 16-byte slot, or byte 16 of a 32-byte slot
      const uint4 c = (H.xoff == 16) ? *(const uint4*)(sp + 16) : make_uint4(a.z, a.w, 0u, 0u);
      if constexpr (VT<V>::cx && sizeof(typename VT<V>::R) == 8) {
        xv.re = __hiloint2double((int)c.y, (int)c.x);
        xv.im = __hiloint2double((int)c.w, (int)c.z);
      } else if constexpr (VT<V>::cx) {
        xv.re = __uint_as_float(c.x);
        xv.im = __uint_as_float(c.y);
      } else if constexpr (sizeof(V) == 8) {
        xv = __hiloint2double((int)c.y, (int)c.x);
      } else {
        xv = __uint_as_float(c.x);
      }
      return true;
    }
    if (k == kHashEmpty) return false;
    h = (h + 1) & H.mask;
  }
}

template <typename V, int MODE, bool HASH, bool TRANS>
__global__ __launch_bounds__(kBlock) void k_mv_pairs(FastOpDev F, const V* __restrict__ x, V* __restrict__ y,
                                                     int accumulate) {
  extern __shared__ unsigned char lds[];
  const OpDev& op = F.base;
  GroupLds L;
  if constexpr (MODE != kGroupNone) L = stage_group(op.g, op.lay, lds);
  const int ng = F.pg.n_groups;
  uint32_t* meta = (uint32_t*)(lds + F.pg_lds);
  double* dre = (double*)(lds + F.pg_lds + (uint32_t)(kMaxPairGroups * 4));
  double* dim_ = dre + 4 * ng;
  double* ore = dim_ + 4 * ng;
  double* oim = ore + 4 * ng;
  lds_copy((unsigned char*)meta, F.pg.meta, (uint32_t)(ng * 4));
  lds_copy((unsigned char*)dre, F.pg.dre, (uint32_t)(ng * 32));
  lds_copy((unsigned char*)dim_, F.pg.dim, (uint32_t)(ng * 32));
  lds_copy((unsigned char*)ore, F.pg.ore, (uint32_t)(ng * 32));
  lds_copy((unsigned char*)oim, F.pg.oim, (uint32_t)(ng * 32));
  lds_copy(lds + op.w_lds, op.ix.W, op.w_bytes);
  const uint64_t* Wl = (const uint64_t*)(lds + op.w_lds);
  __syncthreads();
  const int64_t row = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (row >= op.dim) return;
  using R = typename VT<V>::R;
  const uint64_t b = op.rows[row];

  // one pass over the groups: summed diagonal and the matched groups
  double dr = F.pg.dconst_re, di = F.pg.dconst_im;
  uint64_t bw[kMaxPairGroups / 64];
#pragma unroll
  for (int w = 0; w < kMaxPairGroups / 64; ++w) {
    uint64_t bits = 0;
    const int g0 = w * 64;
    const int gn = (ng - g0) < 64 ? (ng - g0) : 64;
    for (int j = 0; j < gn; ++j) {
      const uint32_t m = meta[g0 + j];
      const uint32_t code = (uint32_t)((b >> (m & 63u)) & 1ull) | ((uint32_t)((b >> ((m >> 8) & 63u)) & 1ull) << 1);
      dr += dre[(g0 + j) * 4 + code];
      di += dim_[(g0 + j) * 4 + code];
      bits |= (uint64_t)((m >> (16 + code)) & 1u) << j;
    }
    bw[w] = bits;
  }
  V acc = vzero<V>();
  vfma(acc, (R)dr, (R)di, x[row]);
  double sqrt_own = 1.0;
  if constexpr (MODE != kGroupNone) sqrt_own = sqrt((double)op.orbit[row]);

  for (;;) {
    int g;
    if (bw[0]) {
      g = __builtin_ctzll(bw[0]);
      bw[0] &= bw[0] - 1;
    } else if (bw[1]) {
      g = 64 + __builtin_ctzll(bw[1]);
      bw[1] &= bw[1] - 1;
    } else if (bw[2]) {
      g = 128 + __builtin_ctzll(bw[2]);
      bw[2] &= bw[2] - 1;
    } else if (bw[3]) {
      g = 192 + __builtin_ctzll(bw[3]);
      bw[3] &= bw[3] - 1;
    } else {
      break;
    }
    const uint32_t m = meta[g];
    const uint32_t p = m & 63u, q = (m >> 8) & 63u;
    const uint32_t code = (uint32_t)((b >> p) & 1ull) | ((uint32_t)((b >> q) & 1ull) << 1);
    const uint32_t d = code ^ ((m >> (24 + 2 * code)) & 3u);
    const uint64_t img = b ^ (((uint64_t)(d & 1u) << p) | ((uint64_t)(d >> 1) << q));
    uint64_t s;
    int h, cnt;
    if constexpr (MODE == kGroupTorus) {
      if (op.g.pair_table)
        canonicalize_torus_pairs(op.g, L, img, s, h, cnt);
      else
        canonicalize_torus_reg<kLyMax>(op.g, L, img, s, h, cnt);
    } else {
      s = img;
      h = 0;
      cnt = 1;
    }
    V xv;
    bool found;
    if constexpr (HASH) {
      found = hash_fetch(F.hash, s, xv);
    } else {
      const int64_t idx = index_of(op.ix, Wl, s);
      found = idx >= 0;
      if (found) xv = x[idx];
    }
    if (found) {
      const double ar = ore[g * 4 + code], ai = oim[g * 4 + code];
      double wr = ar, wi = ai;
      if constexpr (MODE != kGroupNone) {
        const double pr = L.ph_re[h];
        const double pi = TRANS ? -L.ph_im[h] : L.ph_im[h];
        const double sc = sqrt_own * L.sqcnt[cnt];
        wr = (ar * pr - ai * pi) * sc;
        wi = (ar * pi + ai * pr) * sc;
      }
      vfma(acc, (R)wr, (R)wi, xv);
    }
  }
  y[row] = accumulate ? vadd(y[row], acc) : acc;
}

template <typename V, int MODE, bool HASH>
hipError_t launch_pairs_t(const FastOpDev& op, int transpose, int accumulate, const void* x, void* y, hipStream_t st) {
  dim3 grid((unsigned)((op.base.dim + kBlock - 1) / kBlock));
  if (transpose)
    hipLaunchKernelGGL((k_mv_pairs<V, MODE, HASH, true>), grid, dim3(kBlock), op.lds_total, st, op, (const V*)x, (V*)y,
                       accumulate);
  else
    hipLaunchKernelGGL((k_mv_pairs<V, MODE, HASH, false>), grid, dim3(kBlock), op.lds_total, st, op, (const V*)x,
                       (V*)y, accumulate);
  return hipGetLastError();
}

template <typename V>
hipError_t launch_pairs_v(const FastOpDev& op, int transpose, int accumulate, const void* x, void* y, hipStream_t st) {
  const bool hash = op.hash.slots != nullptr;
  if (op.base.g.mode == kGroupTorus)
    return hash ? launch_pairs_t<V, kGroupTorus, true>(op, transpose, accumulate, x, y, st)
                : launch_pairs_t<V, kGroupTorus, false>(op, transpose, accumulate, x, y, st);
  return hash ? launch_pairs_t<V, kGroupNone, true>(op, transpose, accumulate, x, y, st)
              : launch_pairs_t<V, kGroupNone, false>(op, transpose, accumulate, x, y, st);
}

__global__ __launch_bounds__(kBlock) void k_hash_build(const uint64_t* __restrict__ keys, int64_t dim, HashDev H,
                                                       uint32_t* __restrict__ slot_of) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= dim) return;
  const uint64_t s = keys[i];
  uint32_t h = hash_word(s) & H.mask;
  for (;;) {
    unsigned long long* kp = (unsigned long long*)(H.slots + (size_t)h * (size_t)H.slot_bytes);
    const unsigned long long prev = atomicCAS(kp, (unsigned long long)kHashEmpty, (unsigned long long)s);
    if (prev == (unsigned long long)kHashEmpty || prev == (unsigned long long)s) {
      slot_of[i] = h;
      return;
    }
    h = (h + 1) & H.mask;
  }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_hash_scatter(int64_t dim, const V* __restrict__ x,
                                                         const uint32_t* __restrict__ slot_of, HashDev H) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < dim; i += (int64_t)gridDim.x * kBlock) {
    *(V*)(H.slots + (size_t)slot_of[i] * (size_t)H.slot_bytes + H.xoff) = x[i];
  }
}

}  // namespace

hipError_t launch_matvec_pairs(const FastOpDev& op, int prec, int transpose, int accumulate, const void* x, void* y,
                               hipStream_t st) {
  if (op.base.dim == 0) return hipSuccess;
  switch (prec) {
    case kF64:
      return launch_pairs_v<double>(op, transpose, accumulate, x, y, st);
    case kC128:
      return launch_pairs_v<Cx<double>>(op, transpose, accumulate, x, y, st);
    case kF32:
      return launch_pairs_v<float>(op, transpose, accumulate, x, y, st);
    case kC64:
      return launch_pairs_v<Cx<float>>(op, transpose, accumulate, x, y, st);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_hash_build(const uint64_t* d_keys, int64_t dim, const HashDev& h, uint32_t* d_slot_of,
                             hipStream_t st) {
  if (dim == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hash_build, dim3((unsigned)((dim + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, d_keys, dim, h,
                     d_slot_of);
  return hipGetLastError();
}

hipError_t launch_hash_scatter(int prec, int64_t dim, const void* x, const uint32_t* d_slot_of, const HashDev& h,
                               hipStream_t st) {
  if (dim == 0) return hipSuccess;
  int64_t nb = (dim + kBlock - 1) / kBlock;
  if (nb > 16384) nb = 16384;
  switch (prec) {
    case kF64:
      hipLaunchKernelGGL(k_hash_scatter<double>, dim3((unsigned)nb), dim3(kBlock), 0, st, dim, (const double*)x, d_slot_of, h);
      break;
    case kC128:
      hipLaunchKernelGGL(k_hash_scatter<Cx<double>>, dim3((unsigned)nb), dim3(kBlock), 0, st, dim, (const Cx<double>*)x,
                         d_slot_of, h);
      break;
    case kF32:
      hipLaunchKernelGGL(k_hash_scatter<float>, dim3((unsigned)nb), dim3(kBlock), 0, st, dim, (const float*)x, d_slot_of, h);
      break;
    case kC64:
      hipLaunchKernelGGL(k_hash_scatter<Cx<float>>, dim3((unsigned)nb), dim3(kBlock), 0, st, dim, (const Cx<float>*)x,
                         d_slot_of, h);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace edhip
