/*
 * CPU oracle: a plain-C restatement of the reference algorithms on the
 * symmetry-reduced Hamiltonian-application path of ExactDiagonalization.jl
 * (v0.14.3).  TEST INFRASTRUCTURE ONLY: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker or the timed CPU baseline, never as part of the product path.
 *
 * Every function follows one reference function literally (loop order,
 * last-writer-wins dictionary semantics, chop rule) and cites it.
 * Scalars are complex double throughout (the reference's ComplexF64 route);
 * real inputs are passed with zero imaginary parts.
 *
 * Indices are 0-based here; -1 marks "not found / not in the reduced space"
 * exactly where the reference uses -1 (or 0 from fs_index).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ---------------------------------------------------------------------- */
/* FrozenSortedArrayIndex: fs_index / get(arr, key, -1)                    */
/* src/frozensortedarray.jl:29-32 (searchsortedfirst then equality test)   */
/* and :45-48 (default on miss).                                           */
/* ---------------------------------------------------------------------- */
int64_t orc_lookup(const uint64_t* keys, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n; /* searchsortedfirst: first i with keys[i] >= key */
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && keys[lo] == key) ? lo : -1;
}

void orc_lookup_many(const uint64_t* keys, int64_t n, int64_t m, const uint64_t* q, int64_t* out) {
  for (int64_t i = 0; i < m; ++i) out[i] = orc_lookup(keys, n, q[i]);
}

/* Validation of FrozenSortedArrayIndex{K}(keys): sorted and unique.
 * src/frozensortedarray.jl:14-22.  Returns 0 if valid, 1 if unsorted,
 * 2 if it contains duplicates. */
int orc_validate_sorted_unique(const uint64_t* keys, int64_t n) {
  for (int64_t i = 1; i < n; ++i) {
    if (keys[i - 1] > keys[i]) return 1;
  }
  for (int64_t i = 1; i < n; ++i) {
    if (keys[i - 1] == keys[i]) return 2;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* make_bitmask(bitwidth): src/util.jl:49-60                               */
/* ---------------------------------------------------------------------- */
static uint64_t make_bitmask(int w) { return (w >= 64) ? ~0ull : ((1ull << w) - 1ull); }

/* ---------------------------------------------------------------------- */
/* symmetry_apply(hs, ::SitePermutation, bitrep): src/Symmetry/            */
/* symmetry_apply.jl:82-92.  map[i] is the 0-based image site of site i.   */
/* ---------------------------------------------------------------------- */
uint64_t orc_symmetry_apply(int n_sites, const int32_t* bitoffsets, const int32_t* bitwidths,
                            const int32_t* map, uint64_t b) {
  uint64_t out = 0;
  for (int i = 0; i < n_sites; ++i) {
    out |= ((b >> bitoffsets[i]) & make_bitmask(bitwidths[i])) << bitoffsets[map[i]];
  }
  return out;
}

/* ---------------------------------------------------------------------- */
/* symmetry_reduce_serial(hsr, symops_and_amplitudes; tol):                */
/* src/Symmetry/symmetry_reduce_generic.jl:22-106.                          */
/*                                                                          */
/* maps: n_group x n_sites (0-based images), identity first.               */
/* flips (orc_symmetry_reduce_x, may be NULL): per element, a word XORed     */
/* into the permuted word: GlobalBitFlip(true) composed with the permutation */
/* (bitflipsymmetry.jl:26-38, DirectProductOperation symmetry_apply.jl:65-76;*/
/* a global flip commutes with every site permutation).                     */
/* chi:  the amplitudes y of (symop, y); the stored phase is conj(y*sgn)    */
/*       (:67) with sgn = 1 for site permutations (symmetry_apply.jl:91).   */
/* Outputs: reduced basis (ascending), mapping index (parent length, -1 if  */
/* excluded) and mapping amplitude (parent length, 0 if excluded).          */
/* Returns the reduced dimension, -1 if some |y| != 1 (ArgumentError, :27-29)*/
/* or -2 if an image falls outside the parent basis (KeyError at :83).      */
/* ---------------------------------------------------------------------- */
int64_t orc_symmetry_reduce_x(int64_t n_basis, const uint64_t* basis, int n_sites,
                              const int32_t* bitoffsets, const int32_t* bitwidths, int n_group,
                              const int32_t* maps, const uint64_t* flips, const double* chi_re,
                              const double* chi_im, double tol, uint64_t* reduced_out,
                              int64_t* mapping_index, double* amp_re, double* amp_im) {
  const double rtol_default = sqrt(2.220446049250313e-16);
  for (int g = 0; g < n_group; ++g) {
    double a = cabs(chi_re[g] + I * chi_im[g]);
    /* isapprox(abs(y), one(abs(y))) with the default rtol */
    if (!(fabs(a - 1.0) <= rtol_default * fmax(a, 1.0))) return -1;
  }
  int64_t* rep_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_basis > 0 ? n_basis : 1));
  unsigned char* visited = (unsigned char*)calloc((size_t)(n_basis > 0 ? n_basis : 1), 1);
  uint64_t* states = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n_group);
  cplx* phases = (cplx*)malloc(sizeof(cplx) * (size_t)n_group);
  int64_t n_red = 0;
  int64_t status = 0;
  for (int64_t p = 0; p < n_basis; ++p) {
    rep_of[p] = -1;
    amp_re[p] = 0.0;
    amp_im[p] = 0.0;
  }
  for (int g = 0; g < n_group; ++g) phases[g] = 1.0; /* basis_phases = ones(...) */

  for (int64_t ivec = 0; ivec < n_basis && status == 0; ++ivec) {
    if (visited[ivec]) continue;
    uint64_t bvec = basis[ivec];
    int compatible = 1;
    for (int i = 1; i < n_group; ++i) {
      uint64_t bp = orc_symmetry_apply(n_sites, bitoffsets, bitwidths, maps + (size_t)i * n_sites, bvec);
      if (flips) bp ^= flips[i];
      cplx ampl = chi_re[i] + I * chi_im[i];
      if (bp < bvec) {
        compatible = 0;
        break;
      } else if (bp == bvec && !(cabs(ampl - 1.0) <= tol)) { /* isapprox(ampl*sgn, 1; atol=tol) */
        compatible = 0;
        break;
      }
      states[i] = bp;
      phases[i] = conj(ampl);
    }
    if (!compatible) continue;
    states[0] = bvec;
    reduced_out[n_red++] = bvec;

    /* Dict{BR,S}: basis_amplitudes[state] = phase for i = 1..G, last writer
     * wins; its length is the orbit size. */
    int n_distinct = 0;
    for (int i = 0; i < n_group; ++i) {
      int later = 0;
      for (int j = i + 1; j < n_group; ++j)
        if (states[j] == states[i]) { later = 1; break; }
      if (!later) ++n_distinct;
    }
    double inv_norm = 1.0 / sqrt((double)n_distinct);
    for (int i = 0; i < n_group; ++i) {
      int later = 0;
      for (int j = i + 1; j < n_group; ++j)
        if (states[j] == states[i]) { later = 1; break; }
      if (later) continue;
      int64_t ip = orc_lookup(basis, n_basis, states[i]);
      if (ip < 0) { status = -2; break; }
      visited[ip] = 1;
      rep_of[ip] = ivec;
      cplx v = phases[i] * inv_norm;
      amp_re[ip] = creal(v);
      amp_im[ip] = cimag(v);
    }
  }
  if (status == 0) {
    /* basis_mapping_index: :88-101 */
    for (int64_t p = 0; p < n_basis; ++p) mapping_index[p] = -1;
    for (int64_t r = 0; r < n_red; ++r) mapping_index[orc_lookup(basis, n_basis, reduced_out[r])] = r;
    for (int64_t p = 0; p < n_basis; ++p) {
      int64_t rp = rep_of[p];
      if (rp < 0) continue;
      if (rp == p) continue;
      mapping_index[p] = mapping_index[rp];
    }
  }
  free(rep_of);
  free(visited);
  free(states);
  free(phases);
  return status == 0 ? n_red : status;
}

int64_t orc_symmetry_reduce(int64_t n_basis, const uint64_t* basis, int n_sites,
                            const int32_t* bitoffsets, const int32_t* bitwidths, int n_group,
                            const int32_t* maps, const double* chi_re, const double* chi_im,
                            double tol, uint64_t* reduced_out, int64_t* mapping_index,
                            double* amp_re, double* amp_im) {
  return orc_symmetry_reduce_x(n_basis, basis, n_sites, bitoffsets, bitwidths, n_group, maps, NULL, chi_re,
                               chi_im, tol, reduced_out, mapping_index, amp_re, amp_im);
}

/* ---------------------------------------------------------------------- */
/* Operator representation (plain or reduced).                              */
/* ---------------------------------------------------------------------- */
typedef struct {
  int64_t n_parent;
  const uint64_t* parent; /* HilbertSpaceRepresentation.basis_list */
  int reduced;
  int64_t n_red;
  const uint64_t* red;        /* ReducedHilbertSpaceRepresentation.basis_list */
  const int64_t* map_idx;     /* basis_mapping_index (parent length)          */
  const double* map_re;       /* basis_mapping_amplitude (parent length)      */
  const double* map_im;
  int64_t n_terms;
  const uint64_t* mask;
  const uint64_t* row;
  const uint64_t* col;
  const double* a_re;
  const double* a_im;
} orc_opr;

static int64_t opr_dim(const orc_opr* o) { return o->reduced ? o->n_red : o->n_parent; }

/* Row iterator.  Plain: src/Representation/operator_representation.jl:66-79
 * over the SumOperator row iterator src/Operator/operator_iterator.jl:48-54.
 * Reduced: src/Symmetry/reduced_operator_representation.jl:57-85.
 * Writes up to n_terms (index, value) pairs, index -1 when invalid. */
static int64_t row_elements(const orc_opr* o, int64_t irow, int64_t* idx, cplx* val) {
  int64_t n = 0;
  if (!o->reduced) {
    uint64_t brow = o->parent[irow];
    for (int64_t t = 0; t < o->n_terms; ++t) {
      if ((brow & o->mask[t]) != o->row[t]) continue;
      uint64_t bcol = (brow & ~o->mask[t]) | o->col[t];
      idx[n] = orc_lookup(o->parent, o->n_parent, bcol);
      val[n] = o->a_re[t] + I * o->a_im[t];
      ++n;
    }
    return n;
  }
  uint64_t brow = o->red[irow];
  int64_t irow_p = orc_lookup(o->parent, o->n_parent, brow);
  cplx ampl_row = o->map_re[irow_p] + I * o->map_im[irow_p];
  cplx inv_ampl_row = 1.0 / ampl_row;
  for (int64_t t = 0; t < o->n_terms; ++t) {
    if ((brow & o->mask[t]) != o->row[t]) continue;
    uint64_t bcol = (brow & ~o->mask[t]) | o->col[t];
    cplx ampl = o->a_re[t] + I * o->a_im[t];
    int64_t icol_p = orc_lookup(o->parent, o->n_parent, bcol);
    int64_t icol_r = (icol_p >= 0) ? o->map_idx[icol_p] : -1;
    if (icol_r < 0) {
      idx[n] = -1;
      val[n] = ampl;
    } else {
      cplx ampl_col = o->map_re[icol_p] + I * o->map_im[icol_p];
      idx[n] = icol_r;
      val[n] = ampl * ampl_col * inv_ampl_row;
    }
    ++n;
  }
  return n;
}

/* Column iterator.  Plain: operator_representation.jl:90-103 over
 * operator_iterator.jl:57-63.  Reduced: reduced_operator_representation.jl:88-116. */
static int64_t col_elements(const orc_opr* o, int64_t icol, int64_t* idx, cplx* val) {
  int64_t n = 0;
  if (!o->reduced) {
    uint64_t bcol = o->parent[icol];
    for (int64_t t = 0; t < o->n_terms; ++t) {
      if ((bcol & o->mask[t]) != o->col[t]) continue;
      uint64_t brow = (bcol & ~o->mask[t]) | o->row[t];
      idx[n] = orc_lookup(o->parent, o->n_parent, brow);
      val[n] = o->a_re[t] + I * o->a_im[t];
      ++n;
    }
    return n;
  }
  uint64_t bcol = o->red[icol];
  int64_t icol_p = orc_lookup(o->parent, o->n_parent, bcol);
  cplx ampl_col = o->map_re[icol_p] + I * o->map_im[icol_p];
  cplx inv_ampl_col = 1.0 / conj(ampl_col);
  for (int64_t t = 0; t < o->n_terms; ++t) {
    if ((bcol & o->mask[t]) != o->col[t]) continue;
    uint64_t brow = (bcol & ~o->mask[t]) | o->row[t];
    cplx ampl = o->a_re[t] + I * o->a_im[t];
    int64_t irow_p = orc_lookup(o->parent, o->n_parent, brow);
    int64_t irow_r = (irow_p >= 0) ? o->map_idx[irow_p] : -1;
    if (irow_r < 0) {
      idx[n] = -1;
      val[n] = ampl;
    } else {
      cplx ampl_row = conj(o->map_re[irow_p] + I * o->map_im[irow_p]);
      idx[n] = irow_r;
      val[n] = ampl * ampl_row * inv_ampl_col;
    }
    ++n;
  }
  return n;
}

static void fill_opr(orc_opr* o, int64_t n_parent, const uint64_t* parent, int reduced, int64_t n_red,
                     const uint64_t* red, const int64_t* map_idx, const double* map_re,
                     const double* map_im, int64_t n_terms, const uint64_t* mask,
                     const uint64_t* row, const uint64_t* col, const double* a_re,
                     const double* a_im) {
  o->n_parent = n_parent;
  o->parent = parent;
  o->reduced = reduced;
  o->n_red = n_red;
  o->red = red;
  o->map_idx = map_idx;
  o->map_re = map_re;
  o->map_im = map_im;
  o->n_terms = n_terms;
  o->mask = mask;
  o->row = row;
  o->col = col;
  o->a_re = a_re;
  o->a_im = a_im;
}

#define ORC_OPR_ARGS                                                                         \
  int64_t n_parent, const uint64_t *parent, int reduced, int64_t n_red, const uint64_t *red, \
      const int64_t *map_idx, const double *map_re, const double *map_im, int64_t n_terms,   \
      const uint64_t *mask, const uint64_t *row, const uint64_t *col, const double *a_re,    \
      const double *a_im
#define ORC_OPR_PASS \
  n_parent, parent, reduced, n_red, red, map_idx, map_re, map_im, n_terms, mask, row, col, a_re, a_im

/* Row / column element lists for tests of the iterators themselves.
 * Returns the number of elements written (<= n_terms). */
int64_t orc_row_iterator(ORC_OPR_ARGS, int64_t irow, int64_t* idx, double* v_re, double* v_im) {
  orc_opr o;
  fill_opr(&o, ORC_OPR_PASS);
  cplx* val = (cplx*)malloc(sizeof(cplx) * (size_t)(n_terms > 0 ? n_terms : 1));
  int64_t n = row_elements(&o, irow, idx, val);
  for (int64_t k = 0; k < n; ++k) {
    v_re[k] = creal(val[k]);
    v_im[k] = cimag(val[k]);
  }
  free(val);
  return n;
}

int64_t orc_col_iterator(ORC_OPR_ARGS, int64_t icol, int64_t* idx, double* v_re, double* v_im) {
  orc_opr o;
  fill_opr(&o, ORC_OPR_PASS);
  cplx* val = (cplx*)malloc(sizeof(cplx) * (size_t)(n_terms > 0 ? n_terms : 1));
  int64_t n = col_elements(&o, icol, idx, val);
  for (int64_t k = 0; k < n; ++k) {
    v_re[k] = creal(val[k]);
    v_im[k] = cimag(val[k]);
  }
  free(val);
  return n;
}

/* apply_serial!(out, opr, state) (transpose = 0):
 *   src/Representation/abstract_operator_representation.jl:296-316
 * apply_serial!(out, state, opr) (transpose = 1): :327-347.
 * Adds into out (it does not zero it).  Vectors are interleaved complex. */
void orc_apply(ORC_OPR_ARGS, int transpose, const double* x, double* y) {
  orc_opr o;
  fill_opr(&o, ORC_OPR_PASS);
  int64_t dim = opr_dim(&o);
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_terms > 0 ? n_terms : 1));
  cplx* val = (cplx*)malloc(sizeof(cplx) * (size_t)(n_terms > 0 ? n_terms : 1));
  const cplx* xs = (const cplx*)x;
  cplx* ys = (cplx*)y;
  for (int64_t i = 0; i < dim; ++i) {
    int64_t n = transpose ? col_elements(&o, i, idx, val) : row_elements(&o, i, idx, val);
    for (int64_t k = 0; k < n; ++k) {
      if (idx[k] >= 0 && idx[k] < dim) {
        if (transpose)
          ys[i] += xs[idx[k]] * val[k];
        else
          ys[i] += val[k] * xs[idx[k]];
      }
    }
  }
  free(idx);
  free(val);
}

/* Row-by-key sort used for the per-column dictionary (stable insertion sort;
 * stability keeps the per-key accumulation order of get(colvec,..)+ampl). */
static void sort_pairs(int64_t n, int64_t* idx, cplx* val) {
  for (int64_t i = 1; i < n; ++i) {
    int64_t ki = idx[i];
    cplx vi = val[i];
    int64_t j = i - 1;
    while (j >= 0 && idx[j] > ki) {
      idx[j + 1] = idx[j];
      val[j + 1] = val[j];
      --j;
    }
    idx[j + 1] = ki;
    val[j + 1] = vi;
  }
}

/* sparse_serial(opr; tol): abstract_operator_representation.jl:145-169 with
 * choptol! (src/util.jl:88-94): per column, merge duplicate rows by summing
 * in iteration order, drop |v| < tol, sort rows ascending.  CSC, 0-based.
 * cap: capacity of rowval/nz; returns nnz, or -(needed) if cap is too small. */
int64_t orc_sparse(ORC_OPR_ARGS, double tol, int64_t* colptr, int64_t* rowval, double* nz_re,
                   double* nz_im, int64_t cap) {
  orc_opr o;
  fill_opr(&o, ORC_OPR_PASS);
  int64_t dim = opr_dim(&o);
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_terms > 0 ? n_terms : 1));
  cplx* val = (cplx*)malloc(sizeof(cplx) * (size_t)(n_terms > 0 ? n_terms : 1));
  int64_t nnz = 0;
  int overflow = 0;
  colptr[0] = 0;
  for (int64_t c = 0; c < dim; ++c) {
    int64_t n = col_elements(&o, c, idx, val);
    /* keep valid rows only: `if 1 <= irow <= m` */
    int64_t k2 = 0;
    for (int64_t k = 0; k < n; ++k) {
      if (idx[k] >= 0 && idx[k] < dim) {
        idx[k2] = idx[k];
        val[k2] = val[k];
        ++k2;
      }
    }
    sort_pairs(k2, idx, val);
    int64_t k = 0;
    while (k < k2) {
      int64_t r = idx[k];
      cplx s = 0.0;
      while (k < k2 && idx[k] == r) {
        s = s + val[k];
        ++k;
      }
      if (cabs(s) < tol) continue; /* choptol!: abs(v) < tol */
      if (nnz < cap) {
        rowval[nnz] = r;
        nz_re[nnz] = creal(s);
        nz_im[nnz] = cimag(s);
      } else {
        overflow = 1;
      }
      ++nnz;
    }
    colptr[c + 1] = nnz;
  }
  free(idx);
  free(val);
  return overflow ? -nnz : nnz;
}

/* sparse_parallel(opr; tol): abstract_operator_representation.jl:172-204.
 * The columns are split into one contiguous block per thread (splitrange);
 * each block builds its columns as sparse_serial does (merge in iteration
 * order, choptol!, rows ascending) into block-local lists, and the lists are
 * concatenated in block order after the cumulative sum of the column sizes.
 * Same CSC as orc_sparse; OpenMP threads stand for Julia's threads.  The
 * arrays are allocated here (colptr: dim + 1, the others: nnz entries) and
 * released with orc_free.  Returns nnz, or -1 when an allocation failed. */
int64_t orc_sparse_parallel(ORC_OPR_ARGS, double tol, int n_threads, int64_t** colptr_out, int64_t** rowval_out,
                            double** nz_re_out, double** nz_im_out) {
  orc_opr o;
  fill_opr(&o, ORC_OPR_PASS);
  const int64_t dim = opr_dim(&o);
  int nblocks = n_threads > 0 ? n_threads : 1;
  if (nblocks > dim && dim > 0) nblocks = (int)dim;
  int64_t** lrow = (int64_t**)calloc((size_t)nblocks, sizeof(int64_t*));
  cplx** lval = (cplx**)calloc((size_t)nblocks, sizeof(cplx*));
  int64_t* lsize = (int64_t*)calloc((size_t)nblocks, sizeof(int64_t));
  int64_t* colptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(dim + 1));
  if (!lrow || !lval || !lsize || !colptr) {
    free(lrow);
    free(lval);
    free(lsize);
    free(colptr);
    return -1;
  }
  int failed = 0;
#pragma omp parallel for num_threads(nblocks) schedule(static, 1) reduction(| : failed)
  for (int ib = 0; ib < nblocks; ++ib) {
    const int64_t c0 = dim * ib / nblocks, c1 = dim * (ib + 1) / nblocks;
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_terms > 0 ? n_terms : 1));
    cplx* val = (cplx*)malloc(sizeof(cplx) * (size_t)(n_terms > 0 ? n_terms : 1));
    int64_t capb = 1024, nb = 0;
    int64_t* rb = (int64_t*)malloc(sizeof(int64_t) * (size_t)capb);
    cplx* vb = (cplx*)malloc(sizeof(cplx) * (size_t)capb);
    int bad = !rb || !vb || !idx || !val;
    for (int64_t c = c0; c < c1 && !bad; ++c) {
      int64_t n = col_elements(&o, c, idx, val);
      int64_t k2 = 0;
      for (int64_t k = 0; k < n; ++k) {
        if (idx[k] >= 0 && idx[k] < dim) { /* `if 1 <= irow <= m` */
          idx[k2] = idx[k];
          val[k2] = val[k];
          ++k2;
        }
      }
      sort_pairs(k2, idx, val);
      const int64_t before = nb;
      int64_t k = 0;
      while (k < k2 && !bad) {
        const int64_t r = idx[k];
        cplx sum = 0.0;
        while (k < k2 && idx[k] == r) {
          sum = sum + val[k];
          ++k;
        }
        if (cabs(sum) < tol) continue; /* choptol! */
        if (nb == capb) {
          capb *= 2;
          int64_t* rb2 = (int64_t*)realloc(rb, sizeof(int64_t) * (size_t)capb);
          if (rb2) rb = rb2;
          cplx* vb2 = (cplx*)realloc(vb, sizeof(cplx) * (size_t)capb);
          if (vb2) vb = vb2;
          if (!rb2 || !vb2) {
            bad = 1;
            break;
          }
        }
        rb[nb] = r;
        vb[nb] = sum;
        ++nb;
      }
      colptr[c + 1] = nb - before; /* column size, summed below */
    }
    failed |= bad;
    lrow[ib] = rb;
    lval[ib] = vb;
    lsize[ib] = nb;
    free(idx);
    free(val);
  }
  colptr[0] = 0;
  for (int64_t c = 0; c < dim; ++c) colptr[c + 1] += colptr[c];
  const int64_t nnz = colptr[dim];
  int64_t* rowval = NULL;
  double *nr = NULL, *ni = NULL;
  if (!failed) {
    rowval = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
    nr = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    ni = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    failed = !rowval || !nr || !ni;
  }
  if (!failed) {
    int64_t off = 0;
    for (int ib = 0; ib < nblocks; ++ib) {
      for (int64_t k = 0; k < lsize[ib]; ++k) {
        rowval[off + k] = lrow[ib][k];
        nr[off + k] = creal(lval[ib][k]);
        ni[off + k] = cimag(lval[ib][k]);
      }
      off += lsize[ib];
    }
  }
  for (int ib = 0; ib < nblocks; ++ib) {
    free(lrow[ib]);
    free(lval[ib]);
  }
  free(lrow);
  free(lval);
  free(lsize);
  if (failed) {
    free(colptr);
    free(rowval);
    free(nr);
    free(ni);
    return -1;
  }
  *colptr_out = colptr;
  *rowval_out = rowval;
  *nz_re_out = nr;
  *nz_im_out = ni;
  return nnz;
}

void orc_free(void* p) { free(p); }

/* get_element(opr, irow, icol).  Plain: operator_representation.jl:109-119
 * (the SumOperator element, operator_iterator.jl:66-83).  Reduced:
 * reduced_operator_representation.jl:120-138 (sum over the column iterator). */
void orc_get_element(ORC_OPR_ARGS, int64_t irow, int64_t icol, double* out_re, double* out_im) {
  orc_opr o;
  fill_opr(&o, ORC_OPR_PASS);
  cplx s = 0.0;
  if (!reduced) {
    uint64_t br = parent[irow], bc = parent[icol];
    for (int64_t t = 0; t < n_terms; ++t) {
      if (((br & mask[t]) == row[t]) && (((br & ~mask[t]) | col[t]) == bc)) s += a_re[t] + I * a_im[t];
    }
  } else {
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_terms > 0 ? n_terms : 1));
    cplx* val = (cplx*)malloc(sizeof(cplx) * (size_t)(n_terms > 0 ? n_terms : 1));
    int64_t n = col_elements(&o, icol, idx, val);
    for (int64_t k = 0; k < n; ++k)
      if (idx[k] == irow) s = s + val[k];
    free(idx);
    free(val);
  }
  *out_re = creal(s);
  *out_im = cimag(s);
}

/* symmetry_reduce_serial(rhsr, large_vector): src/Symmetry/symmetry_reduce.jl:38-56
 * small[r] += conj(U[p]) * large[p]; and symmetry_unreduce (:208-225)
 * large[p] += U[p] * small[r]. Interleaved complex vectors. */
void orc_vector_reduce(int64_t n_parent, const int64_t* map_idx, const double* map_re,
                       const double* map_im, const double* large, double* small) {
  const cplx* L = (const cplx*)large;
  cplx* S = (cplx*)small;
  for (int64_t p = 0; p < n_parent; ++p) {
    int64_t r = map_idx[p];
    if (r >= 0) S[r] += conj(map_re[p] + I * map_im[p]) * L[p];
  }
}

void orc_vector_unreduce(int64_t n_parent, const int64_t* map_idx, const double* map_re,
                         const double* map_im, const double* small, double* large) {
  const cplx* S = (const cplx*)small;
  cplx* L = (cplx*)large;
  for (int64_t p = 0; p < n_parent; ++p) {
    int64_t r = map_idx[p];
    if (r >= 0) L[p] += (map_re[p] + I * map_im[p]) * S[r];
  }
}
