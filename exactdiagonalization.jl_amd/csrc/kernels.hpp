// Host-side launchers of the edhip kernels (implemented in the .hip files).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"
#include "host_logic.hpp"

namespace edhip {

// ---- space.hip -------------------------------------------------------------
// words[i] = unrank(first + i), i < count
hipError_t launch_enumerate_sector(uint64_t first, uint64_t count, int n_bits, const uint64_t* d_W, uint64_t* d_out,
                                   hipStream_t st);
// One site of a multi-quantum-number sector walk (k_enumerate_qn): its table
// entries start at off, ns states, bit offset shift, n_upper upper sums.
struct QnSite {
  int32_t off, ns, shift, n_upper;
};
// Tables up to this size are staged in LDS by the enumeration kernel.
constexpr size_t kQnStageBytes = 48 * 1024;
// words[i] = i-th word of a multi-QN sector in ascending order, i < count
hipError_t launch_enumerate_qn(uint64_t count, int n_sites, const QnSite* d_sites, int n_entries, const uint64_t* d_cnt,
                               const int32_t* d_nxt, uint64_t* d_out, hipStream_t st);
// flags[key >> shift] = 1 for every key
hipError_t launch_mark_tops(const uint64_t* d_keys, int64_t dim, int shift, uint32_t* d_flags, hipStream_t st);
// x spread over the word positions of X (before each product); with d_orbit,
// x / sqrt(orbit) (the orbit factor of the images, X.scaled).
hipError_t launch_xmap_fill(int prec, const XMapDev& X, const uint64_t* d_keys, int64_t dim, const void* x,
                            const uint16_t* d_orbit, hipStream_t st);
// prefix[j] = lower_bound(keys, j << shift), j <= n_buckets
hipError_t launch_build_prefix(const uint64_t* d_keys, int64_t dim, int shift, uint64_t n_buckets, uint32_t* d_prefix,
                               hipStream_t st);

struct ReduceSource {
  // either a popcount sector (unranked on the fly) or an explicit device list
  const uint64_t* list;  // nullptr for a sector
  uint64_t n_words;
  int n_bits;
  int single_popcount;   // 1: Gosper stepping is valid inside a chunk
  const uint64_t* W;     // completion counts (sector)
};

// pass 0: per-thread counts of representatives; pass 1: write them (ascending)
hipError_t launch_reduce_count(const ReduceSource& src, uint64_t chunk, uint64_t n_threads, const GroupDev& g,
                               const GroupLdsLayout& lay, const uint64_t* d_Wsrc, uint32_t* d_counts, hipStream_t st);
hipError_t launch_reduce_write(const ReduceSource& src, uint64_t chunk, uint64_t n_threads, const GroupDev& g,
                               const GroupLdsLayout& lay, const uint64_t* d_offsets, uint64_t* d_reps,
                               uint16_t* d_orbit, hipStream_t st);

// For every parent word p: mapping index (rep index or -1) and amplitude
// U[p] = phase[h] / sqrt(N_s) (0 when excluded).
hipError_t launch_export_mapping(const ReduceSource& src, const GroupDev& g, const GroupLdsLayout& lay,
                                 const IndexDev& reduced_index, const uint16_t* d_orbit, int64_t* d_idx, double* d_re,
                                 double* d_im, hipStream_t st);

hipError_t launch_lookup(const IndexDev& ix, int64_t n, const uint64_t* d_keys, int64_t* d_idx, hipStream_t st);

// Vector maps between a parent space and its reduction, in precision prec:
// large[p] = U[p] small[r(p)] (unreduce), small[r] = sum conj(U[p]) large[p]
// over the orbit of representative r (reduce).
hipError_t launch_vec_unreduce(const ReduceSource& src, const GroupDev& g, const GroupLdsLayout& lay,
                               const IndexDev& reduced_index, const uint16_t* d_orbit, int prec, const void* d_small,
                               void* d_large, hipStream_t st);
hipError_t launch_vec_reduce(const uint64_t* d_reps, const uint16_t* d_orbit, int64_t dim, const GroupDev& g,
                             const GroupLdsLayout& lay, const IndexDev& parent_index, int prec, const void* d_large,
                             void* d_small, hipStream_t st);

// Blocked rank index over the words [0, 384 n_blocks) (zeroed blocks in,
// layout as in IndexDev); d_totals (u32) and d_base (u64) hold n_blocks
// entries of scratch.  Two-phase: call with d_tmp == nullptr to size the
// scan's scratch.
hipError_t launch_bitmap_build(const uint64_t* d_keys, int64_t dim, uint64_t* d_blocks, uint64_t n_blocks,
                               uint32_t* d_totals, uint64_t* d_base, void* d_tmp, size_t& tmp_bytes, hipStream_t st);
hipError_t exclusive_scan_u32_to_u64(const uint32_t* d_in, uint64_t* d_out, uint64_t n, void* d_tmp, size_t& tmp_bytes,
                                     hipStream_t st);
hipError_t exclusive_scan_i64(const int64_t* d_in, int64_t* d_out, uint64_t n, void* d_tmp, size_t& tmp_bytes,
                              hipStream_t st);

// ---- matvec.hip ------------------------------------------------------------
// Off-diagonal term as the matvec sees it: the row matches iff
// (b & mask) == match; its image is (b & ~mask) | image.
struct OffTerm {
  uint64_t mask, match, image;
  double are, aim;
};

struct DiagTerm {
  uint64_t mask, match;
  double are, aim;
};

struct OpDev {
  IndexDev ix;           // lookup of canonical images
  GroupDev g;            // symmetry group
  GroupLdsLayout lay;    // LDS layout of the group tables
  const uint64_t* rows;  // basis words (the space's sorted keys)
  const uint16_t* orbit; // orbit sizes (reduced) or nullptr
  int64_t dim;           // rows of the space (products: the end of the rows computed)
  int64_t row_begin;     // products: first row computed; y holds rows row_begin .. dim - 1 from y[0]
  int n_off;             // number of off-diagonal terms (in the row or column sense)
  int n_diag;
  int word_bits;         // bits of the words and of every diagonal term's mask and match
  int xcd_chunk;         // products: consecutive blocks each XCD takes per window of 8 xcd_chunk (<= 1: plain order)
  const OffTerm* off;    // row form: match = R, image = C ; column form: match = C, image = R
  // The matvec's copy of off: every run of equal masks padded to an even
  // length (padding terms carry the run's mask, a match outside it and no
  // flip), so terms 2q and 2q + 1 share a mask and are matched together.
  const OffTerm* mv_off;
  int n_mv;              // padded length (0: the matvec matches off term by term)
  int pad_mv_;
  const DiagTerm* diag;
  const double* diag_re; // per-row diagonal element (summed diagonal terms), or nullptr
  const double* diag_im; // its imaginary part, or nullptr when every diagonal amplitude is real
  uint32_t off_lds;      // LDS offset of the staged off-diagonal terms
  uint32_t w_lds;        // LDS offset of the staged rank table (rank mode)
  uint32_t w_bytes;
  XMapDev xm;            // matvec: x read from its spread over word positions (xm.values != null)
  uint32_t codes_lds;    // LDS offset of xm's code table, after lds_total (0: no room)
  uint32_t lds_total;    // LDS of the kernels (the matvec adds xm's code table)
};

enum Prec : int { kF64 = 0, kC128 = 1, kF32 = 2, kC64 = 3 };

// y = (accumulate ? y : 0) + A x   (transpose=0: row form; 1: column form, y = x A)
hipError_t launch_matvec(const OpDev& op, int prec, int transpose, int accumulate, const void* x, void* y,
                         hipStream_t st);

// ---- assembled rows (row form): entries col | cls << col_bits per image ----
constexpr int kAsmMaxClasses = 256;
struct AsmFillDev {
  const uint64_t* rowptr;  // dim + 1 offsets of the rows' entries
  uint32_t* ent;
  const uint32_t* acls;    // per matvec term: amplitude class, 1-based
  const uint32_t* pcls;    // per group element: phase class (nullptr: every phase 1)
  int n_ph;
  int col_bits;
};
struct AsmDev {
  const uint64_t* rowptr;
  const uint32_t* ent;
  const double* w_re;      // n_cls weights, class 0 = 0
  const double* w_im;
  int n_cls;
  int col_bits;
};
// counts[r] = entries of row r (terms matching it)
hipError_t launch_asm_count(const OpDev& op, uint32_t* d_counts, hipStream_t st);
hipError_t launch_asm_fill(const OpDev& op, const AsmFillDev& a, hipStream_t st);
// xt = x / sqrt(orbit) (orbit != nullptr), in precision prec
hipError_t launch_asm_scale(int prec, int64_t dim, const uint16_t* d_orbit, const void* x, void* xt, hipStream_t st);
// y = (accumulate ? y : 0) + A x from the entries; xt = x scaled as above (x itself without orbits);
// lanes = lanes per row (4, 8, 16 or 32), batch = entries per lane loaded together (4, 8 or 16)
hipError_t launch_asm_spmv(const OpDev& op, const AsmDev& a, int prec, int lanes, int batch, int accumulate,
                           const void* x, const void* xt, void* y, hipStream_t st);

// Diagonal element of every row: re[r] (+ i im[r] when im != nullptr) = sum of
// the amplitudes of the diagonal terms that match row r.
hipError_t launch_diag_values(const OpDev& op, double* d_re, double* d_im, hipStream_t st);

// Total number of off-diagonal images the row form generates (atomic u64 add).
hipError_t launch_count_images(const OpDev& op, unsigned long long* d_total, hipStream_t st);

// Raw iterator elements (index or -1, value) in term order for the given rows.
hipError_t launch_iterate(const OpDev& op, const OffTerm* all_terms, int n_all, int transpose, int64_t n_rows,
                          const int64_t* d_rows, int64_t* d_idx, double* d_re, double* d_im, hipStream_t st);

// Merged, chopped, ascending entries of rows (row form, transpose = 0) or
// columns (column form, transpose = 1): entry k is vector d_sel[k] (k itself
// when d_sel is null), up to slot entries each, d_count[k] kept.
hipError_t launch_sparse_columns(const OpDev& op, int transpose, double tol, int slot, int64_t n_sel,
                                 const int64_t* d_sel, int64_t* d_scratch_idx, double* d_scratch_re,
                                 double* d_scratch_im, int64_t* d_count, hipStream_t st);
// Dense column-major matrix (column form operator) added into a cleared
// dim x dim buffer of doubles (complex_out: interleaved re, im pairs).
hipError_t launch_dense_columns(const OpDev& op, int complex_out, double* d_out, hipStream_t st);
hipError_t launch_sparse_compact(int64_t dim, int slot, const int64_t* d_colptr, const int64_t* d_scratch_idx,
                                 const double* d_scratch_re, const double* d_scratch_im, int64_t* d_rowval,
                                 double* d_re, double* d_im, hipStream_t st);

// Per-term bilinear forms g_t = sum_r conj(phi_r) sum_{images of t} w psi_s,
// deterministic two-stage reduction.  n_terms = n_diag + n_off.
hipError_t launch_term_expect(const OpDev& op, int prec, const void* phi, const void* psi, double* d_partial,
                              int64_t n_blocks, double* d_out, hipStream_t st);
int64_t term_expect_blocks(const OpDev& op);
// rows of n_terms * 2 doubles the kernel needs in d_partial (per-block sums
// and the reduction's chunk sums); terms per pass for vectors of nc real
// components (0: the tables leave no LDS room)
int64_t term_expect_scratch_rows(const OpDev& op);

// Fixed device scratch of the allocation-free term pass (one array per GPU,
// held until the process ends): room for the per-block partial sums and the
// result of at most kNoAllocMaxTerms terms.  term_expect_noalloc_setup runs
// once per device with that device current (the context's creation) and
// keeps the device address of the array; term_expect_noalloc_scratch returns
// it (nullptr before the setup).
constexpr int kNoAllocMaxTerms = 512;
hipError_t term_expect_noalloc_setup(int device);
double* term_expect_noalloc_scratch(int device);
size_t term_expect_noalloc_doubles();
int term_expect_window(const OpDev& op, int nc);

// ---- fast path: bond-pair term groups + register canonicaliser + hash lookup
// Terms grouped by bit mask (1 or 2 bits).  For a row b the group's pattern
// code = b[p] | b[q] << 1 (q = p for one-bit masks) selects the summed
// diagonal amplitude and at most one off-diagonal entry (new pattern code +
// amplitude).  meta[g] = p | q << 8 | offmask << 16 | newcode(4 x 2 bits) << 24.
struct PairGroupsDev {
  int n_groups;
  int pad_;
  const uint32_t* meta;
  const double* dre;  // 4 per group
  const double* dim;
  const double* ore;  // 4 per group
  const double* oim;
  double dconst_re, dconst_im;  // identity terms (empty mask)
};

struct FastOpDev {
  OpDev base;  // index, group, layout, rows, orbit, dim, rank table offsets (off terms unused)
  PairGroupsDev pg;
  HashDev hash;      // slots == nullptr: lookup through base.ix, x gathered separately
  uint32_t pg_lds;   // LDS offset of the staged groups
  uint32_t lds_total;
};

constexpr int kMaxPairGroups = 256;

hipError_t launch_matvec_pairs(const FastOpDev& op, int prec, int transpose, int accumulate, const void* x, void* y,
                               hipStream_t st);
hipError_t launch_hash_build(const uint64_t* d_keys, int64_t dim, const HashDev& h, uint32_t* d_slot_of,
                             hipStream_t st);
hipError_t launch_hash_scatter(int prec, int64_t dim, const void* x, const uint32_t* d_slot_of, const HashDev& h,
                               hipStream_t st);

// ---- vecops.hip ------------------------------------------------------------
// All reductions accumulate in double and are deterministic (fixed tree).
hipError_t vec_dot(int prec, int64_t n, const void* x, const void* y, double* d_partial, double* d_out2,
                   hipStream_t st);  // <x, y> (conjugated x), d_out2 = {re, im}
hipError_t vec_multi_dot(int prec, int64_t n, int k, const void* V, int64_t ldv, const void* w, double* d_partial,
                         double* d_out, hipStream_t st);  // d_out[2j..] = <V_j, w>
hipError_t vec_multi_axpy(int prec, int64_t n, int k, const void* V, int64_t ldv, const double* d_coef, void* w,
                          hipStream_t st);  // w -= sum_j coef_j V_j (complex coef)
hipError_t vec_axpby(int prec, int64_t n, double ar, double ai, const void* x, double br, double bi, void* y,
                     hipStream_t st);  // y = a x + b y
hipError_t vec_convert(int prec_in, int prec_out, int64_t n, const void* x, void* y, hipStream_t st);
hipError_t vec_gaussian(int prec, int64_t n, uint64_t seed, void* x, hipStream_t st);
hipError_t vec_fill_zero(int prec, int64_t n, void* x, hipStream_t st);
// out_l = sum_j coef[j*k + l] V_j for l < k (real coefficients, double accumulation)
hipError_t vec_combine(int prec_v, int prec_out, int64_t n, int m, const void* V, int64_t ldv, int k,
                       const double* d_coef, void* out, int64_t ldo, hipStream_t st);
int64_t vec_partial_blocks(int64_t n);

}  // namespace edhip
