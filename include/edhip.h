/*
 * edhip: MI355X-native (gfx950, HIP) Hamiltonian application on Sz-sector
 * and symmetry-reduced bases — the C ABI.
 *
 * This is the drop-in boundary for the ExactDiagonalization.jl (v0.14.3)
 * path  HilbertSpaceSector -> represent -> symmetry_reduce -> represent(rhsr, H)
 * -> mul!/apply!/sparse -> eigensolver.  Each entry point names the reference
 * interface it replaces (file:line, relative to the reference root).
 *
 * Conventions
 *  - Every call returns an int status (EDHIP_OK = 0); edhip_last_error(ctx)
 *    gives the message of the last failure on that context.  No C++
 *    exception crosses the ABI.
 *  - Indices are 0-based and 64-bit; -1 marks "not present", as the
 *    reference's -1 (its 1-based wrapper adds one).
 *  - Words (basis states) are uint64 bit patterns: site i (0-based) occupies
 *    bits [bitoffset_i, bitoffset_i + bitwidth_i), as in HilbertSpace
 *    (src/HilbertSpace/hilbert_space.jl:25-41).
 *  - The library owns device memory behind opaque handles.  Pointers passed
 *    with a *_host suffix are host memory borrowed for the call; all others
 *    are device pointers, ordered on the context's stream.
 *  - A reduced space keeps a pointer to its parent: the parent must outlive it.
 *    A handle is not re-entrant; calls on one context are serialised by the
 *    caller.
 */
#ifndef EDHIP_H
#define EDHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (mapped by the host wrappers onto the reference's exceptions) */
#define EDHIP_OK 0
#define EDHIP_ERR_ARGUMENT 1  /* ArgumentError */
#define EDHIP_ERR_DIMENSION 2 /* DimensionMismatch */
#define EDHIP_ERR_MEMORY 3    /* device allocation failed */
#define EDHIP_ERR_DEVICE 4    /* HIP runtime error / no device */
#define EDHIP_ERR_KEY 5       /* KeyError */
#define EDHIP_ERR_BOUNDS 6    /* BoundsError */
#define EDHIP_ERR_NOCONV 7    /* iterative solver hit its iteration limit */

/* vector element types */
#define EDHIP_F64 0  /* Float64 */
#define EDHIP_C128 1 /* ComplexF64 */
#define EDHIP_F32 2  /* Float32 */
#define EDHIP_C64 3  /* ComplexF32 */

typedef struct edhip_ctx edhip_ctx;
typedef struct edhip_space edhip_space;
typedef struct edhip_op edhip_op;

/* ---- context ------------------------------------------------------------ */
int edhip_create(int device, edhip_ctx** out);
void edhip_destroy(edhip_ctx* ctx);
const char* edhip_last_error(const edhip_ctx* ctx);
/* Use an external hipStream_t (NULL restores the context's own stream). */
int edhip_set_stream(edhip_ctx* ctx, void* hip_stream);
/* The hipStream_t the context's calls run on (its own non-blocking stream,
 * or the one set by edhip_set_stream), so that a caller can order its work
 * against the library's with events instead of host synchronisation.  No
 * counterpart in the reference (Julia has no streams at this boundary). */
int edhip_get_stream(const edhip_ctx* ctx, void** hip_stream);
int edhip_synchronize(edhip_ctx* ctx);
/* Device memory of the context's device (plain hipMalloc / hipFree). */
int edhip_malloc(edhip_ctx* ctx, uint64_t bytes, void** out);
int edhip_free(edhip_ctx* ctx, void* ptr);
/* The eigensolver and MINRES keep their vectors (Krylov basis, restart
 * scratch, Ritz vectors) in one device workspace per context that grows to
 * the largest request and is reused by later calls; this frees it. */
int edhip_release_workspace(edhip_ctx* ctx);
int edhip_memcpy(edhip_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind /* hipMemcpyKind */);

/* ---- spaces ------------------------------------------------------------- */
/* represent(HilbertSpaceSector(hs, qn)) for sites of bit width 1 whose
 * sector fixes the number of set bits:
 *   src/Representation/hilbert_space_representation.jl:215-230 (represent,
 *   represent_array) over hs_get_basis_list(hss, BR) at :127-206.
 * popcounts: the allowed numbers of set bits (ascending or not).  The basis is
 * the ascending list of n_bits-bit words with an allowed popcount; it is
 * enumerated on the device by unranking and materialised on first use. */
int edhip_space_sector(edhip_ctx* ctx, int n_bits, int n_popcounts, const int32_t* popcounts_host,
                       edhip_space** out);

/* represent(HilbertSpaceSector(hs, qns)) and represent(hs) for any sites:
 *   hs_get_basis_list(hss, BR) at src/Representation/hilbert_space_representation.jl:127-206
 *   (and hs_get_basis_list(hs, BR) at :109-119 when every reachable total is allowed).
 * Site j has n_states[j] states and occupies bits [bitoffsets[j], bitoffsets[j+1])
 * (bitoffsets has n_sites + 1 entries, HilbertSpace.bitoffsets); its state s is
 * stored as the number s in those bits and carries the quantum-number tuple
 * state_qns[(sum_{i<j} n_states[i] + s) * qn_dim + 0 .. qn_dim-1].
 * allowed_qns: n_allowed tuples of qn_dim components; a word belongs to the
 * sector iff the sum of its states' tuples is one of them.  The ascending basis
 * is enumerated on the device by unranking over per-site completion counts. */
int edhip_space_sector_qn(edhip_ctx* ctx, int n_sites, const int32_t* bitoffsets, const int32_t* n_states, int qn_dim,
                          const int32_t* state_qns, int n_allowed, const int32_t* allowed_qns, edhip_space** out);

/* represent(hs, basis_list) / represent_array(hs, basis_list):
 *   hilbert_space_representation.jl:239-256.  The list must be sorted and
 * unique (FrozenSortedArrayIndex validation, src/frozensortedarray.jl:14-22);
 * EDHIP_ERR_ARGUMENT otherwise. */
int edhip_space_from_list(edhip_ctx* ctx, int n_bits, int64_t n, const uint64_t* basis_host, edhip_space** out);

/* symmetry_reduce(hsr, symops_and_amplitudes; tol):
 *   src/Symmetry/symmetry_reduce_generic.jl:7-106 (and the IrrepComponent
 *   route, symmetry_reduce_translation.jl:10-106, after conj of the
 *   characters).
 * bit_perm: n_group x n_bits, bit_perm[g*n_bits + i] = image bit of bit i
 *   (a SitePermutation expanded to bits, symmetry_apply.jl:82-92); element 0
 *   must be the identity and the set must be closed under composition.
 * chi: the amplitudes y_g, |y_g| = 1 (EDHIP_ERR_ARGUMENT otherwise, :27-29).
 * A word is kept iff it is the smallest of its orbit and every element fixing
 * it has |y_g - 1| <= tol.  The kept words are ascending; their orbit sizes
 * are recorded.  Translation groups of an Lx x Ly torus are detected and use
 * the pruned canonicaliser; every other group uses LDS permutation tables. */
int edhip_space_reduce(edhip_ctx* ctx, const edhip_space* parent, int n_group, const int32_t* bit_perm_host,
                       const double* chi_re_host, const double* chi_im_host, double tol, edhip_space** out);

/* symmetry_reduce with operations that may include the global spin flip:
 * GlobalBitFlip (src/Symmetry/bitflipsymmetry.jl:8-38) composed with a site
 * permutation (DirectProductOperation, symmetry_apply.jl:65-76), i.e.
 * g(b) = perm_g(b) ^ flip[g] with flip[g] = 0 or all n_bits bits.
 * flip_host: n_group words or NULL (then as edhip_space_reduce).  The parent
 * must be closed under the flip (a sector whose popcounts p and n_bits - p
 * are allowed together, or a basis list equal to its complement). */
int edhip_space_reduce_ex(edhip_ctx* ctx, const edhip_space* parent, int n_group, const int32_t* bit_perm_host,
                          const uint64_t* flip_host, const double* chi_re_host, const double* chi_im_host,
                          double tol, edhip_space** out);

void edhip_space_destroy(edhip_space* space);
/* dimension(hsr) / dimension(rhsr): hilbert_space_representation.jl:90,
 * reduced_hilbert_space_representation.jl:38 */
int edhip_space_dim(const edhip_space* space, int64_t* dim);
/* hsr.basis_list / rhsr.basis_list (host copy) */
int edhip_space_copy_basis(const edhip_space* space, uint64_t* out_host);
/* device pointer to the sorted basis (materialises a sector basis) */
int edhip_space_device_basis(const edhip_space* space, const uint64_t** out);
/* orbit size |G|/|Stab| of every representative (1 for an unreduced space) */
int edhip_space_copy_orbit(const edhip_space* space, uint16_t* out_host);
/* get(basis_lookup, key, -1) for many keys: src/frozensortedarray.jl:29-48 */
int edhip_space_lookup(const edhip_space* space, int64_t n, const uint64_t* keys_host, int64_t* idx_host);
/* basis_mapping_index / basis_mapping_amplitude of a reduced space over its
 * parent (reduced_hilbert_space_representation.jl:13-22): for every parent
 * word p, the index of its representative (or -1) and U[p] (0 if excluded).
 * Arrays have the parent's dimension.  For parity checks; O(parent). */
int edhip_space_export_mapping(const edhip_space* reduced, int64_t* idx_host, double* amp_re_host,
                               double* amp_im_host);
/* symmetry_reduce(rhsr, large_vector) (src/Symmetry/symmetry_reduce.jl:38-56):
 * small[r] = sum over the parent words p of the orbit of r of conj(U[p]) large[p].
 * symmetry_unreduce(rhsr, small_vector) (symmetry_reduce.jl:208-225):
 * large[p] = U[p] small[r(p)], 0 for parent words outside the reduced space.
 * Device vectors of precision prec; n_large must be the parent's dimension and
 * n_small the reduced one (EDHIP_ERR_DIMENSION otherwise).  Complex
 * characters need a complex precision (the caller promotes, as promote_type
 * does at :43 and :215). */
int edhip_space_vector_reduce(const edhip_space* reduced, int prec, int64_t n_large, const void* large,
                              int64_t n_small, void* small);
int edhip_space_vector_unreduce(const edhip_space* reduced, int prec, int64_t n_small, const void* small,
                                int64_t n_large, void* large);
/* 1 if the space is symmetry reduced; group order and canonicaliser kind
 * (0 none, 1 generic LDS tables, 2 torus translations) */
int edhip_space_info(const edhip_space* space, int* reduced, int* group_order, int* canon_kind);

/* ---- operators ---------------------------------------------------------- */
/* represent(hsr, op) / represent(rhsr, op):
 *   src/Representation/operator_representation.jl:34-36,
 *   src/Symmetry/reduced_operator_representation.jl:40-42.
 * The terms are those of a simplified SumOperator (operator_simplify.jl:32-81)
 * as PureOperator(bitmask, bitrow, bitcol, amplitude) (pure_operator.jl:25-52);
 * bitrow/bitcol outside bitmask give EDHIP_ERR_ARGUMENT (pure_operator.jl:32-36). */
int edhip_op_create(edhip_ctx* ctx, const edhip_space* space, int64_t n_terms, const uint64_t* mask_host,
                    const uint64_t* row_host, const uint64_t* col_host, const double* amp_re_host,
                    const double* amp_im_host, edhip_op** out);
void edhip_op_destroy(edhip_op* op);
/* 1 if the operator or the space's characters need complex arithmetic */
int edhip_op_is_complex(const edhip_op* op, int* is_complex);

/* apply! / mul!:
 *   transpose = 0: out += opr * state  (apply!(out, opr, state),
 *                  abstract_operator_representation.jl:260-267,296-316,358-378)
 *   transpose = 1: out += state * opr  (apply!(out, state, opr), :278-285,327-347)
 *   accumulate = 1 adds into y (apply!), 0 overwrites it (mul!, :110-118).
 * x, y: device vectors of n elements of type prec (EDHIP_F64 ...).  n must
 * equal the dimension (EDHIP_ERR_DIMENSION otherwise).  Real precisions
 * require a real operator and real characters (EDHIP_ERR_ARGUMENT). */
int edhip_apply(edhip_op* op, int prec, int transpose, int accumulate, int64_t n_x, const void* x, int64_t n_y,
                void* y);
/* same on host vectors (copied to and from the device) */
int edhip_apply_host(edhip_op* op, int prec, int transpose, int accumulate, int64_t n_x, const void* x_host,
                     int64_t n_y, void* y_host);
/* One share of a product sharded by rows (no counterpart in the reference;
 * the per-GPU part of a product over contiguous row ranges, with the whole
 * input vector resident, as after an all-gather): y[i] = (opr * state)[row_begin
 * + i] (transpose = 1: (state * opr)[row_begin + i]) for i < row_end -
 * row_begin, overwriting y, by the term-list kernel whatever the strategy.
 * Rows outside [0, dimension] give EDHIP_ERR_ARGUMENT; n_x must equal the
 * dimension and n_y the row count (EDHIP_ERR_DIMENSION).  Every row comes out
 * bitwise as in the whole product. */
int edhip_apply_rows(edhip_op* op, int prec, int transpose, int64_t row_begin, int64_t row_end, int64_t n_x,
                     const void* x, int64_t n_y, void* y);

/* Kernel and lookup used by apply/mul/eigensolver calls (performance knob;
 * every choice gives the same matrix).  The automatic choice (0, 0) is the
 * product path; the bond-pair kernel (kernel 2) and the open-addressing
 * table (lookup 2) are never chosen automatically, measured slower on every
 * workload tried (e.g. 514 vs 97 ms per 6x6 product for the table), and are
 * kept only as alternative implementations the parity tests cross-check
 * against (test-only).
 *   kernel: 0 auto (term list), 1 term-list kernel, 2 bond-pair kernel
 *           (operators whose terms act on one or two bits, plain or
 *           torus-translation spaces; test-only),
 *           3 assembled rows: the request itself finds every row's
 *             off-diagonal images once, with the term-list kernel's
 *             matching, canonicalisation and index, and keeps them as
 *             4-byte entries (column | weight class << column bits, at
 *             most 2^31 states and 256 classes: distinct term amplitudes
 *             times distinct character phases); row-form products then
 *             stream the entries (the column form keeps the term list).
 *             The arrays take 4 bytes per image plus 8 per row and stay
 *             until another kernel is requested or the operator is
 *             destroyed; EDHIP_ERR_MEMORY when they do not fit (nothing
 *             kept).  An eigensolver or MINRES workspace that does not fit
 *             beside them drops them and sets the strategy back to auto.
 *   lookup: 0 auto (term-list kernel: x spread over word positions when
 *             it fits, for full popcount sectors, where the spread is at
 *             most 256 MB, and where it holds at most 24 (4-byte
 *             elements) or 8 (wider ones) positions per basis word and
 *             the vector itself exceeds 256 MB; else
 *             bitmap, else sorted search, or rank arithmetic for a full
 *             popcount sector),
 *           1 sorted search behind the prefix table (rank arithmetic for a
 *             full popcount sector),
 *           2 open-addressing table of the basis with the input vector
 *             embedded (bond-pair kernel only; built on first use;
 *             test-only),
 *           3 bitmap rank index (words of at most 36 bits),
 *           4 x spread over the positions of the words (term-list kernel):
 *             one load per image, no rank; the spread (at most 32 elements
 *             per basis word) is allocated per element size on first use
 *             and refreshed before every product (and before the term pass
 *             of edhip_term_expect, which reads psi through it);
 *   on a full popcount sector lookups 1 to 3 use rank arithmetic.
 * edhip_op_last_strategy reports what the last product used. */
int edhip_op_set_strategy(edhip_op* op, int kernel, int lookup);
int edhip_op_last_strategy(const edhip_op* op, int* kernel, int* lookup);
/* Assembled rows of the operator (kernel 3): entries, device bytes and the
 * seconds their construction took (zeros when none are held). */
int edhip_op_assembled_info(const edhip_op* op, uint64_t* entries, uint64_t* bytes, double* build_seconds);

/* Number of off-diagonal images the row iterators generate over all rows
 * (terms whose bitrow matches), i.e. the off-diagonal elements one apply!
 * visits; used for the element-throughput metric. */
int edhip_op_count_images(edhip_op* op, int64_t* n_images);

/* get_row_iterator / get_column_iterator for a list of rows (columns):
 *   operator_representation.jl:66-103, reduced_operator_representation.jl:57-116.
 * For entry k and term slot j (< n_terms, in term order) writes the element
 * index (or -1 for an invalid image) and value; unused slots get index -2. */
int edhip_iterate(edhip_op* op, int transpose, int64_t n_rows, const int64_t* rows_host, int64_t* idx_host,
                  double* re_host, double* im_host);

/* sparse(opr; tol) (abstract_operator_representation.jl:136-204, choptol!
 * src/util.jl:88-94): builds the CSC matrix on the device (per column:
 * merge duplicates, drop |v| < tol, rows ascending) and returns nnz;
 * edhip_sparse_export copies it (0-based colptr[dim+1], rowval[nnz]). */
int edhip_sparse_build(edhip_op* op, double tol, int64_t* nnz);
int edhip_sparse_export(edhip_op* op, int64_t* colptr_host, int64_t* rowval_host, double* re_host,
                        double* im_host);

/* get_row / get_column (abstract_operator_representation.jl:217-243) and the
 * column sums behind get_element (operator_representation.jl:109-119,
 * reduced_operator_representation.jl:120-138): for each of the n indices in
 * which_host, the row (transpose = 0, row iterator) or column (transpose = 1,
 * column iterator) of the matrix with duplicates merged, entries with |v| < tol
 * dropped (choptol!, src/util.jl:88-94; tol < 0 keeps every merged entry) and
 * ascending element indices.  Entry k fills counts_host[k] slots starting at
 * k * (n_terms + 1) of idx_host / re_host / im_host (n_terms as passed to
 * edhip_op_create).  EDHIP_ERR_BOUNDS for an index outside [0, dim). */
int edhip_get_vectors(edhip_op* op, int transpose, int64_t n, const int64_t* which_host, double tol,
                      int64_t* counts_host, int64_t* idx_host, double* re_host, double* im_host);

/* Matrix(opr) (abstract_operator_representation.jl:121-132): the dense matrix,
 * column-major (Julia's layout), into the device buffer out of dim * dim
 * doubles (real operators) or dim * dim interleaved (re, im) pairs (complex
 * operators, edhip_op_is_complex).  Dimensions above 2^20 are refused. */
int edhip_matrix(edhip_op* op, void* out_device);

/* ---- eigensolver (the reference's caller uses Arpack.eigs(opr; nev,
 * which=:SR), docs/src/examples/spinhalf.md:26, through mul!) ------------- */
#define EDHIP_LANCZOS_FP64 0  /* FP64 (or ComplexF64) vectors throughout */
#define EDHIP_LANCZOS_MIXED 1 /* FP32 (ComplexF32) Krylov storage + FP64 refinement */

typedef struct {
  int nev;           /* lowest eigenvalues wanted */
  int max_basis;     /* Krylov vectors kept before a thick restart */
  int max_restarts;
  int mode;          /* EDHIP_LANCZOS_FP64 or EDHIP_LANCZOS_MIXED */
  int want_vectors;  /* 1: write the eigenvectors (nev x dim, FP64/C128) */
  int refine_basis;  /* mixed mode: Krylov size of one FP64 refinement cycle */
  int max_refine;    /* mixed mode: refinement cycles per eigenvalue */
  int check_every;   /* Ritz-value convergence check period (iterations);
                        bases over 2^22 elements check every iteration */
  int criterion;     /* 0: residual ||H y - theta y|| <= tol*max(1,|theta|);
                        1: also accept r^2/gap <= tol*max(1,|theta|) with theta
                        stationary between checks (eigenvalue accuracy) */
  int reorth;        /* 0 (auto): 2 on bases over 2^22 elements, else 3;
                        1: two classical Gram-Schmidt passes against the
                        basis at every step (CGS2);
                        2: partial reorthogonalisation (Simon 1984) in the
                        first cycle: the two latest vectors at every step,
                        the whole basis only when the omega recurrence
                        estimates an overlap above sqrt(eps) of the storage
                        precision (that step and the next); every step after
                        a thick restart or an injected direction as 3;
                        3: one pass against the basis at every step, a
                        second one when the first removed more than 99% of
                        ||w||^2.  The FP64 refinement of the mixed mode
                        uses 1 or 3. */
  uint64_t seed;     /* start vector: SplitMix64 + Box-Muller Gaussian */
  double tol;        /* see criterion */
  double tol_fp32;   /* mixed mode: eigenvalue-criterion tolerance of the FP32
                        phase, max(tol, tol_fp32); 0 selects max(tol, 1e-10) */
  int assemble;      /* 0 (auto): with the kernel strategy on auto, a reduced
                        space over 2^22 words gets assembled rows (kernel 3)
                        for the time of the solve when they take at most
                        16 GiB and fit beside the workspace with 4 GiB to
                        spare; 1: never */
} edhip_lanczos_opts;

typedef struct {
  int iterations;     /* Lanczos steps (storage precision) */
  int restarts;
  int matvecs;        /* matvecs in storage precision */
  int matvecs_fp64;   /* FP64 matvecs of the refinement */
  int converged;
  int second_passes;  /* steps whose reorthogonalisation took a second pass */
  double residual[16];     /* final ||H y - theta y|| / max(1,|theta|) (FP64) */
  double unrefined[16];    /* mixed mode: Ritz values before refinement */
  double t_total;          /* seconds from the start vector to the eigenvalues
                              (the workspace allocation excluded) */
  double t_matvec;         /* seconds in storage-precision matvecs */
  double t_refine;         /* seconds in the FP64 refinement */
  double t_alloc;          /* seconds getting the Krylov workspace (0 when a
                              previous call's workspace is reused) */
  double t_ortho;          /* device seconds (HIP events) of the
                              orthogonalisation kernels against the basis,
                              counted for vectors over 2^22 elements */
  double ortho_bytes;      /* their algorithmic bytes: per pass (2k + 3)
                              vectors of n elements for k basis vectors */
  int full_reorth_steps;   /* steps orthogonalised against the whole basis
                              (every step unless reorth is partial) */
  int assembled;           /* 1 if the solve assembled the operator's rows */
} edhip_lanczos_stats;

void edhip_lanczos_default_opts(edhip_lanczos_opts* opts);
/* evals_host: nev doubles, ascending.  evecs: device pointer to nev*dim
 * FP64 (real op) or C128 (complex op) elements, or NULL. */
int edhip_lanczos(edhip_op* op, const edhip_lanczos_opts* opts, double* evals_host, void* evecs,
                  edhip_lanczos_stats* stats);

/* ---- differentiation (no counterpart in the reference; used by the
 * rrule of the ground-state energy and of a correlator) ------------------ */
/* Per-term bilinear forms with unit amplitude in one fused basis pass:
 *   g_t = sum_r conj(phi_r) sum_{images b' of term t from b_r} w_t * psi_s
 * (w_t the representation weight without the term amplitude).  out_host:
 * 2*n_terms doubles (re, im), in the order of the terms given to
 * edhip_op_create.  phi, psi: device vectors of type prec. */
int edhip_term_expect(edhip_op* op, int prec, const void* phi, const void* psi, double* out_host);
/* The same per-term forms with no device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a fixed per-GPU scratch array (about 8.5 MB, held
 * until the process ends), which limits the operator to 512 terms
 * (n_diag + n_off; more: EDHIP_ERR_ARGUMENT); psi is read through the
 * operator's index as built by edhip_op_create (never through a spread of
 * psi, which would be allocated).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams, other threads,
 * other contexts) race. */
int edhip_term_expect_noalloc(edhip_op* op, int prec, const void* phi, const void* psi, double* out_host);

typedef struct {
  int iterations;
  int converged;
  double residual; /* final ||b - A x|| / ||b|| estimate */
  int assembled;   /* 1 if the solve assembled the operator's rows */
} edhip_minres_stats;

/* MINRES (Paige & Saunders) on A = P (H - shift) P, P = 1 - sum_i d_i d_i^H
 * over n_deflate orthonormal device vectors d (may be 0):
 * solves A x = P b.  FP64/C128 vectors (prec EDHIP_F64 or EDHIP_C128).
 * With the kernel strategy on auto, a reduced space over 2^22 words gets
 * assembled rows for the solve as in edhip_lanczos (at most 16 GiB, beside
 * the workspace with 4 GiB to spare); another strategy turns that off. */
int edhip_minres(edhip_op* op, int prec, double shift, int n_deflate, const void* deflate, const void* b, void* x,
                 double tol, int max_iter, edhip_minres_stats* stats);

/* ---- vector helpers on the device (FP64-accumulated, deterministic) ----- */
int edhip_vec_dot(edhip_ctx* ctx, int prec, int64_t n, const void* x, const void* y, double* out_re_im_host);
int edhip_vec_gaussian(edhip_ctx* ctx, int prec, int64_t n, uint64_t seed, void* x);
int edhip_vec_convert(edhip_ctx* ctx, int prec_in, int prec_out, int64_t n, const void* x, void* y);

/* Timing on the context's stream (hipEvent pairs). */
int edhip_event_record(edhip_ctx* ctx, int slot); /* slot 0..15 (14, 15 taken while kernel timing is on) */
int edhip_event_elapsed(edhip_ctx* ctx, int slot0, int slot1, double* ms);
/* Testing (no counterpart in the reference): the next `count` device
 * allocations of the given kinds fail as if the device were out of memory
 * (process-wide; count 0 turns it off).  kinds: bit 0 the spread of an
 * input vector over the word positions (the products then read x through
 * the rank index or the sorted search), bit 1 the eigensolver's Krylov
 * workspace (on a failure the spreads are freed and the allocation tried
 * once more; a second failure returns EDHIP_ERR_MEMORY), bit 2 the entries
 * of an operator's assembled rows (the strategy request returns
 * EDHIP_ERR_MEMORY and keeps nothing). */
#define EDHIP_ALLOC_SPREAD 1
#define EDHIP_ALLOC_WORKSPACE 2
#define EDHIP_ALLOC_ASSEMBLED 4
int edhip_debug_fail_allocs(int kinds, int count);
/* Tuning (no counterpart in the reference): lanes per row of the
 * assembled-row product, 4, 8, 16 or 32, and entries each lane loads
 * together, 4, 8 or 16 (0: the defaults, lanes chosen from the mean entries
 * per row).  Process-wide. */
int edhip_debug_asm_shape(int lanes, int batch);
/* Tuning (no counterpart in the reference): block order of the products
 * (term-list and assembled-row kernels).  Workgroups go to the 8 XCDs in
 * turn; with chunk C > 1 each XCD takes C consecutive blocks of every window
 * of 8 C blocks (0 or 1: plain order; at most 64; default 32).
 * Process-wide; every order gives the same results. */
int edhip_debug_block_order(int chunk);

/* Measurement (no counterpart in the reference): with timing on, every
 * term-list matvec kernel launch and every fused term pass (edhip_term_expect:
 * its kernel and the reduction of its partial sums) of the context is
 * bracketed by events (slots 14 and 15) and waited for, and its duration is
 * added to its operator's total, which edhip_op_kernel_time reports (and
 * clears when reset != 0).  The steps before the kernel (e.g. spreading x)
 * are not included. */
int edhip_set_kernel_timing(edhip_ctx* ctx, int on);
int edhip_op_kernel_time(edhip_op* op, int reset, double* ms_total, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* EDHIP_H */
