// Device-side building blocks shared by the edhip kernels (gfx950, wave64).
//
//  * sector unranking / ranking by completion counts (the ascending order of
//    popcount-constrained words, i.e. the order hs_get_basis_list produces,
//    src/Representation/hilbert_space_representation.jl:127-206);
//  * the sorted-key lookup (FrozenSortedArrayIndex semantics,
//    src/frozensortedarray.jl:29-48) as prefix table + branch-free search;
//  * the two canonicalisers of a word under a symmetry group: a generic one
//    that applies bit permutations through 4-bit lookup tables staged in LDS,
//    and a torus-translation one with top-row pruning.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edhip {

// ----------------------------------------------------------------------------
// Scalars
// ----------------------------------------------------------------------------
template <typename R>
struct Cx {
  R re, im;
};

template <typename V>
struct VT;
template <>
struct VT<double> {
  using R = double;
  static constexpr bool cx = false;
};
template <>
struct VT<float> {
  using R = float;
  static constexpr bool cx = false;
};
template <>
struct VT<Cx<double>> {
  using R = double;
  static constexpr bool cx = true;
};
template <>
struct VT<Cx<float>> {
  using R = float;
  static constexpr bool cx = true;
};

// Block order of the products.  Workgroups are dispatched to the 8 XCDs in
// turn (dispatch b runs on XCD b mod 8), each XCD with its own L2.  With
// chunk C > 1, a window of 8 C dispatches covers 8 C consecutive blocks and
// XCD x takes blocks x C .. x C + C - 1 of it, so that neighbouring rows,
// whose images share lines of x, meet in one L2 while every XCD still works
// inside the same window of rows (which the shared last-level cache serves).
// A bijection on [0, n_blocks); blocks of a last partial window keep their order.
__device__ __forceinline__ int64_t xcd_chunked_block(int64_t b, int64_t n_blocks, int C) {
  const int64_t w = 8 * (int64_t)C;
  if (C <= 1 || b >= (n_blocks / w) * w) return b;
  const int64_t r = b % w;
  return b - r + (r & 7) * C + (r >> 3);
}

template <typename V>
__device__ __forceinline__ V vzero() {
  if constexpr (VT<V>::cx) {
    return V{0, 0};
  } else {
    return V(0);
  }
}

// acc += w * x, w a complex weight given as (wr, wi)
template <typename V, typename R>
__device__ __forceinline__ void vfma(V& acc, R wr, R wi, const V& x) {
  if constexpr (VT<V>::cx) {
    acc.re += wr * x.re - wi * x.im;
    acc.im += wr * x.im + wi * x.re;
  } else {
    (void)wi;
    acc += wr * x;
  }
}

template <typename V>
__device__ __forceinline__ V vadd(const V& a, const V& b) {
  if constexpr (VT<V>::cx) {
    return V{a.re + b.re, a.im + b.im};
  } else {
    return a + b;
  }
}

// ----------------------------------------------------------------------------
// Sector counting: W[i*(n+1)+ones] = number of fillings of bits [0,i) such
// that ones + popcount(filling) is an allowed popcount.
// ----------------------------------------------------------------------------
__device__ __forceinline__ uint64_t unrank_word(uint64_t rank, int n, const uint64_t* W) {
  uint64_t w = 0;
  int ones = 0;
  for (int i = n - 1; i >= 0; --i) {
    uint64_t c0 = W[i * (n + 1) + ones];
    if (rank >= c0) {
      rank -= c0;
      w |= (1ull << i);
      ++ones;
    }
  }
  return w;
}

// Rank of w among the allowed words (w must be an allowed word).
__device__ __forceinline__ uint64_t rank_word(uint64_t w, int n, const uint64_t* W) {
  uint64_t r = 0;
  int ones = 0;
  while (w) {
    int i = 63 - __builtin_clzll(w);
    r += W[i * (n + 1) + ones];
    ++ones;
    w &= ~(1ull << i);
  }
  return r;
}

// Next larger word of the same popcount (Gosper); v != 0.
__device__ __forceinline__ uint64_t next_same_popcount(uint64_t v) {
  uint64_t t = v | (v - 1);
  return (t + 1) | (((~t & (t + 1)) - 1) >> (__builtin_ctzll(v) + 1));
}

// ----------------------------------------------------------------------------
// Lookup in a sorted key array behind a prefix table.
//   bucket(key) = key >> shift; keys with that bucket live in
//   [prefix[bucket], prefix[bucket+1]).  Inside, a branch-free binary search.
// Returns the 0-based index or -1.
// ----------------------------------------------------------------------------
struct IndexDev {
  const uint64_t* keys;
  const uint32_t* prefix;
  int64_t dim;
  uint64_t n_buckets;
  int shift;
  int rank_mode;  // 1: the space is a full popcount sector; index == rank
  int n_bits;
  int pad_;
  uint64_t pop_allowed;  // bit p set iff popcount p is allowed (rank mode)
  const uint64_t* W;     // completion counts (rank mode)
  // blocked rank index (bm_bits > 0) over the words [0, 2^bm_bits): block j
  // (64 bytes) covers words [384 j, 384 j + 384): u64 0..5 hold their
  // membership bits, u64 6 the number of basis words below 384 j, u64 7 the
  // bit counts of u64 0..w-1 for w = 1..5 (9 bits each).  One cache line per
  // lookup.
  const uint64_t* bitmap;
  const uint32_t* bm_counts;  // unused
  int bm_bits;
  int pad2_;
};

constexpr uint32_t kRankBlockWords = 384;  // membership bits per 64-byte block

// Index of key through the blocked rank index (membership and rank from one
// 64-byte block).
__device__ __forceinline__ int64_t bitmap_index(const IndexDev& ix, uint64_t key) {
  if ((key >> ix.bm_bits) != 0) return -1;
  const uint32_t q = (uint32_t)(key >> 7);         // < 2^29
  const uint32_t j = __umulhi(q, 0xAAAAAAABu) >> 1;  // q / 3 = key / 384
  const uint32_t within = (uint32_t)(key - (uint64_t)j * kRankBlockWords);
  const uint32_t w = within >> 6, bit = within & 63u;
  const uint64_t* blk = ix.bitmap + (uint64_t)j * 8u;
  // the three reads hit one cache line; issue them together
  const uint64_t word = blk[w];
  const ulonglong2 tail = *(const ulonglong2*)(blk + 6);  // base rank, packed sub-counts
  const uint32_t sub = w ? (uint32_t)(tail.y >> (9u * (w - 1u))) & 511u : 0u;
  const int64_t r = (int64_t)(tail.x + sub + (uint64_t)__builtin_popcountll(word & ((1ull << bit) - 1ull)));
  return ((word >> bit) & 1ull) ? r : -1;
}

__device__ __forceinline__ int64_t search_sorted(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ prefix,
                                                 uint64_t n_buckets, int shift, uint64_t key) {
  uint64_t bucket = key >> shift;
  if (bucket >= n_buckets) return -1;
  int64_t base = prefix[bucket];
  int64_t len = (int64_t)prefix[bucket + 1] - base;
  if (len <= 0) return -1;
  while (len > 1) {
    int64_t half = len >> 1;
    base = (keys[base + half] <= key) ? base + half : base;
    len -= half;
  }
  return (keys[base] == key) ? base : -1;
}

__device__ __forceinline__ int64_t index_of(const IndexDev& ix, const uint64_t* W_lds, uint64_t key) {
  if (ix.bm_bits) return bitmap_index(ix, key);
  if (ix.rank_mode) {
    if ((key >> ix.n_bits) != 0) return -1;
    int pc = __builtin_popcountll(key);
    if (!((ix.pop_allowed >> pc) & 1ull)) return -1;
    return (int64_t)rank_word(key, ix.n_bits, W_lds);
  }
  return search_sorted(ix.keys, ix.prefix, ix.n_buckets, ix.shift, key);
}

// Two lookups with their loads issued together.
__device__ __forceinline__ void index_of2(const IndexDev& ix, const uint64_t* W_lds, uint64_t k1, uint64_t k2,
                                          int64_t& i1, int64_t& i2) {
  if (ix.bm_bits) {
    const bool in1 = (k1 >> ix.bm_bits) == 0, in2 = (k2 >> ix.bm_bits) == 0;
    const uint64_t a1 = in1 ? k1 : 0, a2 = in2 ? k2 : 0;
    const uint32_t j1 = __umulhi((uint32_t)(a1 >> 7), 0xAAAAAAABu) >> 1;
    const uint32_t j2 = __umulhi((uint32_t)(a2 >> 7), 0xAAAAAAABu) >> 1;
    const uint32_t w1 = (uint32_t)(a1 - (uint64_t)j1 * kRankBlockWords);
    const uint32_t w2 = (uint32_t)(a2 - (uint64_t)j2 * kRankBlockWords);
    const uint64_t* b1 = ix.bitmap + (uint64_t)j1 * 8u;
    const uint64_t* b2 = ix.bitmap + (uint64_t)j2 * 8u;
    const uint64_t word1 = b1[w1 >> 6], word2 = b2[w2 >> 6];
    const ulonglong2 t1 = *(const ulonglong2*)(b1 + 6);
    const ulonglong2 t2 = *(const ulonglong2*)(b2 + 6);
    const uint32_t s1 = (w1 >> 6) ? (uint32_t)(t1.y >> (9u * ((w1 >> 6) - 1u))) & 511u : 0u;
    const uint32_t s2 = (w2 >> 6) ? (uint32_t)(t2.y >> (9u * ((w2 >> 6) - 1u))) & 511u : 0u;
    const uint32_t bit1 = w1 & 63u, bit2 = w2 & 63u;
    const int64_t r1 = (int64_t)(t1.x + s1 + (uint64_t)__builtin_popcountll(word1 & ((1ull << bit1) - 1ull)));
    const int64_t r2 = (int64_t)(t2.x + s2 + (uint64_t)__builtin_popcountll(word2 & ((1ull << bit2) - 1ull)));
    i1 = (in1 && ((word1 >> bit1) & 1ull)) ? r1 : -1;
    i2 = (in2 && ((word2 >> bit2) & 1ull)) ? r2 : -1;
    return;
  }
  i1 = index_of(ix, W_lds, k1);
  i2 = index_of(ix, W_lds, k2);
}

// ----------------------------------------------------------------------------
// Symmetry group description on the device.
// ----------------------------------------------------------------------------
enum GroupMode : int { kGroupNone = 0, kGroupGeneric = 1, kGroupTorus = 2 };
// kernel specialisations of kGroupTorus for groups with the row-pair table:
// any shape, and the common shapes fixed at compile time
constexpr int kGroupTorusPairs = 3;
constexpr int kTorusPairs6x6 = 4;
constexpr int kTorusPairs4x6 = 5;
constexpr int kTorusPairs6x4 = 6;
constexpr int kTorusPairs4x4 = 7;
// the same shapes for the translations alone (one coset): kernels without
// the coset code, which the larger space-group kernels carry
constexpr int kTransPairs6x6 = 8;
constexpr int kTransPairs4x6 = 9;
constexpr int kTransPairs6x4 = 10;
constexpr int kTransPairs4x4 = 11;

constexpr int kMaxLy = 64;

struct GroupDev {
  int mode;
  int G;
  int n_nib;  // generic: number of 4-bit chunks of the word
  int Lx, Ly, nb;
  int minrot_table;     // torus: 1 if the per-row minimal-rotation table is used
  int pair_table;       // torus: 1 if the adjacent-row-pair table is used (Lx <= 6, 2 <= Ly <= 8)
  int unit_phases;      // 1 if every phase is exactly 1 (all characters 1)
  int has_xor;          // some element XORs a mask into the permuted word (global spin flip)
  int n_cosets;         // torus: G / (Lx Ly) cosets T p of the translations (1: translations only)
  int nib_global;       // generic: the permutation tables are read from nib (global), not LDS
  // torus cosets in kernel arguments (coset_regs = 1: at most 16 cosets, 16
  // permutations and one XOR word): 2 bits of kind and 4 bits of permutation
  // per coset, bit p of cx_bits when coset p XORs cxword
  int coset_regs;
  uint32_t ck_bits;
  uint64_t cp_bits;
  uint32_t cx_bits;
  uint32_t pad_cx_;
  uint64_t cxword;
  uint64_t row_mask;    // torus: (1 << Lx) - 1
  uint64_t full_mask;   // (1 << nb) - 1
  const uint64_t* nib;  // generic: G * n_nib * 16 images of 4-bit chunks
  const uint32_t* minrot;  // torus: 2^Lx entries, low 16: min rotation value, high 16: mask of rotations
  // torus: 2^(2 Lx) entries indexed by (hi << Lx) | lo for two adjacent rows
  // (hi above lo): min_a (rot_a(hi) << Lx | rot_a(lo)) << 8 | mask of the a reaching it
  const uint32_t* pair;
  const uint64_t* rot_hi;  // torus: Lx entries, per-row bits kept by a left rotation by a
  const uint64_t* rot_lo;  // torus: Lx entries, per-row bits wrapped around
  const int32_t* g_of_ab;  // torus: Lx*Ly -> group index of T(a, b)
  const int32_t* inv;      // G: index of the inverse element
  const uint64_t* xr;      // G: word XORed after the permutation (has_xor)
  // torus cosets: every element is T(pos) p for a translation T(pos), pos =
  // a + Lx b, and a coset representative p (p = 0: the identity)
  const uint64_t* coset_nib;   // distinct permutations of the representatives, n_nib x 16 each
  const int32_t* coset_pidx;   // n_cosets: which of them p uses (kind 2)
  const uint64_t* coset_xr;    // n_cosets: word XORed after the permutation
  const int32_t* coset_kind;   // n_cosets: 0 identity, 1 XOR only, 2 permutation (+ XOR)
  const int32_t* coset_elem;   // n_cosets x (Lx Ly): element index of T(pos) p
  const int32_t* coset_inv;    // n_cosets x (Lx Ly): inv[coset_elem]
  const uint8_t* trivial;  // G: |chi(g) - 1| <= tol
  const double* ph_re;     // G: phase[h] = conj(chi(inv(h)))
  const double* ph_im;
  const double* invsqrt;   // G+1: 1/sqrt(N)
  const double* sqcnt;     // G+1: sqrt(cnt/G) = 1/sqrt(G/cnt) for a stabiliser of cnt elements
};

// Pointers to the LDS copies of the group tables (same meaning as GroupDev).
struct GroupLds {
  const double* ph_re;
  const double* ph_im;
  const double* sqcnt;
  const uint64_t* nib;
  const uint32_t* minrot;
  const uint32_t* pair;
  const uint64_t* rot_hi;
  const uint64_t* rot_lo;
  const int32_t* g_of_ab;
  const int32_t* inv;
  const uint8_t* trivial;
  const uint64_t* xr;
  const uint64_t* coset_nib;
  const int32_t* coset_pidx;
  const uint64_t* coset_xr;
  const int32_t* coset_kind;
  const int32_t* coset_elem;
  const int32_t* coset_inv;
};

// Image of b under a bit permutation given as n_nib tables of 16 entries (the
// image of each value of 4-bit chunk k at tab[16 k + v]).  The reads go out
// four at a time, so that a chunk's read does not wait for the previous one;
// in the last, partial group the chunks past n_nib read entry 0, the image of
// no bits (0).
__device__ __forceinline__ uint64_t perm_apply_nib(const uint64_t* tab, int n_nib, uint64_t b) {
  uint64_t out = 0;
  int k0 = 0;
  for (; k0 + 4 <= n_nib; k0 += 4) {
    uint64_t v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = tab[((k0 + q) << 4) + (int)((b >> (4 * (k0 + q))) & 15ull)];
    out |= (v[0] | v[1]) | (v[2] | v[3]);
  }
  if (k0 < n_nib) {  // 1 to 3 chunks left
    uint64_t v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int k = k0 + q;
      v[q] = tab[k < n_nib ? (k << 4) + (int)((b >> ((4 * k) & 63)) & 15ull) : 0];
    }
    out |= v[0] | v[1] | v[2];
  }
  return out;
}

__device__ __forceinline__ uint64_t rows_rotl(uint64_t w, int a, int Lx, const uint64_t* hi, const uint64_t* lo) {
  if (a == 0) return w;
  return ((w << a) & hi[a]) | ((w >> (Lx - a)) & lo[a]);
}

__device__ __forceinline__ uint64_t rows_shift(uint64_t w, int b, int Lx, int nb, uint64_t full) {
  if (b == 0) return w;
  int s = b * Lx;
  return ((w << s) | (w >> (nb - s))) & full;
}

// minimal left rotation of an Lx-bit row value and the mask of rotations reaching it
__device__ __forceinline__ void row_minrot(const GroupDev& g, const GroupLds& L, uint64_t v, uint64_t& mv, uint64_t& amask) {
  if (g.minrot_table) {
    uint32_t e = L.minrot[(int)v];
    mv = e & 0xffffu;
    amask = e >> 16;
    return;
  }
  mv = v;
  amask = 1;
  for (int a = 1; a < g.Lx; ++a) {
    uint64_t r = ((v << a) | (v >> (g.Lx - a))) & g.row_mask;
    if (r < mv) {
      mv = r;
      amask = 1ull << a;
    } else if (r == mv) {
      amask |= 1ull << a;
    }
  }
}

// Torus canonicaliser over the adjacent-row-pair table (g.pair_table set).
// The smallest image has the smallest (top row, second row) pair, so only the
// (row y, rotation a) whose pair (rot_a(row_y), rot_a(row_{y-1})) reaches
// that minimum can give it; they are collected in one mask (8 bits per row:
// bit 8y + a) and walked in a single loop, so lanes of a wave stay together
// unless one of them has several such candidates.  Same s, h and cnt as the
// generic torus path below.
// LXC, LYC > 0 fix the torus shape at compile time; 0 reads it from g.
// Translation part.  Returns the smallest translation image s of b, the
// position pos = a + Lx * bb of a translation T(pos) reaching it (among ties
// the one with the largest invt[pos]) and cnt = #{T : T(b) = s}.
// invt: inverse indices of the elements T(pos) p of the coset being scanned.
template <int LXC = 0, int LYC = 0, bool NEED_H = true>
__device__ __forceinline__ void canonicalize_torus_pairs_t(const GroupDev& g, const GroupLds& L, const int32_t* invt,
                                                           uint64_t b, uint64_t& s, int& pos, int& cnt) {
  const int Lx = LXC ? LXC : g.Lx;
  const int Ly = LYC ? LYC : g.Ly;
  const int NB = LXC ? LXC * LYC : g.nb;
  const uint64_t FULL = LXC ? ((1ull << (LXC * LYC)) - 1ull) : g.full_mask;
  const uint32_t m2 = (1u << (2 * Lx)) - 1u;
  // row Ly-1 copied below row 0: the window at y*Lx holds (row_y << Lx) | row_{y-1}
  const uint64_t w = (b << Lx) | (b >> (NB - Lx));
  uint32_t e[8];
  uint32_t lo = 0xffffffffu;  // the smallest entry: its value part is the smallest pair
#pragma unroll
  for (int y = 0; y < 8; ++y) {
    e[y] = 0xffffffffu;
    if (y < Ly) {
      e[y] = L.pair[(uint32_t)(w >> (y * Lx)) & m2];
      lo = min(lo, e[y]);
    }
  }
  // an entry of the smallest value, less that value, is its rotation mask
  // (below 256); every other entry is at least 256 above: min(e - base, 256)
  // leaves the mask or 0 in the low byte, and the low bytes of the eight rows
  // are packed into the candidate mask (v_perm_b32, selectors 0x0c: zero)
  const uint32_t base = lo & ~0xffu;
  uint32_t m[8];
#pragma unroll
  for (int y = 0; y < 8; ++y) m[y] = min(e[y] - base, 256u);
  const uint32_t cand_lo = __builtin_amdgcn_perm(m[1], m[0], 0x0c0c0400u) |
                           (__builtin_amdgcn_perm(m[3], m[2], 0x0c0c0400u) << 16);
  const uint32_t cand_hi = __builtin_amdgcn_perm(m[5], m[4], 0x0c0c0400u) |
                           ((LYC > 0 && LYC <= 6) ? 0u : (__builtin_amdgcn_perm(m[7], m[6], 0x0c0c0400u) << 16));
  uint64_t cand = ((uint64_t)cand_hi << 32) | cand_lo;
  s = ~0ull;
  cnt = 0;
  int win = 0;  // bit 8y + a of the candidate giving s
  const uint32_t rmask = (1u << Lx) - 1u;
  bool first = true;
  while (cand) {
    const int bit = __builtin_ctzll(cand);
    cand &= cand - 1;
    const int a = bit & 7;
    // Every candidate shares the top two rows of its image; after the first,
    // one whose third row (row y - 2, rotated by a) exceeds that of s cannot
    // reach s and is skipped before its full image is formed (most lanes
    // have a single candidate: the loop runs as long as the lane with the
    // most candidates, and the skipped ones cost a few operations).
    bool full = true;
    if (Ly >= 3 && !first) {
      const int y3 = ((bit >> 3) + Ly - 2) % Ly;
      // row offsets as 24-bit products (v_mul_u32_u24, full rate; a 32-bit
      // v_mul_lo_u32 issues at a quarter of the rate)
      const uint32_t r3 = (uint32_t)(b >> __umul24((uint32_t)y3, (uint32_t)Lx)) & rmask;
      const uint32_t t3 = ((r3 << a) | (r3 >> (Lx - a))) & rmask;
      full = t3 <= ((uint32_t)(s >> ((Ly - 3) * Lx)) & rmask);
    }
    first = false;
    if (!full) continue;
    const int sh = (int)__umul24((uint32_t)(Ly - 1 - (bit >> 3)), (uint32_t)Lx);  // row y moves to the top
    const uint64_t t = ((b << sh) | (b >> (NB - sh))) & FULL;
    const uint64_t img = ((t << a) & L.rot_hi[a]) | ((t >> (Lx - a)) & L.rot_lo[a]);
    const bool lt = img < s, eq = img == s;
    if (NEED_H && eq) {  // a tie (s has a nontrivial stabiliser): keep the largest inverse index
      const int p_new = a + Lx * (Ly - 1 - (bit >> 3));
      const int p_cur = (win & 7) + Lx * (Ly - 1 - (win >> 3));
      if (invt[p_new] > invt[p_cur]) win = bit;
    }
    cnt = lt ? 1 : cnt + (eq ? 1 : 0);
    if (NEED_H) win = lt ? bit : win;
    s = lt ? img : s;
  }
  pos = NEED_H ? (win & 7) + Lx * (Ly - 1 - (win >> 3)) : 0;
}


// Translation part without the row-pair table (same outputs).
__device__ __forceinline__ void canonicalize_torus_rows(const GroupDev& g, const GroupLds& L, const int32_t* invt,
                                                        uint64_t b, uint64_t& s, int& pos, int& cnt) {
  // pass 1: the smallest possible top row
  uint64_t best_top = ~0ull;
  for (int y = 0; y < g.Ly; ++y) {
    uint64_t mv, am;
    row_minrot(g, L, (b >> (y * g.Lx)) & g.row_mask, mv, am);
    best_top = mv < best_top ? mv : best_top;
  }
  // pass 2: full images of the candidates reaching it
  s = ~0ull;
  pos = 0;
  cnt = 0;
  for (int y = 0; y < g.Ly; ++y) {
    uint64_t mv, am;
    row_minrot(g, L, (b >> (y * g.Lx)) & g.row_mask, mv, am);
    if (mv != best_top) continue;
    int bb = g.Ly - 1 - y;  // this row lands on top under T(a, bb)
    uint64_t shifted = rows_shift(b, bb, g.Lx, g.nb, g.full_mask);
    while (am) {
      int a = __builtin_ctzll(am);
      am &= am - 1;
      uint64_t img = rows_rotl(shifted, a, g.Lx, L.rot_hi, L.rot_lo);
      const int q = a + g.Lx * bb;
      if (img < s) {
        s = img;
        pos = q;
        cnt = 1;
      } else if (img == s) {
        ++cnt;
        if (invt[q] > invt[pos]) pos = q;
      }
    }
  }
}

// perm_apply_nib with the number of chunks fixed: all NNIB table reads are
// issued before the first is used (the loop form waits for each in turn).
template <int NNIB>
__device__ __forceinline__ uint64_t perm_apply_nib_t(const uint64_t* tab, uint64_t b) {
  uint64_t out = 0;
#pragma unroll
  for (int k = 0; k < NNIB; ++k) out |= tab[(k << 4) + (int)((b >> (4 * k)) & 15ull)];
  return out;
}

// p(b) for the coset representative p.  NB > 0: words of NB bits, known at
// compile time (the chunk count g.n_nib is then (NB + 3) / 4).  UNIFORM: p is
// the same in every lane of the wave, so its kind and table are read once
// into scalar registers and the branches on them are taken by the whole wave.
template <int NB = 0, bool UNIFORM = false>
__device__ __forceinline__ uint64_t apply_coset(const GroupDev& g, const GroupLds& L, int p, uint64_t b) {
  if (g.coset_regs) {  // kind, permutation and XOR flag from the kernel arguments
    const int kind = (int)((g.ck_bits >> (2 * p)) & 3u);
    uint64_t w = b;
    if (kind == 2) {
      const uint64_t* tab = L.coset_nib + (int)((g.cp_bits >> (4 * p)) & 15u) * (g.n_nib << 4);
      if constexpr (NB > 0)
        w = perm_apply_nib_t<(NB + 3) / 4>(tab, b);
      else
        w = perm_apply_nib(tab, g.n_nib, b);
    }
    return w ^ (((g.cx_bits >> p) & 1u) ? g.cxword : 0ull);
  }
  int kind = L.coset_kind[p];
  if (UNIFORM) kind = __builtin_amdgcn_readfirstlane(kind);
  if (kind == 0) return b;
  uint64_t w = b;
  if (kind == 2) {
    int pidx = L.coset_pidx[p];
    if (UNIFORM) pidx = __builtin_amdgcn_readfirstlane(pidx);
    const uint64_t* tab = L.coset_nib + pidx * (g.n_nib << 4);
    if constexpr (NB > 0)
      w = perm_apply_nib_t<(NB + 3) / 4>(tab, b);
    else
      w = perm_apply_nib(tab, g.n_nib, b);
  }
  return w ^ L.coset_xr[p];
}

// Smallest (top row, second row) pair over the translation images of w, as
// the row-pair table's value part (entry | 0xff).
template <int LXC = 0, int LYC = 0>
__device__ __forceinline__ uint32_t torus_pair_min(const GroupDev& g, const GroupLds& L, uint64_t w) {
  const int Lx = LXC ? LXC : g.Lx;
  const int Ly = LYC ? LYC : g.Ly;
  const int NB = LXC ? LXC * LYC : g.nb;
  const uint32_t m2 = (1u << (2 * Lx)) - 1u;
  const uint64_t ww = (w << Lx) | (w >> (NB - Lx));
  uint32_t best = 0xffffffffu;
#pragma unroll
  for (int y = 0; y < 8; ++y)
    if (y < Ly) best = min(best, L.pair[(uint32_t)(ww >> (y * Lx)) & m2] | 0xffu);
  return best;
}

// Torus group: min over the cosets T p of the translation-canonical form of
// p(b); s, h (largest inverse index among ties) and cnt = |Stab(s)|.  With
// the row-pair table and several cosets, a first pass finds the cosets whose
// images reach the smallest (top, second) row pair; only those are scanned.
template <int LXC = 0, int LYC = 0, bool PAIRS = true, bool NEED_H = true>
__device__ __forceinline__ void canonicalize_torus_cosets(const GroupDev& g, const GroupLds& L, uint64_t b,
                                                          uint64_t& s, int& h, int& cnt) {
  const int nT = (LXC ? LXC : g.Lx) * (LYC ? LYC : g.Ly);
  constexpr int NBC = LXC * LYC;  // word bits when the shape is fixed, else 0
  if (PAIRS && g.n_cosets > 1 && g.n_cosets <= 64) {
    uint32_t best = 0xffffffffu;
    uint64_t sel = 0;  // cosets reaching best
    // p(b) of the last permutation applied: cosets that share a permutation
    // (a representative and its products with XOR words) follow each other
    int last = -1;
    uint64_t pb = 0;
    for (int p = 0; p < g.n_cosets; ++p) {
      // the coset's kind and permutation: from the kernel arguments, or read
      // once per wave from LDS
      const int kind = g.coset_regs ? (int)((g.ck_bits >> (2 * p)) & 3u)
                                    : __builtin_amdgcn_readfirstlane(L.coset_kind[p]);
      uint64_t w = b;
      if (kind == 2) {
        const int pidx = g.coset_regs ? (int)((g.cp_bits >> (4 * p)) & 15u)
                                      : __builtin_amdgcn_readfirstlane(L.coset_pidx[p]);
        if (pidx != last) {
          const uint64_t* tab = L.coset_nib + pidx * (g.n_nib << 4);
          if constexpr (NBC > 0)
            pb = perm_apply_nib_t<(NBC + 3) / 4>(tab, b);
          else
            pb = perm_apply_nib(tab, g.n_nib, b);
          last = pidx;
        }
        w = pb;
      }
      if (g.coset_regs)
        w ^= ((g.cx_bits >> p) & 1u) ? g.cxword : 0ull;
      else if (kind != 0)
        w ^= L.coset_xr[p];
      const uint32_t v = torus_pair_min<LXC, LYC>(g, L, w);
      sel = v < best ? (1ull << p) : (v == best ? sel | (1ull << p) : sel);
      best = v < best ? v : best;
    }
    s = ~0ull;
    h = 0;
    cnt = 0;
    while (sel) {
      const int p = __builtin_ctzll(sel);
      sel &= sel - 1;
      uint64_t s2;
      int pos2, c2;
      canonicalize_torus_pairs_t<LXC, LYC, NEED_H>(g, L, L.coset_inv + p * nT, apply_coset<NBC>(g, L, p, b), s2, pos2, c2);
      const int h2 = NEED_H ? L.coset_elem[p * nT + pos2] : 0;
      if (s2 < s) {
        s = s2;
        h = h2;
        cnt = c2;
      } else if (s2 == s) {
        cnt += c2;
        if (NEED_H && L.inv[h2] > L.inv[h]) h = h2;
      }
    }
    return;
  }
  int pos;
  if constexpr (PAIRS)
    canonicalize_torus_pairs_t<LXC, LYC, NEED_H>(g, L, L.coset_inv, b, s, pos, cnt);
  else
    canonicalize_torus_rows(g, L, L.coset_inv, b, s, pos, cnt);
  h = NEED_H ? L.coset_elem[pos] : 0;
  for (int p = 1; p < g.n_cosets; ++p) {
    const uint64_t w = apply_coset<NBC, true>(g, L, p, b);
    uint64_t s2;
    int pos2, c2;
    if constexpr (PAIRS)
      canonicalize_torus_pairs_t<LXC, LYC, NEED_H>(g, L, L.coset_inv + p * nT, w, s2, pos2, c2);
    else
      canonicalize_torus_rows(g, L, L.coset_inv + p * nT, w, s2, pos2, c2);
    const int h2 = NEED_H ? L.coset_elem[p * nT + pos2] : 0;
    if (s2 < s) {
      s = s2;
      h = h2;
      cnt = c2;
    } else if (s2 == s) {
      cnt += c2;
      if (NEED_H && L.inv[h2] > L.inv[h]) h = h2;
    }
  }
}

// NEED_H = false: h is not tracked (returned as 0), for callers that only use
// s and cnt (every character 1).
template <int LXC = 0, int LYC = 0, bool NEED_H = true>
__device__ __forceinline__ void canonicalize_torus_pairs(const GroupDev& g, const GroupLds& L, uint64_t b, uint64_t& s,
                                                         int& h, int& cnt) {
  canonicalize_torus_cosets<LXC, LYC, true, NEED_H>(g, L, b, s, h, cnt);
}

// Canonical form s = min_h h(b), an element h with h(b) = s (among ties, the
// one whose inverse has the largest index: the reference's last writer of the
// orbit dictionary), and cnt = #{h : h(b) = s} = |Stab(s)|.
// Calls f(tables) with the generic group's 4-bit permutation tables: the LDS
// copy, or the global array for groups too large for LDS (two separate
// inlined calls, so that each reads through its own kind of load).
template <typename F>
__device__ __forceinline__ void with_nib_tables(const GroupDev& g, const GroupLds& L, F&& f) {
  if (g.nib_global)
    f(g.nib);
  else
    f(L.nib);
}

__device__ __forceinline__ void canonicalize(const GroupDev& g, const GroupLds& L, uint64_t b, uint64_t& s, int& h,
                                             int& cnt) {
  if (g.mode == kGroupGeneric) {
    s = b;
    h = 0;
    cnt = 1;
    const int stride = g.n_nib << 4;
    with_nib_tables(g, L, [&](const uint64_t* nib) {
      for (int e = 1; e < g.G; ++e) {
        uint64_t img = perm_apply_nib(nib + e * stride, g.n_nib, b);
        if (g.has_xor) img ^= L.xr[e];
        if (img < s) {
          s = img;
          h = e;
          cnt = 1;
        } else if (img == s) {
          ++cnt;
          if (L.inv[e] > L.inv[h]) h = e;
        }
      }
    });
    return;
  }
  if (g.mode == kGroupTorus) {
    if (g.pair_table)
      canonicalize_torus_cosets<0, 0, true>(g, L, b, s, h, cnt);
    else
      canonicalize_torus_cosets<0, 0, false>(g, L, b, s, h, cnt);
    return;
  }
  s = b;
  h = 0;
  cnt = 1;
}

// Representative test of symmetry_reduce (symmetry_reduce_generic.jl:56-69):
// b is kept iff no image is smaller and every element fixing b has a trivial
// character.  On success orbit = G / |Stab(b)|.
// Translation images of w against a candidate representative b: false if
// one is smaller than b; for each image T(pos)(w) equal to b, the element
// elem[pos] must have a trivial character, and stab counts it.
__device__ __forceinline__ bool torus_scan(const GroupDev& g, const GroupLds& L, uint64_t w, uint64_t b,
                                           const int32_t* elem, int& stab) {
  const uint64_t top = (b >> ((g.Ly - 1) * g.Lx)) & g.row_mask;
  uint64_t best_top = ~0ull;
  for (int y = 0; y < g.Ly; ++y) {
    uint64_t mv, am;
    row_minrot(g, L, (w >> (y * g.Lx)) & g.row_mask, mv, am);
    best_top = mv < best_top ? mv : best_top;
  }
  if (best_top < top) return false;
  if (best_top > top) return true;  // every image is larger than b
  for (int y = 0; y < g.Ly; ++y) {
    uint64_t mv, am;
    row_minrot(g, L, (w >> (y * g.Lx)) & g.row_mask, mv, am);
    if (mv != best_top) continue;
    int bb = g.Ly - 1 - y;
    uint64_t shifted = rows_shift(w, bb, g.Lx, g.nb, g.full_mask);
    while (am) {
      int a = __builtin_ctzll(am);
      am &= am - 1;
      uint64_t img = rows_rotl(shifted, a, g.Lx, L.rot_hi, L.rot_lo);
      if (img < b) return false;
      if (img == b) {
        if (!L.trivial[elem[a + g.Lx * bb]]) return false;
        ++stab;
      }
    }
  }
  return true;
}

// Representative test of symmetry_reduce (symmetry_reduce_generic.jl:56-69):
// b is kept iff no image is smaller and every element fixing b has a trivial
// character.  On success orbit = G / |Stab(b)|.
__device__ __forceinline__ bool is_representative(const GroupDev& g, const GroupLds& L, uint64_t b, int& orbit) {
  if (g.mode == kGroupGeneric) {
    int stab = 1;
    bool ok = true;
    const int stride = g.n_nib << 4;
    with_nib_tables(g, L, [&](const uint64_t* nib) {
      for (int e = 1; e < g.G; ++e) {
        uint64_t img = perm_apply_nib(nib + e * stride, g.n_nib, b);
        if (g.has_xor) img ^= L.xr[e];
        if (img < b || (img == b && !L.trivial[e])) {
          ok = false;
          return;
        }
        if (img == b) ++stab;
      }
    });
    if (!ok) return false;
    orbit = g.G / stab;
    return true;
  }
  if (g.mode == kGroupTorus) {
    int stab = 0;
    const int nT = g.Lx * g.Ly;
    for (int p = 0; p < g.n_cosets; ++p)
      if (!torus_scan(g, L, apply_coset(g, L, p, b), b, L.coset_elem + p * nT, stab)) return false;
    orbit = g.G / stab;
    return true;
  }
  orbit = 1;
  return true;
}

// ----------------------------------------------------------------------------
// LDS staging of the group tables.  The host computes the byte offsets.
// ----------------------------------------------------------------------------
struct GroupLdsLayout {
  uint32_t off_nib, off_minrot, off_pair, off_rot_hi, off_rot_lo, off_g_of_ab, off_inv, off_trivial;
  uint32_t off_ph_re, off_ph_im, off_invsqrt, off_sqcnt, off_xr, off_cnib, off_cpidx, off_cxr, off_ckind, off_celem,
      off_cinv;
  uint32_t bytes_nib, bytes_minrot, bytes_pair, bytes_rot, bytes_g_of_ab, bytes_inv, bytes_trivial, bytes_xr;
  uint32_t bytes_cnib, bytes_cxr, bytes_ckind, bytes_celem;
  uint32_t bytes_ph, bytes_invsqrt;
  uint32_t total;
};

__device__ __forceinline__ void lds_copy(unsigned char* dst, const void* src, uint32_t bytes) {
  if (!src || bytes == 0) return;
  const uint32_t n4 = bytes >> 2;  // all staged tables are multiples of 4 bytes
  const uint32_t* s = (const uint32_t*)src;
  uint32_t* d = (uint32_t*)dst;
  for (uint32_t i = threadIdx.x; i < n4; i += blockDim.x) d[i] = s[i];
}

__device__ __forceinline__ GroupLds stage_group(const GroupDev& g, const GroupLdsLayout& lay, unsigned char* lds) {
  GroupLds L;
  lds_copy(lds + lay.off_nib, g.nib, lay.bytes_nib);
  lds_copy(lds + lay.off_minrot, g.minrot, lay.bytes_minrot);
  lds_copy(lds + lay.off_pair, g.pair, lay.bytes_pair);
  lds_copy(lds + lay.off_rot_hi, g.rot_hi, lay.bytes_rot);
  lds_copy(lds + lay.off_rot_lo, g.rot_lo, lay.bytes_rot);
  lds_copy(lds + lay.off_g_of_ab, g.g_of_ab, lay.bytes_g_of_ab);
  lds_copy(lds + lay.off_inv, g.inv, lay.bytes_inv);
  lds_copy(lds + lay.off_trivial, g.trivial, lay.bytes_trivial);
  lds_copy(lds + lay.off_ph_re, g.ph_re, lay.bytes_ph);
  lds_copy(lds + lay.off_ph_im, g.ph_im, lay.bytes_ph);
  lds_copy(lds + lay.off_sqcnt, g.sqcnt, lay.bytes_invsqrt);
  lds_copy(lds + lay.off_xr, g.xr, lay.bytes_xr);
  lds_copy(lds + lay.off_cnib, g.coset_nib, lay.bytes_cnib);
  lds_copy(lds + lay.off_cpidx, g.coset_pidx, lay.bytes_ckind);
  lds_copy(lds + lay.off_cxr, g.coset_xr, lay.bytes_cxr);
  lds_copy(lds + lay.off_ckind, g.coset_kind, lay.bytes_ckind);
  lds_copy(lds + lay.off_celem, g.coset_elem, lay.bytes_celem);
  lds_copy(lds + lay.off_cinv, g.coset_inv, lay.bytes_celem);
  L.sqcnt = (const double*)(lds + lay.off_sqcnt);
  L.ph_re = (const double*)(lds + lay.off_ph_re);
  L.ph_im = (const double*)(lds + lay.off_ph_im);
  L.nib = (const uint64_t*)(lds + lay.off_nib);
  L.minrot = (const uint32_t*)(lds + lay.off_minrot);
  L.pair = (const uint32_t*)(lds + lay.off_pair);
  L.rot_hi = (const uint64_t*)(lds + lay.off_rot_hi);
  L.rot_lo = (const uint64_t*)(lds + lay.off_rot_lo);
  L.g_of_ab = (const int32_t*)(lds + lay.off_g_of_ab);
  L.inv = (const int32_t*)(lds + lay.off_inv);
  L.trivial = (const uint8_t*)(lds + lay.off_trivial);
  L.xr = (const uint64_t*)(lds + lay.off_xr);
  L.coset_nib = (const uint64_t*)(lds + lay.off_cnib);
  L.coset_pidx = (const int32_t*)(lds + lay.off_cpidx);
  L.coset_xr = (const uint64_t*)(lds + lay.off_cxr);
  L.coset_kind = (const int32_t*)(lds + lay.off_ckind);
  L.coset_elem = (const int32_t*)(lds + lay.off_celem);
  L.coset_inv = (const int32_t*)(lds + lay.off_cinv);
  return L;
}


// ----------------------------------------------------------------------------
// Torus canonicaliser with the per-row table entries kept in registers
// (Ly <= LYMAX, minimal-rotation table in LDS).  Same result as
// canonicalize() for kGroupTorus with g.minrot_table set.
// ----------------------------------------------------------------------------
template <int LYMAX>
__device__ __forceinline__ void canonicalize_torus_reg(const GroupDev& g, const GroupLds& L, uint64_t b, uint64_t& s,
                                                       int& h, int& cnt) {
  uint32_t e[LYMAX];
  uint32_t best = 0xffffffffu;
#pragma unroll
  for (int y = 0; y < LYMAX; ++y) {
    if (y < g.Ly) {
      e[y] = L.minrot[(uint32_t)((b >> (y * g.Lx)) & g.row_mask)];
      const uint32_t mv = e[y] & 0xffffu;
      best = mv < best ? mv : best;
    }
  }
  s = ~0ull;
  h = 0;
  cnt = 0;
#pragma unroll
  for (int y = 0; y < LYMAX; ++y) {
    if (y < g.Ly && (e[y] & 0xffffu) == best) {
      const int bb = g.Ly - 1 - y;
      const uint64_t shifted = rows_shift(b, bb, g.Lx, g.nb, g.full_mask);
      uint32_t am = e[y] >> 16;
      while (am) {
        const int a = __builtin_ctz(am);
        am &= am - 1;
        const uint64_t img = rows_rotl(shifted, a, g.Lx, L.rot_hi, L.rot_lo);
        const int el = L.g_of_ab[a + g.Lx * bb];
        if (img < s) {
          s = img;
          h = el;
          cnt = 1;
        } else if (img == s) {
          ++cnt;
          if (L.inv[el] > L.inv[h]) h = el;
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Open-addressing table of the representatives (linear probing) whose slots
// carry the current input vector next to each key: one cache line per image.
// ----------------------------------------------------------------------------
constexpr uint64_t kHashEmpty = ~0ull;

__device__ __forceinline__ uint32_t hash_word(uint64_t s) {
  uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
  uint32_t h = lo * 0x9E3779B1u ^ (hi + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

// The input vector spread over word positions (the matvec's lookup): a
// word's position is its bits below top_shift under the code of its top bits
// (topcode: the rank of the top value among those of the basis words, 0xffff
// for none; the codes keep the order).  values[position] holds the element
// of x of the basis word there, 0 at every other position, so that a
// canonical image finds its x element with one load and no rank; an image
// outside the basis adds w * 0.  Refreshed before every product.
// The spread of x over the word positions.  Plain: position = (code of the
// word's top bits) << top_shift | (the bits below).  Colex (every word of the
// space has one popcount, colex != null): the low 12 bits are replaced by
// their colex rank among the 12-bit words of their popcount (which the
// higher bits fix), position = ((code << hbits | bits 12..top_shift-1) *
// 924 + rank: 4.4 times fewer positions, so that more of the lines the
// images touch are shared (6x6 J1-J2: 4.1 GB instead of 18.3 GB in FP32).
struct XMapDev {
  void* values;             // null: not used
  const uint16_t* topcode;  // 2^top_bits codes (global memory; the kernels copy it to LDS)
  const uint16_t* colex;    // colex rank of every 12-bit word, u16[4096] (global; copied to LDS after the codes), or null
  int bm_bits;              // words must be below 2^bm_bits
  int top_bits, top_shift;
  int hbits;                // colex: top_shift - 12
  int pop;                  // colex: the popcount of every word (a word of another popcount has no position);
                            // -1 where every image keeps the popcount (no test needed)
  int scaled;               // 1: the spread holds x / sqrt(N) of each representative (orbit size N)
};

// LDS bytes of the code table and, when used, the colex tables after it.
__host__ __device__ __forceinline__ uint32_t xmap_lds_bytes(const XMapDev& X) {
  return (2u << X.top_bits) + (X.colex ? 2u * 4096u : 0u);
}

// Position of word k whose top bits have code c; cx: the colex rank table
// (LDS or global), used when X.colex is set.
__device__ __forceinline__ uint64_t xmap_pos(const XMapDev& X, const uint16_t* cx, uint32_t c, uint64_t k) {
  if (X.colex) {
    // (c 2^hbits + bits 12 .. 11 + hbits) 924 + the rank of the low 12 bits,
    // in 32 bits: the host keeps the colex form to words of at most 44 bits
    // (hbits <= 20) and spreads of fewer than 2^32 positions
    const uint32_t lo = (uint32_t)k;
    const uint32_t hc = (c << X.hbits) | __builtin_amdgcn_ubfe(lo, 12u, (uint32_t)X.hbits);
    return (uint64_t)(__umul24(hc, 924u) + (uint32_t)cx[lo & 0xfffu]);
  }
  return ((uint64_t)c << X.top_shift) | (k & ((1ull << X.top_shift) - 1ull));
}

// x element of one canonical image k (false: the word has no position).
// codes: the code table in LDS, followed by the colex rank table when used.
template <typename V>
__device__ __forceinline__ bool xmap_fetch1(const XMapDev& X, const uint16_t* codes, uint64_t k, V& x) {
  const bool in = (k >> X.bm_bits) == 0 && (!X.colex || X.pop < 0 || __builtin_popcountll(k) == X.pop);
  const uint32_t c = codes[in ? (uint32_t)(k >> X.top_shift) : 0u];
  const bool f = in && c != 0xffffu;
  x = ((const V*)X.values)[f ? xmap_pos(X, codes + (1u << X.top_bits), c, k) : 0ull];
  return f;
}

// x elements of two canonical images k0, k1 (f0, f1: the word has a
// position), without branches: both code reads are issued, then both loads;
// an image without a position loads position 0 and is dropped by the caller.
template <typename V>
__device__ __forceinline__ void xmap_fetch2(const XMapDev& X, const uint16_t* codes, uint64_t k0, uint64_t k1, V& x0,
                                            V& x1, bool& f0, bool& f1) {
  // (colex positions are unique within one popcount only)
  bool in0 = (k0 >> X.bm_bits) == 0, in1 = (k1 >> X.bm_bits) == 0;
  if (X.colex && X.pop >= 0) {  // the same branch in every lane
    in0 = in0 && __builtin_popcountll(k0) == X.pop;
    in1 = in1 && __builtin_popcountll(k1) == X.pop;
  }
  const uint32_t c0 = codes[in0 ? (uint32_t)(k0 >> X.top_shift) : 0u];
  const uint32_t c1 = codes[in1 ? (uint32_t)(k1 >> X.top_shift) : 0u];
  f0 = in0 && c0 != 0xffffu;
  f1 = in1 && c1 != 0xffffu;
  const uint16_t* cx = codes + (1u << X.top_bits);
  const uint64_t p0 = f0 ? xmap_pos(X, cx, c0, k0) : 0ull;
  const uint64_t p1 = f1 ? xmap_pos(X, cx, c1, k1) : 0ull;
  x0 = ((const V*)X.values)[p0];
  x1 = ((const V*)X.values)[p1];
}

// Stages the code table and the colex tables of X into LDS at codes.
__device__ __forceinline__ void xmap_stage(const XMapDev& X, uint16_t* codes) {
  lds_copy((unsigned char*)codes, X.topcode, 2u << X.top_bits);
  if (X.colex) lds_copy((unsigned char*)(codes + (1u << X.top_bits)), X.colex, 2u * 4096u);
}

struct HashDev {
  unsigned char* slots;  // n_slots * slot_bytes; key at offset 0, vector element at offset xoff
  uint32_t mask;         // n_slots - 1 (power of two)
  int slot_bytes;        // 16 or 32
  int xoff;              // 8 or 16
  int pad_;
};

}  // namespace edhip
