// The edhip C ABI (include/edhip.h): contexts, spaces, operators, apply,
// iterators, sparse export, Lanczos, term expectations and MINRES.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/edhip.h"
#include "kernels.hpp"

using namespace edhip;

// ============================================================================
// Errors
// ============================================================================
namespace {

struct Error {
  int code;
  std::string msg;
};

inline size_t elem_size(int prec) {
  switch (prec) {
    case EDHIP_F64:
      return 8;
    case EDHIP_C128:
      return 16;
    case EDHIP_F32:
      return 4;
    case EDHIP_C64:
      return 8;
    default:
      return 0;
  }
}

inline bool prec_is_complex(int prec) { return prec == EDHIP_C128 || prec == EDHIP_C64; }

#define HIP_OR_THROW(expr)                                                                                      \
  do {                                                                                                          \
    hipError_t e_ = (expr);                                                                                     \
    if (e_ != hipSuccess) {                                                                                     \
      int code_ = (e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation) ? EDHIP_ERR_MEMORY              \
                                                                                : EDHIP_ERR_DEVICE;             \
      throw Error{code_, std::string(#expr) + ": " + hipGetErrorString(e_)};                                    \
    }                                                                                                           \
  } while (0)

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

// Injected allocation failures (edhip_debug_fail_allocs): the next
// g_fail_count allocations of the kinds in g_fail_kinds fail as if the
// device were out of memory, before any device call is made.
constexpr int kAllocSpread = 1, kAllocWorkspace = 2, kAllocAssembled = 4;
std::atomic<int> g_fail_kinds{0}, g_fail_count{0};
void injected_alloc_failure(int kind) {
  if (!(g_fail_kinds.load() & kind)) return;
  int c = g_fail_count.load();
  while (c > 0 && !g_fail_count.compare_exchange_weak(c, c - 1)) {
  }
  if (c > 0) fail(EDHIP_ERR_MEMORY, "injected allocation failure");
}

// Device buffer owned by a handle.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    HIP_OR_THROW(hipMalloc(&p, n));
    bytes = n;
  }
  void ensure(size_t n) {
    if (bytes < n) alloc(n);
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

template <typename T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIP_OR_THROW(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

}  // namespace

// ============================================================================
// Context
// ============================================================================
struct edhip_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string last_error;
  DevBuf partial;  // reduction partials
  DevBuf small;    // small results
  // Eigensolver / MINRES vectors, kept between calls (grow-only): repeated
  // allocations of hundreds of GB after frees cost seconds per call.
  DevBuf krylov;
  hipEvent_t ev[16] = {};
  hipEvent_t ev_ortho[2] = {};  // the eigensolver's timing of its orthogonalisation kernels
  bool kernel_timing = false;  // time every matvec kernel launch (edhip_set_kernel_timing)

  char* krylov_arena(size_t bytes) {
    if (krylov.bytes < bytes) injected_alloc_failure(kAllocWorkspace);
    krylov.ensure(bytes);
    return krylov.as<char>();
  }

  double* partial_buf(size_t doubles) {
    partial.ensure(doubles * sizeof(double));
    return partial.as<double>();
  }
  double* small_buf(size_t doubles) {
    small.ensure(doubles * sizeof(double));
    return small.as<double>();
  }
  void sync() { HIP_OR_THROW(hipStreamSynchronize(stream)); }
};

namespace {

template <typename F>
int guarded(edhip_ctx* ctx, F&& f) {
  try {
    if (ctx) HIP_OR_THROW(hipSetDevice(ctx->device));
    f();
    if (ctx) ctx->last_error.clear();
    return EDHIP_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->last_error = "host allocation failed";
    return EDHIP_ERR_MEMORY;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return EDHIP_ERR_DEVICE;
  }
}

// binomial-based completion counts W[i*(n+1)+ones]
std::vector<uint64_t> completion_counts(int n, const std::vector<int>& pops) {
  std::vector<std::vector<uint64_t>> C(n + 1, std::vector<uint64_t>(n + 1, 0));
  for (int i = 0; i <= n; ++i) {
    C[i][0] = 1;
    for (int k = 1; k <= i; ++k) C[i][k] = C[i - 1][k - 1] + (k <= i - 1 ? C[i - 1][k] : 0);
  }
  std::vector<uint64_t> W((size_t)(n + 1) * (n + 1), 0);
  for (int i = 0; i <= n; ++i)
    for (int ones = 0; ones <= n; ++ones) {
      uint64_t s = 0;
      for (int p : pops) {
        int need = p - ones;
        if (need >= 0 && need <= i) s += C[i][need];
      }
      W[(size_t)i * (n + 1) + ones] = s;
    }
  return W;
}

inline int bit_length(uint64_t v) { return v == 0 ? 0 : 64 - __builtin_clzll(v); }

}  // namespace

// ============================================================================
// Spaces
// ============================================================================
struct GroupHost {
  int G = 1;
  int mode = kGroupNone;
  int Lx = 0, Ly = 0;
  std::vector<int32_t> perm;  // G x n_bits
  std::vector<uint64_t> xr;   // G: word XORed into the permuted word (0 or all bits: global flip)
  std::vector<std::complex<double>> chi;
  std::vector<int32_t> inv;
  bool complex_phase = false;
  bool has_xor = false;
  int n_cosets = 1;  // torus: cosets T p of the translations
  // device tables
  DevBuf nib, minrot, pair, rot_hi, rot_lo, g_of_ab, inv_d, trivial, ph_re, ph_im, invsqrt, sqcnt, xr_d, coset_nib_d,
      coset_pidx_d, coset_xr_d, coset_kind_d, coset_elem_d, coset_inv_d;
  GroupDev dev{};
  GroupLdsLayout lay{};
};

// Top bits of the words coded in the spread of x (a table of 2^12 u16 codes,
// 8 KiB, staged in LDS by the matvec), and the largest spread, in elements
// per basis word, the matvec builds (6x6 k=(0,0): 4.07 in the colex form).
constexpr int kXTopBits = 12;
constexpr uint64_t kXMapRatio = 32;
// Spreads of 8- and 16-byte elements are used by default when they hold at
// most this many positions per basis word
constexpr uint64_t kXMapDenseRatio = 8;
// and 4-byte spreads with at most this many (6x6 J1-J2 FP32 products: the
// translation sector's 4.1 and 18.2 positions per word (colex and plain
// forms) beat the rank index, 83 and 97 against 116 ms)
constexpr uint64_t kXMapDenseRatio32 = 24;
// The blocked rank index is built while it takes at most this many times the
// bytes of the basis words it indexes (64: the 6x6 space-group A1 sector with
// the even spin flip, 15.8M words below 2^35, gets its 5.7 GB index, and its
// products run 52.4 instead of 62.9 ms through the sorted search)
constexpr uint64_t kBitmapMaxRatio = 64;
// Spreads of 8- and 16-byte elements are used by default only up to this
// size (about what the last-level cache holds)
constexpr uint64_t kXMapCachedBytes = 256ull << 20;

// Generic groups whose 4-bit permutation tables exceed this stay in global
// memory (read through L1/L2) instead of LDS.
constexpr size_t kNibLdsMax = 40 * 1024;

struct edhip_space {
  edhip_ctx* ctx = nullptr;
  int n_bits = 0;
  int64_t dim = 0;
  // sector description (popcount sector)
  bool is_sector = false;
  std::vector<int> popcounts;
  uint64_t pop_allowed = 0;
  DevBuf W;  // completion counts
  std::vector<uint64_t> W_host;
  // basis
  mutable bool materialized = false;
  mutable DevBuf keys;
  // sorted-search index
  mutable DevBuf prefix;
  mutable int shift = 0;
  mutable uint64_t n_buckets = 0;
  mutable bool index_built = false;
  // reduction
  const edhip_space* parent = nullptr;
  std::unique_ptr<GroupHost> group;
  DevBuf orbit;
  // open-addressing table of the basis words with an embedded vector slot
  mutable DevBuf hash_slots, hash_slot_of;
  mutable HashDev hash_dev{};

  // Builds (or widens) the table for vector elements of elem_bytes bytes.
  const HashDev& ensure_hash(size_t elem_bytes) const {
    const int want = elem_bytes > 8 ? 32 : 16;
    if (hash_dev.slots && hash_dev.slot_bytes >= want) return hash_dev;
    materialize();
    uint64_t n_slots = 16;
    while (n_slots < (uint64_t)dim * 2) n_slots <<= 1;
    if (n_slots > (1ull << 32)) fail(EDHIP_ERR_ARGUMENT, "basis too large for the hash lookup");
    hash_slots.alloc((size_t)n_slots * want);
    hash_slot_of.alloc((size_t)std::max<int64_t>(dim, 1) * 4);
    HIP_OR_THROW(hipMemsetAsync(hash_slots.p, 0xFF, (size_t)n_slots * want, ctx->stream));
    HashDev h{};
    h.slots = hash_slots.as<unsigned char>();
    h.mask = (uint32_t)(n_slots - 1);
    h.slot_bytes = want;
    h.xoff = want == 32 ? 16 : 8;
    HIP_OR_THROW(launch_hash_build(keys.as<uint64_t>(), dim, h, hash_slot_of.as<uint32_t>(), ctx->stream));
    hash_dev = h;
    return hash_dev;
  }

  void materialize() const {
    if (materialized) return;
    if (!is_sector) fail(EDHIP_ERR_ARGUMENT, "space has no basis");
    if ((uint64_t)dim > (1ull << 34)) fail(EDHIP_ERR_MEMORY, "sector too large to materialise");
    keys.alloc((size_t)dim * 8);
    HIP_OR_THROW(launch_enumerate_sector(0, (uint64_t)dim, n_bits, W.as<uint64_t>(), keys.as<uint64_t>(), ctx->stream));
    materialized = true;
  }

  void build_index() const {
    if (index_built) return;
    materialize();
    if (dim == 0) {
      shift = 0;
      n_buckets = 1;
      std::vector<uint32_t> z(2, 0);
      prefix.alloc(8);
      HIP_OR_THROW(hipMemcpyAsync(prefix.p, z.data(), 8, hipMemcpyHostToDevice, ctx->stream));
      index_built = true;
      return;
    }
    if ((uint64_t)dim >= 0xffffffffull) fail(EDHIP_ERR_ARGUMENT, "more than 2^32-1 basis states");
    uint64_t max_key = 0;
    HIP_OR_THROW(hipMemcpyAsync(&max_key, keys.as<uint64_t>() + (dim - 1), 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    int bl = std::max(1, bit_length(max_key));
    int want = std::max(1, bit_length((uint64_t)dim) + 1);
    int k = std::min({bl, want, 26});
    shift = bl - k;
    n_buckets = (max_key >> shift) + 1;
    prefix.alloc((size_t)(n_buckets + 1) * 4);
    HIP_OR_THROW(launch_build_prefix(keys.as<uint64_t>(), dim, shift, n_buckets, prefix.as<uint32_t>(), ctx->stream));
    index_built = true;
  }

  // Blocked rank index over [0, 2^bm_bits): built when the largest word has
  // at most 36 bits and the index is at most 24x the size of the key array.
  mutable DevBuf bitmap;
  mutable int bm_bits = 0;
  mutable bool bm_tried = false;

  // The input vector spread over word positions (XMapDev), one array per
  // element size (4, 8, 16 bytes), built on the first product that can use
  // it; positions without a basis word stay 0.
  struct XMapHost {
    DevBuf values;
    XMapDev dev{};
    bool tried = false, ok = false;
  };
  mutable XMapHost xmap[3];
  mutable DevBuf xtop;  // code of each value of the top bits of the words
  mutable int xtop_bits = -1, xtop_shift = 0, xtop_K = 0;
  mutable uint64_t xtop_codes = 0;
  mutable DevBuf xcolex;  // colex rank of every 12-bit word (single-popcount spaces)
  mutable bool xcolex_on = false;

  // The popcount every word of the space has, or -1: a popcount sector with
  // one allowed popcount, or a reduction of one.
  int single_popcount() const {
    if (parent) return parent->single_popcount();
    return (is_sector && popcounts.size() == 1) ? popcounts[0] : -1;
  }

  void release_xmaps() const {
    for (XMapHost& X : xmap) {
      X.values.release();
      X.tried = X.ok = false;
    }
  }

  // Code table of the top bits of the words (built once); returns the number
  // of spread positions, 0 when the space has no words.
  uint64_t ensure_xtop() const {
    if (dim == 0) return 0;
    if (xtop_bits < 0) {
      materialize();
      uint64_t max_key = 0;
      HIP_OR_THROW(hipMemcpyAsync(&max_key, keys.as<uint64_t>() + (dim - 1), 8, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
      const int K = std::max(4, bit_length(max_key));
      // colex positions for the low 12 bits when every word has one popcount
      // (the popcount of those bits is then fixed by the rest of the word);
      // the top code covers what lies above bit 11 (at most kXTopBits bits).
      // The colex form keeps its positions in 32 bits (the device computes
      // them so): words of at most 44 bits and fewer than 2^32 positions,
      // else the plain form.
      std::vector<uint16_t> codes;
      uint64_t nc = 0;
      int t = 0, sh = 0;
      for (int pass = 0; pass < 2; ++pass) {
        xcolex_on = pass == 0 && single_popcount() >= 0 && K >= 16 && K <= 44;
        t = xcolex_on ? std::min(kXTopBits, K - 12) : std::min(kXTopBits, K);
        sh = K - t;
        DevBuf flags;
        flags.alloc(((size_t)1 << t) * 4);
        HIP_OR_THROW(hipMemsetAsync(flags.p, 0, ((size_t)1 << t) * 4, ctx->stream));
        HIP_OR_THROW(launch_mark_tops(keys.as<uint64_t>(), dim, sh, flags.as<uint32_t>(), ctx->stream));
        std::vector<uint32_t> fl((size_t)1 << t);
        HIP_OR_THROW(hipMemcpyAsync(fl.data(), flags.p, fl.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        codes.assign(fl.size(), 0xffff);
        nc = 0;
        for (size_t v = 0; v < fl.size(); ++v)
          if (fl[v]) codes[v] = (uint16_t)nc++;
        if (!xcolex_on || ((nc << (sh - 12)) * (uint64_t)kColexStride >> 32) == 0) break;
      }
      upload(xtop, codes, ctx->stream);
      if (xcolex_on) upload(xcolex, colex12_full(), ctx->stream);
      xtop_bits = t;
      xtop_shift = sh;
      xtop_K = K;
      xtop_codes = nc;
    }
    return xcolex_on ? (xtop_codes << (xtop_shift - 12)) * (uint64_t)kColexStride : xtop_codes << xtop_shift;
  }

  const XMapHost* ensure_xmap(size_t es) const {
    const int slot = es == 4 ? 0 : es == 8 ? 1 : es == 16 ? 2 : -1;
    if (slot < 0) return nullptr;
    XMapHost& X = xmap[slot];
    if (X.tried) return X.ok ? &X : nullptr;
    X.tried = true;
    if (dim == 0) return nullptr;
    try {
      const uint64_t n_pos = ensure_xtop();
      // the spread may take up to kXMapRatio times the memory of the vector
      if (n_pos > (uint64_t)kXMapRatio * (uint64_t)dim && n_pos > (1ull << 20)) return nullptr;
      injected_alloc_failure(kAllocSpread);
      X.values.alloc(n_pos * es);
      HIP_OR_THROW(hipMemsetAsync(X.values.p, 0, n_pos * es, ctx->stream));
      ctx->sync();
      X.dev = XMapDev{X.values.p, xtop.as<uint16_t>(), xcolex_on ? xcolex.as<uint16_t>() : nullptr, xtop_K,
                      xtop_bits, xtop_shift, xcolex_on ? xtop_shift - 12 : 0, single_popcount(), 0};
      X.ok = true;
      return &X;
    } catch (const Error& e) {
      if (e.code != EDHIP_ERR_MEMORY) throw;
      (void)hipGetLastError();  // no room: the products use the rank index
      X.values.release();
      return nullptr;
    }
  }

  bool ensure_bitmap() const {
    if (bm_tried) return bm_bits > 0;
    bm_tried = true;
    build_index();
    if (dim == 0) return false;
    uint64_t max_key = 0;
    HIP_OR_THROW(hipMemcpyAsync(&max_key, keys.as<uint64_t>() + (dim - 1), 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    const int K = std::max(9, bit_length(max_key));
    if (K > 36) return false;
    const uint64_t n_blocks = ((1ull << K) + kRankBlockWords - 1) / kRankBlockWords;
    const uint64_t bytes = n_blocks * 64;
    if (bytes > std::max<uint64_t>(kBitmapMaxRatio * (uint64_t)dim * 8ull, 1ull << 20)) return false;
    bitmap.alloc(bytes);
    HIP_OR_THROW(hipMemsetAsync(bitmap.p, 0, bytes, ctx->stream));
    DevBuf totals, base, tmp;
    totals.alloc(n_blocks * 4);
    base.alloc(n_blocks * 8);
    size_t tb = 0;
    HIP_OR_THROW(launch_bitmap_build(keys.as<uint64_t>(), dim, bitmap.as<uint64_t>(), n_blocks, totals.as<uint32_t>(),
                                     base.as<uint64_t>(), nullptr, tb, ctx->stream));
    tmp.alloc(tb);
    HIP_OR_THROW(launch_bitmap_build(keys.as<uint64_t>(), dim, bitmap.as<uint64_t>(), n_blocks, totals.as<uint32_t>(),
                                     base.as<uint64_t>(), tmp.p, tb, ctx->stream));
    ctx->sync();
    bm_bits = K;
    return true;
  }

  // Index of canonical words in this space.  lookup: 0 auto (bitmap when
  // available, else sorted search), 1 sorted search, 3 bitmap.  A full
  // popcount sector always uses rank arithmetic.
  IndexDev index_dev(int lookup = 0) const {
    IndexDev ix{};
    ix.dim = dim;
    ix.n_bits = n_bits;
    if (is_sector && !group) {
      ix.rank_mode = 1;
      ix.pop_allowed = pop_allowed;
      ix.W = W.as<uint64_t>();
      materialize();
      ix.keys = keys.as<uint64_t>();
      return ix;
    }
    build_index();
    ix.keys = keys.as<uint64_t>();
    ix.prefix = prefix.as<uint32_t>();
    ix.n_buckets = n_buckets;
    ix.shift = shift;
    if (lookup == 0 || lookup == 3) {
      if (ensure_bitmap()) {
        ix.bitmap = bitmap.as<uint64_t>();
        ix.bm_counts = nullptr;
        ix.bm_bits = bm_bits;
      } else if (lookup == 3) {
        fail(EDHIP_ERR_ARGUMENT, "no bitmap index for this space (words wider than 36 bits or too sparse)");
      }
    }
    return ix;
  }

  ReduceSource source() const {
    ReduceSource s{};
    s.n_words = (uint64_t)dim;
    s.n_bits = n_bits;
    if (is_sector) {
      s.list = nullptr;
      s.single_popcount = (popcounts.size() == 1 && popcounts[0] > 0) ? 1 : 0;
      s.W = W.as<uint64_t>();
    } else {
      materialize();
      s.list = keys.as<uint64_t>();
    }
    return s;
  }
};

namespace {

// ---- group analysis --------------------------------------------------------
uint32_t align_up(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

bool match_torus(int nb, int G, const std::vector<int32_t>& perm, int Lx, int Ly, std::vector<int32_t>& g_of_ab) {
  if (Lx * Ly != nb || G != nb) return false;
  g_of_ab.assign((size_t)nb, -1);
  for (int g = 0; g < G; ++g) {
    int img0 = perm[(size_t)g * nb + 0];
    int a = img0 % Lx, b = img0 / Lx;
    for (int y = 0; y < Ly; ++y)
      for (int x = 0; x < Lx; ++x) {
        int src = x + Lx * y;
        int dst = ((x + a) % Lx) + Lx * ((y + b) % Ly);
        if (perm[(size_t)g * nb + src] != dst) return false;
      }
    if (g_of_ab[a + Lx * b] != -1) return false;
    g_of_ab[a + Lx * b] = g;
  }
  return true;
}

void build_group(GroupHost& gh, int nb, double tol, hipStream_t st) {
  const int G = gh.G;
  // identity first
  for (int i = 0; i < nb; ++i)
    if (gh.perm[i] != i) fail(EDHIP_ERR_ARGUMENT, "the first symmetry operation must be the identity");
  // each element is a permutation of the bits
  for (int g = 0; g < G; ++g) {
    std::vector<char> seen(nb, 0);
    for (int i = 0; i < nb; ++i) {
      int j = gh.perm[(size_t)g * nb + i];
      if (j < 0 || j >= nb || seen[j]) fail(EDHIP_ERR_ARGUMENT, "symmetry operation is not a permutation");
      seen[j] = 1;
    }
  }
  // |chi| == 1 (isapprox(abs(y), 1) with the default rtol)
  const double rtol = std::sqrt(2.220446049250313e-16);
  for (int g = 0; g < G; ++g) {
    double a = std::abs(gh.chi[g]);
    if (!(std::fabs(a - 1.0) <= rtol * std::max(a, 1.0))) fail(EDHIP_ERR_ARGUMENT, "all amplitudes need to have norm 1");
  }
  // XOR masks: none, or the global flip (all bits) on some elements
  const uint64_t full = (nb >= 64) ? ~0ull : ((1ull << nb) - 1ull);
  if (gh.xr.empty()) gh.xr.assign(G, 0ull);
  if (gh.xr[0] != 0) fail(EDHIP_ERR_ARGUMENT, "the first symmetry operation must be the identity");
  gh.has_xor = false;
  for (int g = 0; g < G; ++g) {
    if (gh.xr[g] != 0 && gh.xr[g] != full) fail(EDHIP_ERR_ARGUMENT, "a bit flip must flip every bit (GlobalBitFlip)");
    if (gh.xr[g]) gh.has_xor = true;
  }
  // closure and inverses (host_logic.hpp)
  const std::string why = group_inverses(G, nb, gh.perm, gh.xr, gh.inv);
  if (!why.empty()) fail(EDHIP_ERR_ARGUMENT, why);
  // element index by (permutation, XOR word), for the coset decomposition
  std::map<std::pair<std::vector<int32_t>, uint64_t>, int> index;
  for (int g = 0; g < G; ++g)
    index[std::make_pair(std::vector<int32_t>(gh.perm.begin() + (size_t)g * nb, gh.perm.begin() + (size_t)(g + 1) * nb),
                         gh.xr[g])] = g;
  std::vector<int32_t> comp(nb);
  // canonicaliser choice: if every translation of an Lx x Ly torus (no flip)
  // is an element, the group is a union of cosets T p and canonicalises as
  // min over p of the translation-canonical form of p(b); else generic tables.
  auto find_translations = [&](int Lx, int Ly, std::vector<int32_t>& gab) {
    gab.assign((size_t)nb, -1);
    for (int g = 0; g < G; ++g) {
      if (gh.xr[g] != 0) continue;
      const int img0 = gh.perm[(size_t)g * nb + 0];
      const int a = img0 % Lx, b = img0 / Lx;
      bool ok = true;
      for (int y = 0; y < Ly && ok; ++y)
        for (int x = 0; x < Lx && ok; ++x)
          ok = gh.perm[(size_t)g * nb + x + Lx * y] == ((x + a) % Lx) + Lx * ((y + b) % Ly);
      if (ok) gab[a + Lx * b] = g;
    }
    for (int v : gab)
      if (v < 0) return false;
    return true;
  };
  std::vector<int32_t> g_of_ab;
  int bestLx = 0, bestLy = 0;
  std::vector<int32_t> best_map;
  for (int Ly = 1; Ly <= nb && G % nb == 0; ++Ly) {
    if (nb % Ly) continue;
    int Lx = nb / Ly;
    if (Lx < 2) continue;
    if (!find_translations(Lx, Ly, g_of_ab)) continue;
    bool better = false;
    if (bestLx == 0)
      better = true;
    else {
      bool t_new = Lx <= 12, t_old = bestLx <= 12;
      better = (t_new && !t_old) || (t_new == t_old && Lx > bestLx);
    }
    if (better) {
      bestLx = Lx;
      bestLy = Ly;
      best_map = g_of_ab;
    }
  }
  // cosets T p: representatives without an XOR word in element order (the
  // identity first), each followed by the cosets of its products with the
  // group's XOR words (same permutation, so that a canonicaliser applying the
  // cosets in order permutes once for both), then any coset left
  std::vector<int32_t> coset_rep, coset_elem;
  if (bestLx > 0) {
    std::vector<char> covered(G, 0);
    auto add_coset = [&](int g) {
      coset_rep.push_back(g);
      for (int q = 0; q < nb; ++q) {
        const int t = best_map[q];
        for (int i = 0; i < nb; ++i) comp[i] = gh.perm[(size_t)t * nb + gh.perm[(size_t)g * nb + i]];
        auto it = index.find(std::make_pair(comp, gh.xr[g]));  // T(q) p keeps p's flip
        if (it == index.end() || covered[it->second]) fail(EDHIP_ERR_ARGUMENT, "symmetry operations must form a group");
        covered[it->second] = 1;
        coset_elem.push_back(it->second);
      }
    };
    std::vector<uint64_t> xwords;
    for (int g = 0; g < G; ++g)
      if (gh.xr[g] && std::find(xwords.begin(), xwords.end(), gh.xr[g]) == xwords.end()) xwords.push_back(gh.xr[g]);
    for (int g = 0; g < G; ++g) {
      if (covered[g] || gh.xr[g]) continue;
      add_coset(g);
      const std::vector<int32_t> pg(gh.perm.begin() + (size_t)g * nb, gh.perm.begin() + (size_t)(g + 1) * nb);
      for (uint64_t x : xwords) {
        auto it = index.find(std::make_pair(pg, x));
        if (it != index.end() && !covered[it->second]) add_coset(it->second);
      }
    }
    for (int g = 0; g < G; ++g)
      if (!covered[g]) add_coset(g);
  }
  gh.mode = (bestLx > 0) ? kGroupTorus : kGroupGeneric;
  if (G == 1) gh.mode = kGroupGeneric;

  // phases, triviality, 1/sqrt(N)
  std::vector<double> pr(G), pi(G), isq(G + 1, 0.0), sqc(G + 1, 0.0);
  std::vector<uint8_t> triv(align_up((uint32_t)G, 4), 0);
  gh.complex_phase = false;
  for (int h = 0; h < G; ++h) {
    std::complex<double> ph = std::conj(gh.chi[gh.inv[h]]);
    pr[h] = ph.real();
    pi[h] = ph.imag();
    if (std::fabs(pi[h]) > 1e-14) gh.complex_phase = true;
    triv[h] = (std::abs(gh.chi[h] - 1.0) <= tol) ? 1 : 0;
  }
  for (int n = 1; n <= G; ++n) isq[n] = 1.0 / std::sqrt((double)n);
  for (int c = 1; c <= G; ++c) sqc[c] = (G % c == 0) ? isq[G / c] : 0.0;

  GroupDev& d = gh.dev;
  d = GroupDev{};
  d.mode = gh.mode;
  d.G = G;
  d.nb = nb;
  d.unit_phases = 1;
  for (int h = 0; h < G; ++h)
    if (pr[h] != 1.0 || pi[h] != 0.0) d.unit_phases = 0;
  d.has_xor = gh.has_xor ? 1 : 0;
  d.n_cosets = 1;
  d.full_mask = (nb >= 64) ? ~0ull : ((1ull << nb) - 1);
  GroupLdsLayout& L = gh.lay;
  L = GroupLdsLayout{};
  uint32_t off = 0;
  auto place = [&](uint32_t& o, uint32_t bytes) {
    off = align_up(off, 16);
    o = off;
    off += bytes;
  };
  upload(gh.ph_re, pr, st);
  upload(gh.ph_im, pi, st);
  upload(gh.invsqrt, isq, st);
  upload(gh.sqcnt, sqc, st);
  upload(gh.inv_d, gh.inv, st);
  upload(gh.trivial, triv, st);
  d.ph_re = gh.ph_re.as<double>();
  d.ph_im = gh.ph_im.as<double>();
  d.invsqrt = gh.invsqrt.as<double>();
  d.sqcnt = gh.sqcnt.as<double>();
  d.inv = gh.inv_d.as<int32_t>();
  d.trivial = gh.trivial.as<uint8_t>();
  L.bytes_ph = (uint32_t)(G * 8);
  L.bytes_invsqrt = (uint32_t)((G + 1) * 8);
  L.bytes_inv = (uint32_t)(G * 4);
  L.bytes_trivial = (uint32_t)triv.size();
  place(L.off_ph_re, L.bytes_ph);
  place(L.off_ph_im, L.bytes_ph);
  place(L.off_sqcnt, L.bytes_invsqrt);  // 1/sqrt(N) itself is read from global memory
  place(L.off_inv, L.bytes_inv);
  place(L.off_trivial, L.bytes_trivial);
  if (gh.has_xor && gh.mode == kGroupGeneric) {
    upload(gh.xr_d, gh.xr, st);
    d.xr = gh.xr_d.as<uint64_t>();
    L.bytes_xr = (uint32_t)(G * 8);
    place(L.off_xr, L.bytes_xr);
  }
  if (gh.mode == kGroupTorus) {
    gh.Lx = bestLx;
    gh.Ly = bestLy;
    d.Lx = bestLx;
    d.Ly = bestLy;
    d.row_mask = (bestLx >= 64) ? ~0ull : ((1ull << bestLx) - 1);
    d.minrot_table = bestLx <= 12 ? 1 : 0;
    std::vector<uint64_t> hi(bestLx, 0), lo(bestLx, 0);
    for (int a = 0; a < bestLx; ++a)
      for (int y = 0; y < bestLy; ++y) {
        uint64_t rowbits_hi = ((d.row_mask << a) & d.row_mask);
        uint64_t rowbits_lo = (a == 0) ? 0 : ((1ull << a) - 1);
        hi[a] |= rowbits_hi << (bestLx * y);
        lo[a] |= rowbits_lo << (bestLx * y);
      }
    upload(gh.rot_hi, hi, st);
    upload(gh.rot_lo, lo, st);
    upload(gh.g_of_ab, best_map, st);
    // coset tables
    const int nc = (int)coset_rep.size();
    d.n_cosets = nc;
    gh.n_cosets = nc;
    const int n_nib = (nb + 3) / 4;
    d.n_nib = n_nib;
    // 4-bit chunk tables of the distinct permutations among the representatives
    // (a representative and its product with the flip share one)
    std::vector<uint64_t> cnib, cxr(nc, 0);
    std::vector<int32_t> ckind(nc, 0), cpidx(nc, 0), cinv((size_t)nc * nb, 0);
    std::map<std::vector<int32_t>, int> pidx;
    for (int c = 0; c < nc; ++c) {
      const int g = coset_rep[c];
      std::vector<int32_t> pg(gh.perm.begin() + (size_t)g * nb, gh.perm.begin() + (size_t)(g + 1) * nb);
      bool ident = true;
      for (int i = 0; i < nb; ++i) ident = ident && pg[i] == i;
      cxr[c] = gh.xr[g];
      ckind[c] = ident ? (gh.xr[g] ? 1 : 0) : 2;
      if (!ident) {
        auto it = pidx.find(pg);
        if (it == pidx.end()) {
          it = pidx.emplace(pg, (int)pidx.size()).first;
          for (int k = 0; k < n_nib; ++k)
            for (int v = 0; v < 16; ++v) {
              uint64_t o = 0;
              for (int q = 0; q < 4; ++q) {
                int bit = 4 * k + q;
                if (bit < nb && ((v >> q) & 1)) o |= 1ull << pg[bit];
              }
              cnib.push_back(o);
            }
        }
        cpidx[c] = it->second;
      }
      for (int q = 0; q < nb; ++q) cinv[(size_t)c * nb + q] = gh.inv[coset_elem[(size_t)c * nb + q]];
    }
    upload(gh.coset_nib_d, cnib, st);
    upload(gh.coset_pidx_d, cpidx, st);
    upload(gh.coset_xr_d, cxr, st);
    upload(gh.coset_kind_d, ckind, st);
    upload(gh.coset_elem_d, coset_elem, st);
    upload(gh.coset_inv_d, cinv, st);
    d.coset_nib = gh.coset_nib_d.as<uint64_t>();
    d.coset_pidx = gh.coset_pidx_d.as<int32_t>();
    d.coset_xr = gh.coset_xr_d.as<uint64_t>();
    d.coset_kind = gh.coset_kind_d.as<int32_t>();
    d.coset_elem = gh.coset_elem_d.as<int32_t>();
    d.coset_inv = gh.coset_inv_d.as<int32_t>();
    // the same, packed into kernel arguments when small enough
    d.coset_regs = 0;
    uint64_t xw = 0;
    bool one_x = true;
    for (int c = 0; c < nc; ++c)
      if (cxr[c]) {
        if (xw && cxr[c] != xw) one_x = false;
        xw = cxr[c];
      }
    if (nc <= 16 && pidx.size() <= 16 && one_x) {
      d.coset_regs = 1;
      d.ck_bits = 0;
      d.cp_bits = 0;
      d.cx_bits = 0;
      d.cxword = xw;
      for (int c = 0; c < nc; ++c) {
        d.ck_bits |= (uint32_t)ckind[c] << (2 * c);
        d.cp_bits |= (uint64_t)cpidx[c] << (4 * c);
        if (cxr[c]) d.cx_bits |= 1u << c;
      }
    }
    L.bytes_cnib = (uint32_t)(cnib.size() * 8);
    L.bytes_cxr = (uint32_t)(nc * 8);
    L.bytes_ckind = (uint32_t)(nc * 4);
    L.bytes_celem = (uint32_t)(nc * nb * 4);
    if (L.bytes_cnib) place(L.off_cnib, L.bytes_cnib);
    place(L.off_cpidx, L.bytes_ckind);
    place(L.off_cxr, L.bytes_cxr);
    place(L.off_ckind, L.bytes_ckind);
    place(L.off_celem, L.bytes_celem);
    place(L.off_cinv, L.bytes_celem);
    d.rot_hi = gh.rot_hi.as<uint64_t>();
    d.rot_lo = gh.rot_lo.as<uint64_t>();
    d.g_of_ab = gh.g_of_ab.as<int32_t>();
    L.bytes_rot = (uint32_t)(bestLx * 8);
    L.bytes_g_of_ab = (uint32_t)(bestLx * bestLy * 4);
    place(L.off_rot_hi, L.bytes_rot);
    place(L.off_rot_lo, L.bytes_rot);
    place(L.off_g_of_ab, L.bytes_g_of_ab);
    if (d.minrot_table) {
      int nv = 1 << bestLx;
      std::vector<uint32_t> mr(nv);
      for (int v = 0; v < nv; ++v) {
        uint32_t mv = (uint32_t)v, am = 1;
        for (int a = 1; a < bestLx; ++a) {
          uint32_t r = (uint32_t)((((uint64_t)v << a) | ((uint64_t)v >> (bestLx - a))) & d.row_mask);
          if (r < mv) {
            mv = r;
            am = 1u << a;
          } else if (r == mv) {
            am |= 1u << a;
          }
        }
        mr[v] = mv | (am << 16);
      }
      upload(gh.minrot, mr, st);
      d.minrot = gh.minrot.as<uint32_t>();
      L.bytes_minrot = (uint32_t)(nv * 4);
      place(L.off_minrot, L.bytes_minrot);
    }
    d.pair_table = (bestLx <= 6 && bestLy >= 2 && bestLy <= 8) ? 1 : 0;
    if (d.pair_table) {
      const int Lx = bestLx;
      const uint32_t rm = (uint32_t)d.row_mask;
      auto rot = [&](uint32_t v, int a) { return a == 0 ? v : (((v << a) | (v >> (Lx - a))) & rm); };
      const int np = 1 << (2 * Lx);
      std::vector<uint32_t> pt(np);
      for (int v = 0; v < np; ++v) {
        const uint32_t hi = (uint32_t)v >> Lx, lo = (uint32_t)v & rm;
        uint32_t best = 0xffffffffu, am = 0;
        for (int a = 0; a < Lx; ++a) {
          const uint32_t val = (rot(hi, a) << Lx) | rot(lo, a);
          if (val < best) {
            best = val;
            am = 1u << a;
          } else if (val == best) {
            am |= 1u << a;
          }
        }
        pt[v] = (best << 8) | am;
      }
      upload(gh.pair, pt, st);
      d.pair = gh.pair.as<uint32_t>();
      L.bytes_pair = (uint32_t)(np * 4);
    }
  } else {
    int n_nib = (nb + 3) / 4;
    d.n_nib = n_nib;
    std::vector<uint64_t> tab((size_t)G * n_nib * 16, 0);
    for (int g = 0; g < G; ++g)
      for (int k = 0; k < n_nib; ++k)
        for (int v = 0; v < 16; ++v) {
          uint64_t o = 0;
          for (int q = 0; q < 4; ++q) {
            int bit = 4 * k + q;
            if (bit < nb && ((v >> q) & 1)) o |= 1ull << gh.perm[(size_t)g * nb + bit];
          }
          tab[((size_t)g * n_nib + k) * 16 + v] = o;
        }
    upload(gh.nib, tab, st);
    d.nib = gh.nib.as<uint64_t>();
    if (tab.size() * 8 <= kNibLdsMax) {
      L.bytes_nib = (uint32_t)(tab.size() * 8);
      place(L.off_nib, L.bytes_nib);
    } else {
      d.nib_global = 1;  // large groups read their tables through the caches
    }
  }
  // the row-pair table goes last, so that an operator whose terms leave no
  // room for it can drop it (make_opdev) without moving the other tables
  if (L.bytes_pair) place(L.off_pair, L.bytes_pair);
  L.total = align_up(off, 16);
}

GroupDev trivial_group_dev() {
  GroupDev d{};
  d.mode = kGroupNone;
  d.G = 1;
  return d;
}

}  // namespace

// ============================================================================
// Operators
// ============================================================================
struct edhip_op {
  edhip_ctx* ctx = nullptr;
  const edhip_space* space = nullptr;
  int64_t n_terms = 0;
  std::vector<uint64_t> mask, row, col;
  std::vector<double> are, aim;
  bool complex_amp = false;
  bool keeps_popcount = true;  // every off-diagonal term's image has its row's popcount
  // device terms
  DevBuf off_row, off_col, diag, all_row, all_col;
  DevBuf mv_row, mv_col;  // paired copies of off_row / off_col for the matvec
  int n_mv = 0;
  int n_off = 0, n_diag = 0;
  std::vector<int64_t> diag_ids, off_ids;  // original term index of each device term
  OpDev dev_row{}, dev_col{};
  // scratch
  DevBuf hx, hy, iter_rows, iter_idx, iter_re, iter_im;
  // sparse result
  int64_t sp_nnz = -1;
  DevBuf sp_colptr, sp_rowval, sp_re, sp_im;
  // expectation scratch
  DevBuf ex_partial, ex_out;
  // fast path: bond-pair groups (row and column forms)
  bool pairs_ok = false;
  DevBuf pg_meta[2], pg_dre[2], pg_dim[2], pg_ore[2], pg_oim[2];
  PairGroupsDev pg[2]{};
  // per-row diagonal elements for the term-list kernel (built on first use)
  bool diag_tried = false;
  DevBuf diag_re, diag_im;
  int strat_kernel = 0;  // 0 auto, 1 term list, 2 pair groups, 3 assembled rows
  int strat_lookup = 0;  // 0 auto, 1 sorted search (or rank), 2 hash table
  int used_kernel = 0, used_lookup = 0;
  double kt_ms = 0.0;  // summed durations of the timed matvec and term-pass kernel launches
  int64_t kt_n = 0;
  // assembled rows of the row form (strategy 3), built on first use
  bool asm_built = false;
  DevBuf asm_rowptr, asm_ent, asm_w, asm_acls, asm_pcls, asm_xt;
  AsmDev asm_dev{};
  uint64_t asm_nnz = 0;
  double asm_build_s = 0.0;
  void release_assembled() {
    asm_built = false;
    asm_rowptr.release();
    asm_ent.release();
    asm_w.release();
    asm_acls.release();
    asm_pcls.release();
    asm_xt.release();
    asm_dev = AsmDev{};
    asm_nnz = 0;
  }

  bool is_complex() const { return complex_amp || (space->group && space->group->complex_phase); }
};

namespace {

void check_space(const edhip_space* s) {
  if (!s) fail(EDHIP_ERR_ARGUMENT, "null space");
}

// A group with the global flip needs a parent closed under it: a popcount
// sector with p allowed iff n_bits - p is, or a list equal to its complement.
void check_flip_closed(const edhip_space* p) {
  const uint64_t full = (p->n_bits >= 64) ? ~0ull : ((1ull << p->n_bits) - 1ull);
  if (p->is_sector) {
    for (int q : p->popcounts)
      if (std::find(p->popcounts.begin(), p->popcounts.end(), p->n_bits - q) == p->popcounts.end())
        fail(EDHIP_ERR_ARGUMENT, "the global bit flip maps the sector outside itself");
    return;
  }
  p->materialize();
  std::vector<uint64_t> k((size_t)p->dim);
  if (p->dim) HIP_OR_THROW(hipMemcpy(k.data(), p->keys.p, (size_t)p->dim * 8, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < p->dim; ++i)
    if ((k[(size_t)i] ^ full) != k[(size_t)(p->dim - 1 - i)])
      fail(EDHIP_ERR_ARGUMENT, "the global bit flip maps the basis outside itself");
}

OpDev make_opdev(edhip_op* op, bool transpose) {
  const edhip_space* sp = op->space;
  OpDev d{};
  d.ix = sp->index_dev();
  d.dim = sp->dim;
  sp->materialize();
  d.rows = sp->keys.as<uint64_t>();
  if (sp->group) {
    d.g = sp->group->dev;
    d.lay = sp->group->lay;
    d.orbit = sp->orbit.as<uint16_t>();
  } else {
    d.g = trivial_group_dev();
    d.lay = GroupLdsLayout{};
    d.orbit = nullptr;
  }
  d.n_off = op->n_off;
  d.n_diag = op->n_diag;
  d.word_bits = sp->n_bits;
  for (int i = 0; i < op->n_diag; ++i) {
    const int64_t t = op->diag_ids[i];
    d.word_bits = std::max({d.word_bits, bit_length(op->mask[t]), bit_length(op->row[t])});
  }
  d.off = transpose ? op->off_col.as<OffTerm>() : op->off_row.as<OffTerm>();
  d.diag = op->diag.as<DiagTerm>();
  // the matvec stages 36 bytes per padded term, the other kernels 40 per term
  const uint32_t w_bytes = d.ix.rank_mode ? (uint32_t)((sp->n_bits + 1) * (sp->n_bits + 1) * 8) : 0u;
  const uint32_t plain_bytes = align_up((uint32_t)(op->n_off * sizeof(OffTerm)), 16);
  const uint32_t paired_bytes = std::max(plain_bytes, align_up((uint32_t)(op->n_mv * 36), 16));
  const uint32_t group_bytes = align_up(d.lay.total, 16);
  d.mv_off = transpose ? op->mv_col.as<OffTerm>() : op->mv_row.as<OffTerm>();
  d.n_mv = op->n_mv;
  if (group_bytes + paired_bytes + w_bytes > 64 * 1024 && group_bytes + plain_bytes + w_bytes <= 64 * 1024)
    d.n_mv = 0;  // no room for the paired copy: match term by term
  const uint32_t op_bytes = (d.n_mv ? paired_bytes : plain_bytes) + w_bytes;
  if (d.g.pair_table && group_bytes + op_bytes > 64 * 1024) {
    // not enough LDS next to this operator's terms: use the per-row table path
    d.g.pair_table = 0;
    d.lay.bytes_pair = 0;
    d.lay.total = d.lay.off_pair;
  }
  uint32_t off = align_up(d.lay.total, 16);
  d.off_lds = off;
  off += d.n_mv ? paired_bytes : plain_bytes;
  off = align_up(off, 16);
  d.w_lds = off;
  d.w_bytes = d.ix.rank_mode ? (uint32_t)((sp->n_bits + 1) * (sp->n_bits + 1) * 8) : 0u;
  off += d.w_bytes;
  d.lds_total = align_up(off, 16);
  // the matvec alone stages the code table of the spread x after the rest
  d.xm = XMapDev{};
  // (room for the largest code table and the colex rank table)
  d.codes_lds = d.lds_total + (2u << kXTopBits) + 2u * 4096u <= 64 * 1024 ? d.lds_total : 0;
  if (d.lds_total > 64 * 1024)
    fail(EDHIP_ERR_ARGUMENT, "operator and symmetry tables exceed the 64 KiB LDS budget of the matvec kernels (" +
                                 std::to_string(d.lds_total) + " bytes)");
  return d;
}

void check_prec_for_op(const edhip_op* op, int prec) {
  if (elem_size(prec) == 0) fail(EDHIP_ERR_ARGUMENT, "unknown precision code");
  if (!prec_is_complex(prec) && op->is_complex())
    fail(EDHIP_ERR_ARGUMENT, "complex operator or characters need a complex vector type");
}

}  // namespace

namespace {

// Groups the terms by bit mask when every mask has one or two bits and each
// (mask, pattern) has at most one off-diagonal entry.  form 0: row form
// (match bitrow, image bitcol); form 1: column form (roles swapped).
bool build_pair_groups(edhip_op* op, int form, hipStream_t st) {
  std::vector<uint64_t> masks;
  std::map<uint64_t, int> gid;
  double cre = 0.0, cim = 0.0;
  for (int64_t t = 0; t < op->n_terms; ++t) {
    const int pc = __builtin_popcountll(op->mask[t]);
    if (pc == 0) continue;
    if (pc > 2) return false;
    if (!gid.count(op->mask[t])) {
      gid[op->mask[t]] = (int)masks.size();
      masks.push_back(op->mask[t]);
    }
  }
  const int ng = (int)masks.size();
  if (ng > kMaxPairGroups) return false;
  std::vector<uint32_t> meta(ng, 0);
  std::vector<double> dre(4 * (size_t)ng, 0.0), dim(4 * (size_t)ng, 0.0), ore(4 * (size_t)ng, 0.0),
      oim(4 * (size_t)ng, 0.0);
  std::vector<int> pq(2 * (size_t)ng);
  for (int g = 0; g < ng; ++g) {
    const int p = __builtin_ctzll(masks[g]);
    const int q = 63 - __builtin_clzll(masks[g]);
    pq[2 * g] = p;
    pq[2 * g + 1] = q;
    meta[g] = (uint32_t)p | ((uint32_t)q << 8);
  }
  auto code_of = [&](int g, uint64_t bits) -> uint32_t {
    const int p = pq[2 * g], q = pq[2 * g + 1];
    return (uint32_t)((bits >> p) & 1ull) | ((uint32_t)((bits >> q) & 1ull) << 1);
  };
  for (int64_t t = 0; t < op->n_terms; ++t) {
    const uint64_t m = op->mask[t];
    const uint64_t match = form == 0 ? op->row[t] : op->col[t];
    const uint64_t image = form == 0 ? op->col[t] : op->row[t];
    if (m == 0) {
      cre += op->are[t];
      cim += op->aim[t];
      continue;
    }
    const int g = gid[m];
    const uint32_t c = code_of(g, match), nc = code_of(g, image);
    if (match == image) {
      dre[4 * (size_t)g + c] += op->are[t];
      dim[4 * (size_t)g + c] += op->aim[t];
    } else {
      if ((meta[g] >> (16 + c)) & 1u) return false;  // two off-diagonal entries on one pattern
      meta[g] |= 1u << (16 + c);
      meta[g] |= nc << (24 + 2 * c);
      ore[4 * (size_t)g + c] = op->are[t];
      oim[4 * (size_t)g + c] = op->aim[t];
    }
  }
  upload(op->pg_meta[form], meta, st);
  upload(op->pg_dre[form], dre, st);
  upload(op->pg_dim[form], dim, st);
  upload(op->pg_ore[form], ore, st);
  upload(op->pg_oim[form], oim, st);
  PairGroupsDev& d = op->pg[form];
  d.n_groups = ng;
  d.meta = op->pg_meta[form].as<uint32_t>();
  d.dre = op->pg_dre[form].as<double>();
  d.dim = op->pg_dim[form].as<double>();
  d.ore = op->pg_ore[form].as<double>();
  d.oim = op->pg_oim[form].as<double>();
  d.dconst_re = cre;
  d.dconst_im = cim;
  return true;
}

// Whether the pair-group kernel covers this operator and space.
bool fast_path_usable(const edhip_op* op) {
  if (!op->pairs_ok) return false;
  const GroupHost* gh = op->space->group.get();
  if (!gh) return true;
  return gh->mode == kGroupTorus && gh->dev.minrot_table && gh->Ly <= 8 && gh->n_cosets == 1;
}

FastOpDev make_fast(edhip_op* op, int transpose, const HashDev* hash) {
  FastOpDev F{};
  F.base = transpose ? op->dev_col : op->dev_row;
  F.pg = op->pg[transpose ? 1 : 0];
  if (hash) F.hash = *hash;
  uint32_t off = align_up(F.base.lay.total, 16);
  F.pg_lds = off;
  off += (uint32_t)(kMaxPairGroups * 4) + (uint32_t)(F.pg.n_groups * 32 * 4);
  off = align_up(off, 16);
  F.base.w_lds = off;
  off += F.base.w_bytes;
  F.lds_total = align_up(off, 16);
  if (F.lds_total > 64 * 1024) fail(EDHIP_ERR_ARGUMENT, "pair tables exceed the 64 KiB LDS budget");
  return F;
}

// y = (accumulate ? y : 0) + A x (or x A), choosing the kernel and lookup.
// Per-row diagonal elements, computed once per operator when the operator
// has more than one diagonal term and the array (8 or 16 bytes per row) stays
// under 16 GiB; the matvec then reads one number per row instead of looping
// over the diagonal terms.  Returns false when not cached.
bool ensure_diag(edhip_op* op) {
  if (op->diag_tried) return op->diag_re.p != nullptr;
  op->diag_tried = true;
  const int64_t D = op->space->dim;
  if (op->n_diag < 2 || D == 0) return false;
  bool cx = false;
  for (int64_t t : op->diag_ids) cx = cx || op->aim[t] != 0.0;
  const uint64_t bytes = (uint64_t)D * 8ull * (cx ? 2ull : 1ull);
  if (bytes > (16ull << 30)) return false;
  void* pr = nullptr;
  void* pi = nullptr;
  if (hipMalloc(&pr, (size_t)D * 8) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (cx && hipMalloc(&pi, (size_t)D * 8) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(pr);
    return false;
  }
  op->diag_re.p = pr;
  op->diag_re.bytes = (size_t)D * 8;
  op->diag_im.p = pi;
  op->diag_im.bytes = pi ? (size_t)D * 8 : 0;
  const OpDev& d = op->dev_row;
  HIP_OR_THROW(launch_diag_values(d, op->diag_re.as<double>(), op->diag_im.as<double>(), op->ctx->stream));
  return true;
}

// Points d at the spread of x over the word positions, filled with x, when
// the operator's lookup strategy calls for it: lookup 4, and auto on full
// sectors, whose other lookup is rank arithmetic, a chain of table reads per
// image (4x7 J1-J2: FP32 14.9 -> 4.6 ms, FP64 14.9 -> 8.1 ms), where the
// spread stays cache-sized, and where it is dense (at most kXMapDenseRatio32
// positions per basis word for 4-byte elements, 6x6 J1-J2 FP32: 118 -> 87
// ms; kXMapDenseRatio for wider ones, the colex form of single-popcount
// spaces: 6x6 FP64 127 -> 91 ms, where the sparser plain form, 18 positions
// per word, is no faster than the blocked rank index) and x itself does not
// fit the last-level cache.  scaled: the spread holds
// x / sqrt(N) on reduced spaces (the products then weigh an image without
// its orbit size, and their canonicalisers skip counting the stabiliser).
// Returns whether it did.
bool attach_spread(edhip_op* op, int prec, OpDev& d, const void* x, hipStream_t st, bool scaled = false) {
  const edhip_space* sp = op->space;
  const bool rank_mode = sp->is_sector && !sp->group;
  const int idx_lookup = (op->strat_lookup == 2) ? 0 : op->strat_lookup;
  const size_t es = elem_size(prec);
  // a dense spread only for vectors larger than the last-level cache: where x
  // itself fits it, the rank index's x gathers hit it and the index was
  // measured faster (6x6 space-group A1 sector, 31.6M words, 126 MB of FP32
  // x: 57.8 ms against 67.0 ms through its spread)
  const uint64_t ratio = es == 4 ? kXMapDenseRatio32 : kXMapDenseRatio;
  const bool dense = sp->ensure_xtop() <= ratio * (uint64_t)sp->dim && (uint64_t)sp->dim * es > kXMapCachedBytes;
  const bool want = idx_lookup == 4 ||
                    (idx_lookup == 0 && (rank_mode || sp->ensure_xtop() * es <= kXMapCachedBytes || dense));
  if (!want || !d.codes_lds) return false;
  const edhip_space::XMapHost* X = sp->ensure_xmap(es);
  if (!X) return false;
  const uint16_t* orbit = (scaled && sp->group) ? sp->orbit.as<uint16_t>() : nullptr;
  HIP_OR_THROW(launch_xmap_fill(prec, X->dev, sp->keys.as<uint64_t>(), sp->dim, x, orbit, st));
  d.xm = X->dev;
  d.xm.scaled = orbit ? 1 : 0;
  // every image of a term that keeps the popcount has the space's popcount:
  // the products skip the test
  if (op->keeps_popcount) d.xm.pop = -1;
  return true;
}

// The eigensolver assembles the rows of a reduced space over 2^22 words for
// the time of a solve (kernel auto) when their arrays take at most this much
// and fit beside its workspace (6x6 space-group A1 sector: 9.6 GB built in
// 0.1 s, products 57.9 -> 9.8 ms, mixed time to E0 3.40 -> 0.61 s)
constexpr uint64_t kAutoAsmMaxBytes = 16ull << 30;

// Lanes per row and entries per lane per batch of the assembled product
// (0: chosen from the mean entries per row)
std::atomic<int> g_asm_lanes{0}, g_asm_batch{0};
// Consecutive blocks per XCD and window in the products (xcd_chunked_block):
// 32 measured at 6x6 against the plain order (1), same results bitwise:
// term list FP32 82.3 / 83.1 ms, FP64 85.8 / 88.1 ms; assembled rows FP32
// 69.1 / 70.2 ms, FP64 82.0 / 82.7 ms (2 to 64 in between or level)
std::atomic<int> g_xcd_chunk{32};

void asm_shape(const edhip_op* op, int& lanes, int& batch) {
  const double mean = op->space->dim ? (double)op->asm_nnz / (double)op->space->dim : 0.0;
  lanes = g_asm_lanes.load();
  if (!lanes) lanes = mean < 12.0 ? 4 : mean < 40.0 ? 8 : mean < 160.0 ? 16 : 32;
  batch = g_asm_batch.load();
  if (!batch) batch = 8;
}

// Assembled rows of the row form (strategy 3): the entry count of every row
// (matched terms), their offsets, then the entries written by the matvec's own
// matching and canonicalisation.  Weight classes: the distinct amplitudes of
// the terms times the distinct character phases of the group.  Leaves nothing
// allocated when it fails (EDHIP_ERR_MEMORY when the arrays do not fit), or
// when the arrays would take more than max_bytes (then returns false).
bool ensure_assembled(edhip_op* op, uint64_t max_bytes = ~0ull) {
  if (op->asm_built) return true;
  const auto t0 = std::chrono::steady_clock::now();
  const edhip_space* sp = op->space;
  hipStream_t st = op->ctx->stream;
  const int64_t D = sp->dim;
  int col_bits = 1;
  while (col_bits < 32 && (1ull << col_bits) < (uint64_t)D) ++col_bits;
  if (col_bits > 31) fail(EDHIP_ERR_ARGUMENT, "assembled rows hold at most 2^31 basis states");
  OpDev d = op->dev_row;
  d.ix = sp->index_dev(0);
  const int nt = d.n_mv ? d.n_mv : d.n_off;
  std::vector<OffTerm> terms((size_t)nt);
  if (nt)
    HIP_OR_THROW(hipMemcpy(terms.data(), d.n_mv ? d.mv_off : d.off, (size_t)nt * sizeof(OffTerm), hipMemcpyDeviceToHost));
  std::vector<std::complex<double>> amps, phs;
  std::vector<uint32_t> acls((size_t)std::max(nt, 1), 1u);
  for (int t = 0; t < nt; ++t) {
    if (terms[(size_t)t].match == terms[(size_t)t].image) continue;  // padding: never matches
    const std::complex<double> a(terms[(size_t)t].are, terms[(size_t)t].aim);
    const size_t k = (size_t)(std::find(amps.begin(), amps.end(), a) - amps.begin());
    if (k == amps.size()) amps.push_back(a);
    acls[(size_t)t] = 1u + (uint32_t)k;
  }
  const GroupHost* g = sp->group.get();
  const bool use_ph = g && !g->dev.unit_phases;
  std::vector<uint32_t> pcls;
  if (use_ph) {
    pcls.resize((size_t)g->G);
    for (int h = 0; h < g->G; ++h) {
      const std::complex<double> p = std::conj(g->chi[(size_t)g->inv[(size_t)h]]);
      const size_t k = (size_t)(std::find(phs.begin(), phs.end(), p) - phs.begin());
      if (k == phs.size()) phs.push_back(p);
      pcls[(size_t)h] = (uint32_t)k;
    }
  }
  const int n_ph = use_ph ? (int)phs.size() : 1;
  const int n_cls = 1 + (int)std::max<size_t>(amps.size(), 1) * n_ph;
  const int max_cls = (32 - col_bits) >= 8 ? kAsmMaxClasses : (1 << (32 - col_bits));
  if (n_cls > max_cls)
    fail(EDHIP_ERR_ARGUMENT, "assembled rows: " + std::to_string(n_cls) + " element weights exceed the " +
                                 std::to_string(max_cls) + " classes an entry can name at this dimension");
  std::vector<double> w(2 * (size_t)n_cls, 0.0);
  for (size_t a = 0; a < amps.size(); ++a)
    for (int p = 0; p < n_ph; ++p) {
      const std::complex<double> v = use_ph ? amps[a] * phs[(size_t)p] : amps[a];
      w[1 + a * n_ph + p] = v.real();
      w[(size_t)n_cls + 1 + a * n_ph + p] = v.imag();
    }
  try {
    {
      DevBuf cnt, tmp;
      cnt.alloc((size_t)(D + 1) * 4);
      HIP_OR_THROW(hipMemsetAsync(cnt.p, 0, (size_t)(D + 1) * 4, st));
      HIP_OR_THROW(launch_asm_count(d, cnt.as<uint32_t>(), st));
      op->asm_rowptr.alloc((size_t)(D + 1) * 8);
      size_t tb = 0;
      HIP_OR_THROW(exclusive_scan_u32_to_u64(cnt.as<uint32_t>(), op->asm_rowptr.as<uint64_t>(), (uint64_t)D + 1,
                                             nullptr, tb, st));
      tmp.alloc(tb);
      HIP_OR_THROW(exclusive_scan_u32_to_u64(cnt.as<uint32_t>(), op->asm_rowptr.as<uint64_t>(), (uint64_t)D + 1, tmp.p,
                                             tb, st));
      HIP_OR_THROW(hipMemcpyAsync(&op->asm_nnz, op->asm_rowptr.as<uint64_t>() + D, 8, hipMemcpyDeviceToHost, st));
      HIP_OR_THROW(hipStreamSynchronize(st));
    }
    if (op->asm_nnz * 4 + op->asm_rowptr.bytes > max_bytes) {
      op->release_assembled();
      return false;
    }
    injected_alloc_failure(kAllocAssembled);
    op->asm_ent.alloc((size_t)op->asm_nnz * 4);
    upload(op->asm_acls, acls, st);
    if (use_ph) upload(op->asm_pcls, pcls, st);
    upload(op->asm_w, w, st);
    AsmFillDev f{op->asm_rowptr.as<uint64_t>(), op->asm_ent.as<uint32_t>(), op->asm_acls.as<uint32_t>(),
                 use_ph ? op->asm_pcls.as<uint32_t>() : nullptr, n_ph, col_bits};
    HIP_OR_THROW(launch_asm_fill(d, f, st));
    HIP_OR_THROW(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipGetLastError();
    op->release_assembled();
    throw;
  }
  op->asm_dev = AsmDev{op->asm_rowptr.as<uint64_t>(), op->asm_ent.as<uint32_t>(), op->asm_w.as<double>(),
                       op->asm_w.as<double>() + n_cls, n_cls, col_bits};
  op->asm_built = true;
  op->asm_build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

// Row-form product from the assembled rows: x / sqrt(N) first (reduced spaces),
// then the streaming product.
void asm_matvec(edhip_op* op, int prec, int accumulate, const void* x, void* y) {
  hipStream_t st = op->ctx->stream;
  const edhip_space* sp = op->space;
  OpDev d = op->dev_row;
  if (ensure_diag(op)) {
    d.diag_re = op->diag_re.as<double>();
    d.diag_im = op->diag_im.as<double>();
  }
  d.xcd_chunk = g_xcd_chunk.load();
  const size_t es = elem_size(prec);
  if (op->ctx->kernel_timing) HIP_OR_THROW(hipEventRecord(op->ctx->ev[14], st));
  const void* xt = x;
  if (d.orbit) {
    op->asm_xt.ensure((size_t)sp->dim * es);
    HIP_OR_THROW(launch_asm_scale(prec, sp->dim, d.orbit, x, op->asm_xt.p, st));
    xt = op->asm_xt.p;
  }
  int lanes, batch;
  asm_shape(op, lanes, batch);
  HIP_OR_THROW(launch_asm_spmv(d, op->asm_dev, prec, lanes, batch, accumulate, x, xt, y, st));
  if (op->ctx->kernel_timing) {
    HIP_OR_THROW(hipEventRecord(op->ctx->ev[15], st));
    HIP_OR_THROW(hipEventSynchronize(op->ctx->ev[15]));
    float f = 0.f;
    HIP_OR_THROW(hipEventElapsedTime(&f, op->ctx->ev[14], op->ctx->ev[15]));
    op->kt_ms += f;
    ++op->kt_n;
  }
}

// Term-list kernel over the output rows [row_begin, row_end) (y holds them
// from y[0]); x is the whole input vector.
void term_list_matvec(edhip_op* op, int prec, int transpose, int accumulate, const void* x, void* y,
                      int64_t row_begin, int64_t row_end) {
  hipStream_t st = op->ctx->stream;
  const edhip_space* sp = op->space;
  const bool rank_mode = sp->is_sector && !sp->group;
  const int idx_lookup = (op->strat_lookup == 2) ? 0 : op->strat_lookup;  // the hash applies to the pair kernel only
  OpDev d = transpose ? op->dev_col : op->dev_row;
  d.ix = sp->index_dev(idx_lookup == 4 ? 0 : idx_lookup);
  d.xcd_chunk = g_xcd_chunk.load();
  d.row_begin = row_begin;
  d.dim = row_end;
  if (ensure_diag(op)) {
    d.diag_re = op->diag_re.as<double>();
    d.diag_im = op->diag_im.as<double>();
  }
  op->used_kernel = 1;
  op->used_lookup = rank_mode ? 1 : (d.ix.bm_bits ? 3 : 1);
  if (attach_spread(op, prec, d, x, st, true)) op->used_lookup = 4;
  if (idx_lookup == 4 && !d.xm.values) fail(EDHIP_ERR_ARGUMENT, "x cannot be spread over the words of this space");
  if (op->ctx->kernel_timing) HIP_OR_THROW(hipEventRecord(op->ctx->ev[14], st));
  HIP_OR_THROW(launch_matvec(d, prec, transpose ? 1 : 0, accumulate ? 1 : 0, x, y, st));
  if (op->ctx->kernel_timing) {
    HIP_OR_THROW(hipEventRecord(op->ctx->ev[15], st));
    HIP_OR_THROW(hipEventSynchronize(op->ctx->ev[15]));
    float f = 0.f;
    HIP_OR_THROW(hipEventElapsedTime(&f, op->ctx->ev[14], op->ctx->ev[15]));
    op->kt_ms += f;
    ++op->kt_n;
  }
}

void op_matvec(edhip_op* op, int prec, int transpose, int accumulate, const void* x, void* y) {
  hipStream_t st = op->ctx->stream;
  const edhip_space* sp = op->space;
  // assembled rows on request (row form; the column form runs the term list)
  if (op->strat_kernel == 3 && !transpose) {
    ensure_assembled(op);
    op->used_kernel = 3;
    op->used_lookup = 0;
    asm_matvec(op, prec, accumulate, x, y);
    return;
  }
  // auto: the term-list kernel (measured faster on the 6x6 J1-J2 sector: its
  // lanes walk the terms in lock-step, which keeps their lookups in the same
  // cache lines); the pair-group kernel on request
  bool fast = fast_path_usable(op) && op->strat_kernel == 2;
  if (op->strat_kernel == 2 && !fast_path_usable(op))
    fail(EDHIP_ERR_ARGUMENT, "the pair-group kernel does not cover this operator or symmetry group");
  const bool rank_mode = sp->is_sector && !sp->group;
  const int idx_lookup = (op->strat_lookup == 2) ? 0 : op->strat_lookup;  // the hash applies to the pair kernel only
  if (!fast) {
    term_list_matvec(op, prec, transpose, accumulate, x, y, 0, sp->dim);
    return;
  }
  bool use_hash = false;
  if (!rank_mode) {
    use_hash = op->strat_lookup == 2;
  }
  op->used_kernel = 2;
  if (use_hash) {
    op->used_lookup = 2;
    const HashDev& h = sp->ensure_hash(elem_size(prec));
    HIP_OR_THROW(launch_hash_scatter(prec, sp->dim, x, sp->hash_slot_of.as<uint32_t>(), h, st));
    FastOpDev F = make_fast(op, transpose, &h);
    HIP_OR_THROW(launch_matvec_pairs(F, prec, transpose, accumulate, x, y, st));
  } else {
    FastOpDev F = make_fast(op, transpose, nullptr);
    F.base.ix = sp->index_dev(idx_lookup == 4 ? 0 : idx_lookup);
    op->used_lookup = rank_mode ? 1 : (F.base.ix.bm_bits ? 3 : 1);
    HIP_OR_THROW(launch_matvec_pairs(F, prec, transpose, accumulate, x, y, st));
  }
}

}  // namespace

// ============================================================================
// Lanczos (thick restart, full reorthogonalisation) and its host linear algebra
// ============================================================================
namespace {

// Eigen-decomposition of the symmetric m x m matrix T (row-major).
void sym_eig(int m, const std::vector<double>& T, bool tridiagonal, std::vector<double>& w, std::vector<double>& Z) {
  std::vector<double> d(m), e(m, 0.0);
  if (tridiagonal) {
    for (int i = 0; i < m; ++i) d[i] = T[(size_t)i * m + i];
    for (int i = 0; i + 1 < m; ++i) e[i] = T[(size_t)i * m + i + 1];
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[(size_t)i * m + i] = 1.0;
  } else {
    std::vector<double> A = T;
    householder_tridiag(m, A, Z, d, e);
  }
  if (!tridiag_ql(m, d, e, Z)) fail(EDHIP_ERR_NOCONV, "tridiagonal eigensolver did not converge");
  w = d;
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double sec() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// Thick-restart Lanczos on one operator in one storage precision.
// Ritz vectors kept at a thick restart of an mm-vector basis when n_ret
// values are wanted (nondecreasing in both).
inline int keep_count(int mm, int n_ret) {
  return std::max(1, std::min(mm - 1, std::max(n_ret + 2, std::min(mm / 3, n_ret + 16))));
}

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The context's workspace; when it does not fit, the spreads of x that the
// operator's space keeps (caches the next product rebuilds if memory allows)
// are freed and the allocation is tried once more.
char* krylov_arena_for(edhip_op* op, size_t bytes) {
  try {
    return op->ctx->krylov_arena(bytes);
  } catch (const Error& e) {
    if (e.code != EDHIP_ERR_MEMORY) throw;
    (void)hipGetLastError();
    op->space->release_xmaps();
  }
  try {
    return op->ctx->krylov_arena(bytes);
  } catch (const Error& e) {
    // the workspace takes precedence over the operator's assembled rows: they
    // are dropped and the products run matrix-free (strategy back to auto)
    if (e.code != EDHIP_ERR_MEMORY || !op->asm_built) throw;
    (void)hipGetLastError();
    op->release_assembled();
    op->strat_kernel = 0;
    return op->ctx->krylov_arena(bytes);
  }
}

// Assembled rows for the time of a solve (eigensolver, MINRES) with the
// operator's kernel strategy on auto: reduced spaces over 2^22 words whose
// arrays take at most kAutoAsmMaxBytes and fit beside the solver's workspace
// (already allocated) with 4 GiB to spare.  Spaces whose rows an entry cannot
// hold, or that do not fit, stay matrix-free.  The destructor releases the
// arrays and puts the strategy back to auto.
struct SolveAssembly {
  edhip_op* op = nullptr;
  bool begin(edhip_op* o) {
    if (o->strat_kernel != 0 || o->space->dim <= ((int64_t)1 << 22) || !o->space->group) return false;
    size_t free_b = 0, total_b = 0;
    HIP_OR_THROW(hipMemGetInfo(&free_b, &total_b));
    const uint64_t margin = 4ull << 30;
    const uint64_t room = free_b > margin ? (uint64_t)free_b - margin : 0;
    try {
      if (!ensure_assembled(o, std::min<uint64_t>(kAutoAsmMaxBytes, room))) return false;
    } catch (const Error& e) {
      // no room, or rows an entry cannot hold (2^31 states or more, more
      // weight classes than its bits name)
      if (e.code != EDHIP_ERR_MEMORY && e.code != EDHIP_ERR_ARGUMENT) throw;
      (void)hipGetLastError();
      return false;
    }
    o->strat_kernel = 3;
    op = o;
    return true;
  }
  ~SolveAssembly() {
    if (op) {
      op->release_assembled();
      op->strat_kernel = 0;
    }
  }
};

struct TRLanczos {
  edhip_op* op;
  int prec;        // storage precision of the Krylov vectors
  int64_t n;
  int64_t ld = 0;  // stride of the basis vectors (n rounded up to 4 elements: 16-byte aligned vectors)
  int m;           // basis size
  void* Vp = nullptr;    // (m + 1) vectors of stride ld
  void* tmpp = nullptr;  // restart combination scratch
  size_t tmp_bytes = 0;
  DevBuf coef_d;   // device coefficients
  const void* locked = nullptr;  // locked (deflated) vectors, same precision
  int n_locked = 0;
  int matvecs = 0;
  double t_matvec = 0.0;
  // device time (HIP events) and algorithmic bytes of the orthogonalisation
  // kernels (dot products and updates against the basis), large bases only
  double t_ortho = 0.0, ortho_bytes = 0.0;
  int iterations = 0, restarts = 0;

  size_t es() const { return elem_size(prec); }
  void set_size(int64_t dim) {
    n = dim;
    ld = (dim + 3) / 4 * 4;
  }
  void* vec(int j) const { return (char*)Vp + (size_t)j * ld * es(); }

  void matvec(const void* x, void* y) {
    edhip_ctx* ctx = op->ctx;
    Timer t;
    op_matvec(op, prec, 0, 0, x, y);
    ++matvecs;
    if (n > (1 << 22)) {  // time only the large products (avoids a sync per small step)
      ctx->sync();
      t_matvec += t.sec();
    }
  }

  // <x, y> in double
  std::complex<double> dot(const void* x, const void* y) {
    edhip_ctx* ctx = op->ctx;
    double* part = ctx->partial_buf(2 * (size_t)vec_partial_blocks(n));
    double* out = ctx->small_buf(2);
    HIP_OR_THROW(vec_dot(prec, n, x, y, part, out, ctx->stream));
    double h[2];
    HIP_OR_THROW(hipMemcpyAsync(h, out, 16, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    return {h[0], h[1]};
  }

  // Reorthogonalisation: 0 = one classical Gram-Schmidt pass, repeated when
  // the pass removed most of w; 1 = always two passes.
  int reorth = 0;
  // The second pass runs when ||w - B c||^2 < kReorthKeep ||w||^2, i.e. the
  // projection cancelled more than 1 - 1/kappa of w with kappa = 10 (a looser
  // threshold than the Daniel-Gragg-Kaufman-Stewart 1/sqrt(2)): one pass leaves
  // components along B of order kappa * eps relative to the result.  For
  // orthonormal B, ||w - B c||^2 = ||w||^2 - ||c||^2 needs no extra pass.
  static constexpr double kReorthKeep = 1e-2;
  int second_passes = 0;
  // Partial reorthogonalisation (Simon, "The Lanczos algorithm with partial
  // reorthogonalization", Math. Comp. 42 (1984) 115-142) in the first cycle:
  // each step orthogonalises against the two latest vectors only, and the
  // omega recurrence estimates the new vector's overlaps with the older ones
  // from alpha and beta; when one exceeds sqrt(eps) (eps: the storage
  // precision's unit roundoff) the vector, and the next one, are
  // orthogonalised against the whole basis.  The basis stays semi-orthogonal,
  // which keeps the Ritz values accurate to the working precision.  After a
  // thick restart, an injected direction or with locked vectors every step
  // orthogonalises against the whole basis.
  bool partial = false;
  int full_steps = 0;  // steps orthogonalised against the whole basis

  // w -= B (B^H w) over vectors B_0..B_{k-1} (stride ldb), once or twice;
  // returns the first-pass coefficients (host).
  std::vector<double> orthogonalize(const void* B, int64_t ldb, int k, void* w) {
    edhip_ctx* ctx = op->ctx;
    std::vector<double> c1(2 * (size_t)k, 0.0);
    if (k <= 0) return c1;
    const bool adaptive = reorth == 0;
    double* part = ctx->partial_buf(std::max<size_t>((size_t)2 * k * 1024, 2 * (size_t)vec_partial_blocks(n)));
    coef_d.ensure((size_t)(2 * k + 2) * sizeof(double));
    double* nn_d = coef_d.as<double>() + 2 * k;
    std::vector<double> hc(2 * (size_t)k + 2, 0.0);
    // device time of the kernels between two events, and their algorithmic
    // bytes: per pass the basis and w read by the dot products, then again by
    // the update, which writes w (plus one read of w by its norm)
    const bool timed = n > (1 << 22);
    auto t_begin = [&] {
      if (timed) HIP_OR_THROW(hipEventRecord(ctx->ev_ortho[0], ctx->stream));
    };
    auto t_end = [&](double vectors) {
      if (!timed) return;
      HIP_OR_THROW(hipEventRecord(ctx->ev_ortho[1], ctx->stream));
      HIP_OR_THROW(hipEventSynchronize(ctx->ev_ortho[1]));
      float ms = 0.f;
      HIP_OR_THROW(hipEventElapsedTime(&ms, ctx->ev_ortho[0], ctx->ev_ortho[1]));
      t_ortho += 1e-3 * ms;
      ortho_bytes += vectors * (double)n * (double)es();
    };
    const double pass_vectors = 2.0 * (k + 1) + 1.0;
    t_begin();
    if (adaptive) HIP_OR_THROW(vec_dot(prec, n, w, w, part, nn_d, ctx->stream));
    HIP_OR_THROW(vec_multi_dot(prec, n, k, B, ldb, w, part, coef_d.as<double>(), ctx->stream));
    HIP_OR_THROW(hipMemcpyAsync(hc.data(), coef_d.p, hc.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_THROW(vec_multi_axpy(prec, n, k, B, ldb, coef_d.as<double>(), w, ctx->stream));
    t_end(pass_vectors + (adaptive ? 1.0 : 0.0));
    ctx->sync();
    std::copy(hc.begin(), hc.begin() + 2 * (size_t)k, c1.begin());
    bool again = !adaptive;
    if (adaptive) {
      double cc = 0.0;
      for (int i = 0; i < 2 * k; ++i) cc += hc[i] * hc[i];
      const double ww = hc[2 * (size_t)k];
      again = ww - cc < kReorthKeep * ww;
    }
    if (again) {
      ++second_passes;
      t_begin();
      HIP_OR_THROW(vec_multi_dot(prec, n, k, B, ldb, w, part, coef_d.as<double>(), ctx->stream));
      HIP_OR_THROW(vec_multi_axpy(prec, n, k, B, ldb, coef_d.as<double>(), w, ctx->stream));
      t_end(pass_vectors);
      ctx->sync();
    }
    return c1;
  }

  void scale(void* x, double s) { HIP_OR_THROW(vec_axpby(prec, n, s, 0.0, x, 0.0, 0.0, x, op->ctx->stream)); }

  // Convergence rule of a Ritz pair (theta, residual r):
  //   criterion 0: r <= tol * max(1, |theta|)
  //   criterion 1: as 0, or (r^2 / gap <= tol * max(1, |theta|) and theta moved
  //                by at most tol * max(1, |theta|) since the previous check),
  //                gap = distance to the next Ritz value (or gap_hint): the
  //                Hermitian bound |theta - lambda| <= r^2 / gap on the value.
  int criterion = 0;
  double gap_hint = 0.0;
  // stationary: criterion 1 also asks theta to be unchanged since the previous
  // check (the Lanczos residual estimate of an FP32 basis is not a true
  // residual); false where beta is the true FP64 residual ||H y - theta y||,
  // which bounds the eigenvalue error by itself
  bool stationary = true;
  bool accept_invariant = false;

  // Random directions after an invariant subspace (beta ~ 0): the Krylov
  // space of the start vector misses every eigenvector orthogonal to it
  // (degenerate copies included), so, as ARPACK's getv0 does, the iteration
  // goes on from a random unit vector orthogonal to the locked vectors and to
  // the basis, with a zero coupling.
  uint64_t seed = 0;
  int injections = 0;

  void inject(int nb, void* w) {
    ++injections;
    HIP_OR_THROW(vec_gaussian(prec, n, seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)injections), w, op->ctx->stream));
    for (int pass = 0; pass < 2; ++pass) {
      if (n_locked) orthogonalize(locked, n, n_locked, w);
      orthogonalize(Vp, ld, nb, w);
    }
    const double nn = std::sqrt(std::max(dot(w, w).real(), 0.0));
    if (!(nn > 0.0)) fail(EDHIP_ERR_NOCONV, "no direction left orthogonal to the Krylov basis");
    scale(w, 1.0 / nn);
  }

  // Runs the iteration from the unit vector in vec(0).  Returns the lowest
  // n_ret >= nev Ritz values (convergence is required of the lowest nev);
  // Ritz vectors of the lowest nev (if out != nullptr) are written in
  // precision out_prec with stride n.  res: residual norms of the lowest nev.
  // Convergence is accepted at a step whose beta is not tiny, once the basis
  // spans the whole (deflated) space, or at the end of a full basis holding
  // two or more closed Krylov blocks; after a random direction was injected
  // it is tested only when the basis is full, as ARPACK tests at the end of
  // an Arnoldi cycle, so that the eigenvalues of every Krylov block built so
  // far (degenerate copies among them) are in the projected matrix.
  bool run(int nev, double tol, int max_restarts, int check_every, std::vector<double>& theta, std::vector<double>& res,
           void* out, int out_prec, int n_ret = 0) {
    edhip_ctx* ctx = op->ctx;
    const int64_t dim = n;
    int mm = (int)std::min<int64_t>(m, dim - n_locked);
    if (mm < 1) mm = 1;
    nev = std::min(nev, mm);
    n_ret = std::max(nev, std::min(n_ret, mm));
    std::vector<double> T((size_t)mm * mm, 0.0);
    int k = 0;  // kept Ritz vectors after a restart
    bool converged = false;
    std::vector<double> w, Z, prev;
    // omega recurrence: om_cur[q] ~ <v_j, v_q>, om_prev[q] ~ <v_{j-1}, v_q>
    const double eps = (prec == EDHIP_F32 || prec == EDHIP_C64) ? 0x1p-24 : 0x1p-53;
    const double om_thr = std::sqrt(eps);
    std::vector<double> om_prev, om_cur(1, 1.0), om_new;
    bool om_force = false;
    int j = 0;
    double beta_last = 0.0;
    int size = 0;  // current basis size used for the final Ritz step
    bool arrow = false;  // T is arrowhead + tridiagonal after a thick restart, else tridiagonal
    auto ritz = [&](int sz) {  // eigenpairs of the leading sz x sz block of T
      std::vector<double> Tj((size_t)sz * sz);
      for (int a = 0; a < sz; ++a)
        for (int b = 0; b < sz; ++b) Tj[(size_t)a * sz + b] = T[(size_t)a * mm + b];
      sym_eig(sz, Tj, !arrow, w, Z);
    };
    for (;;) {
      arrow = (k > 0);
      for (j = k; j < mm; ++j) {
        void* vj = vec(j);
        void* wv = vec(j + 1);
        matvec(vj, wv);
        ++iterations;
        if (n_locked) orthogonalize(locked, n, n_locked, wv);
        const bool pro = partial && k == 0 && restarts == 0 && injections == 0 && n_locked == 0;
        if (pro) {  // the three-term recurrence: v_{j-1} and v_j only
          const int lo = j > 0 ? j - 1 : 0;
          std::vector<double> c = orthogonalize(vec(lo), ld, j + 1 - lo, wv);
          T[(size_t)j * mm + j] = c[2 * (size_t)(j - lo)];
        } else {
          std::vector<double> c = orthogonalize(Vp, ld, j + 1, wv);
          T[(size_t)j * mm + j] = c[2 * (size_t)j];
          ++full_steps;
        }
        std::complex<double> nn = dot(wv, wv);
        double beta = std::sqrt(std::max(nn.real(), 0.0));
        if (pro && beta > 0.0) {
          const double worst = omega_step(
              j, [&](int q) { return T[(size_t)q * mm + q]; }, [&](int q) { return T[(size_t)(q + 1) * mm + q]; },
              beta, om_cur, om_prev, eps, om_new);
          if (om_force || worst > om_thr) {
            orthogonalize(Vp, ld, j + 1, wv);
            ++full_steps;
            nn = dot(wv, wv);
            beta = std::sqrt(std::max(nn.real(), 0.0));
            for (int q = 0; q <= j; ++q) om_new[(size_t)q] = eps;
            om_force = !om_force;  // the next vector too
          }
          om_prev.swap(om_cur);
          om_cur.swap(om_new);
        }
        size = j + 1;
        const bool last = (j + 1 == mm);
        const bool exhausted = (int64_t)(j + 1) >= dim - n_locked;
        const bool tiny = beta <= 1e-13 * std::max(1.0, std::fabs(T[(size_t)j * mm + j]));
        // settle: a refinement that started on an eigenvector (to working
        // precision) takes it; everything else continues past an invariant
        // subspace from a random direction (unit, orthogonal), zero coupling
        const bool settle = tiny && !exhausted && accept_invariant && j + 1 >= nev;
        const bool invariant = tiny && !exhausted && !settle;
        if (invariant) {
          inject(j + 1, wv);
          beta = 0.0;
        }
        beta_last = beta;
        const bool periodic = (j + 1 - k) % std::max(1, check_every) == 0 && j + 1 >= nev;
        // at the end of the basis also after an invariant step once two or more
        // Krylov blocks are complete (an operator whose blocks all close at
        // once, e.g. zero, has only such steps)
        const bool check = exhausted || settle || (last && (!invariant || injections >= 2)) ||
                           (!invariant && injections == 0 && periodic);
        if (check) {
          ritz(j + 1);
          bool ok = (j + 1 >= nev);
          res.assign(nev, 0.0);
          for (int i = 0; i < nev && i <= j; ++i) {
            res[i] = std::fabs(beta * Z[(size_t)j * (j + 1) + i]);
            const double scale = tol * std::max(1.0, std::fabs(w[i]));
            bool ok_i = res[i] <= scale;
            if (!ok_i && criterion == 1 && (prev.size() > (size_t)i || !stationary)) {
              double gap = (i + 1 <= j) ? (w[i + 1] - w[i]) : gap_hint;
              if (gap_hint > 0.0 && (i + 1 > j || gap > gap_hint)) gap = gap_hint;
              gap = std::max(gap, 1e-12);
              ok_i = (res[i] * res[i] / gap <= scale) && (!stationary || std::fabs(w[i] - prev[i]) <= scale);
            }
            if (!ok_i) ok = false;
          }
          prev.assign(w.begin(), w.begin() + std::min<size_t>(w.size(), (size_t)nev));
          if (ok) {
            converged = true;
            break;
          }
          if (exhausted) break;  // the whole space is in T: nothing left to add
          if (settle) {
            inject(j + 1, wv);
            beta = beta_last = 0.0;
          }
        }
        if (j + 1 < mm) {
          T[(size_t)(j + 1) * mm + j] = T[(size_t)j * mm + j + 1] = beta;
          if (beta > 0.0) scale(wv, 1.0 / beta);
        }
      }
      if (converged || size >= dim - n_locked) break;
      // thick restart
      if (restarts >= max_restarts) break;
      ++restarts;
      ritz(mm);
      int keep = keep_count(mm, n_ret);
      // new basis: V'_i = sum_j Z[j][i] V_j (i < keep), V'_keep = normalised residual
      std::vector<double> coef((size_t)mm * keep);
      for (int a = 0; a < mm; ++a)
        for (int i = 0; i < keep; ++i) coef[(size_t)a * keep + i] = Z[(size_t)a * mm + i];
      coef_d.ensure(coef.size() * sizeof(double));
      HIP_OR_THROW(hipMemcpyAsync(coef_d.p, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream));
      // V[:, :keep] = V Z[:, :keep] in place, a chunk of rows at a time through
      // the scratch (each row of the result needs only that row of V)
      int64_t chunk = (int64_t)std::min<size_t>((size_t)n, tmp_bytes / ((size_t)keep * es()));
      if (chunk < n) chunk &= ~(int64_t)3;  // chunks start on 16-byte boundaries
      if (chunk < 1) fail(EDHIP_ERR_ARGUMENT, "restart scratch too small");
      for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int64_t len = std::min<int64_t>(chunk, n - r0);
        char* vrow = (char*)Vp + (size_t)r0 * es();
        HIP_OR_THROW(vec_combine(prec, prec, len, mm, vrow, ld, keep, coef_d.as<double>(), tmpp, len, ctx->stream));
        HIP_OR_THROW(hipMemcpy2DAsync(vrow, (size_t)ld * es(), tmpp, (size_t)len * es(), (size_t)len * es(), keep,
                                      hipMemcpyDeviceToDevice, ctx->stream));
      }
      // residual direction: vec(mm), unnormalised with norm beta_last, or a
      // unit random direction injected after an invariant subspace (beta_last 0)
      if (beta_last > 0.0) scale(vec(mm), 1.0 / beta_last);
      HIP_OR_THROW(hipMemcpyAsync(vec(keep), vec(mm), (size_t)n * es(), hipMemcpyDeviceToDevice, ctx->stream));
      std::fill(T.begin(), T.end(), 0.0);
      for (int i = 0; i < keep; ++i) {
        T[(size_t)i * mm + i] = w[i];
        double s = beta_last * Z[(size_t)(mm - 1) * mm + i];
        T[(size_t)i * mm + keep] = T[(size_t)keep * mm + i] = s;
      }
      k = keep;
      size = keep;
      ctx->sync();
    }
    // Ritz values / vectors from the current (size x size) block
    if (!converged) {
      ritz(size);
      res.assign(nev, 0.0);
      for (int i = 0; i < nev && i < size; ++i) res[i] = std::fabs(beta_last * Z[(size_t)(size - 1) * size + i]);
    }
    const int nv = std::min(nev, size);
    theta.assign(w.begin(), w.begin() + std::min<size_t>((size_t)n_ret, w.size()));
    if (out) {
      int sz = size;
      std::vector<double> coef((size_t)sz * nev, 0.0);
      for (int a = 0; a < sz; ++a)
        for (int i = 0; i < nv; ++i) coef[(size_t)a * nev + i] = Z[(size_t)a * sz + i];
      coef_d.ensure(coef.size() * sizeof(double));
      HIP_OR_THROW(hipMemcpyAsync(coef_d.p, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream));
      HIP_OR_THROW(vec_combine(prec, out_prec, n, sz, Vp, ld, nev, coef_d.as<double>(), out, n, ctx->stream));
      ctx->sync();
    }
    (void)dim;
    return converged;
  }
};

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int edhip_create(int device, edhip_ctx** out) {
  if (!out) return EDHIP_ERR_ARGUMENT;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return EDHIP_ERR_DEVICE;
  if (device < 0 || device >= n) return EDHIP_ERR_ARGUMENT;
  auto* ctx = new edhip_ctx();
  ctx->device = device;
  int rc = guarded(ctx, [&] {
    HIP_OR_THROW(hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking));
    ctx->stream = ctx->own;
    for (auto& ev : ctx->ev) HIP_OR_THROW(hipEventCreate(&ev));
    for (auto& ev : ctx->ev_ortho) HIP_OR_THROW(hipEventCreate(&ev));
    // this device's address of the allocation-free term pass's scratch
    HIP_OR_THROW(term_expect_noalloc_setup(device));
  });
  if (rc != EDHIP_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return EDHIP_OK;
}

void edhip_destroy(edhip_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : ctx->ev_ortho)
    if (ev) (void)hipEventDestroy(ev);
  ctx->partial.release();
  ctx->small.release();
  if (ctx->own) (void)hipStreamDestroy(ctx->own);
  delete ctx;
}

const char* edhip_last_error(const edhip_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int edhip_set_stream(edhip_ctx* ctx, void* s) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    ctx->sync();
    ctx->stream = s ? (hipStream_t)s : ctx->own;
  });
}

int edhip_get_stream(const edhip_ctx* ctx, void** s) {
  if (!ctx || !s) return EDHIP_ERR_ARGUMENT;
  *s = (void*)ctx->stream;
  return EDHIP_OK;
}

int edhip_synchronize(edhip_ctx* ctx) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] { ctx->sync(); });
}

int edhip_malloc(edhip_ctx* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] { HIP_OR_THROW(hipMalloc(out, bytes ? bytes : 16)); });
}

int edhip_release_workspace(edhip_ctx* ctx) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    ctx->sync();
    ctx->krylov.release();
  });
}

int edhip_free(edhip_ctx* ctx, void* p) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    ctx->sync();
    HIP_OR_THROW(hipFree(p));
  });
}

int edhip_memcpy(edhip_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    HIP_OR_THROW(hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, ctx->stream));
    ctx->sync();
  });
}

// ---- spaces ----------------------------------------------------------------
int edhip_space_sector(edhip_ctx* ctx, int n_bits, int n_pop, const int32_t* pops, edhip_space** out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  *out = nullptr;
  auto sp = std::make_unique<edhip_space>();
  int rc = guarded(ctx, [&] {
    if (n_bits < 0 || n_bits > 63) fail(EDHIP_ERR_ARGUMENT, "type UInt64 not enough to represent the hilbert space");
    if (n_pop < 0 || (n_pop > 0 && !pops)) fail(EDHIP_ERR_ARGUMENT, "bad popcount list");
    std::vector<int> p;
    for (int i = 0; i < n_pop; ++i)
      if (pops[i] >= 0 && pops[i] <= n_bits && std::find(p.begin(), p.end(), (int)pops[i]) == p.end())
        p.push_back(pops[i]);
    std::sort(p.begin(), p.end());
    sp->ctx = ctx;
    sp->n_bits = n_bits;
    sp->is_sector = true;
    sp->popcounts = p;
    for (int q : p) sp->pop_allowed |= 1ull << q;
    sp->W_host = completion_counts(n_bits, p);
    sp->dim = (int64_t)sp->W_host[(size_t)n_bits * (n_bits + 1) + 0];
    upload(sp->W, sp->W_host, ctx->stream);
    ctx->sync();
  });
  if (rc == EDHIP_OK) *out = sp.release();
  return rc;
}

int edhip_space_from_list(edhip_ctx* ctx, int n_bits, int64_t n, const uint64_t* basis, edhip_space** out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  *out = nullptr;
  auto sp = std::make_unique<edhip_space>();
  int rc = guarded(ctx, [&] {
    if (n_bits < 0 || n_bits > 63) fail(EDHIP_ERR_ARGUMENT, "type UInt64 not enough to represent the hilbert space");
    if (n < 0 || (n > 0 && !basis)) fail(EDHIP_ERR_ARGUMENT, "bad basis list");
    for (int64_t i = 1; i < n; ++i)
      if (basis[i - 1] > basis[i]) fail(EDHIP_ERR_ARGUMENT, "vals must be sorted");
    for (int64_t i = 1; i < n; ++i)
      if (basis[i - 1] == basis[i]) fail(EDHIP_ERR_ARGUMENT, "vals contains duplicates");
    sp->ctx = ctx;
    sp->n_bits = n_bits;
    sp->dim = n;
    sp->keys.alloc((size_t)std::max<int64_t>(n, 1) * 8);
    if (n) HIP_OR_THROW(hipMemcpyAsync(sp->keys.p, basis, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
    sp->materialized = true;
    sp->build_index();
    ctx->sync();
  });
  if (rc == EDHIP_OK) *out = sp.release();
  return rc;
}

int edhip_space_sector_qn(edhip_ctx* ctx, int n_sites, const int32_t* bitoffsets, const int32_t* n_states, int qn_dim,
                          const int32_t* state_qns, int n_allowed, const int32_t* allowed_qns, edhip_space** out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  *out = nullptr;
  auto sp = std::make_unique<edhip_space>();
  int rc = guarded(ctx, [&] {
    if (n_sites < 0 || (n_sites > 0 && (!bitoffsets || !n_states || !state_qns)))
      fail(EDHIP_ERR_ARGUMENT, "bad site description");
    if (qn_dim < 1) fail(EDHIP_ERR_ARGUMENT, "quantum numbers need at least one component");
    if (n_allowed < 0 || (n_allowed > 0 && !allowed_qns)) fail(EDHIP_ERR_ARGUMENT, "bad allowed quantum numbers");
    const int n_bits = n_sites ? bitoffsets[n_sites] : 0;
    if (n_sites && bitoffsets[0] != 0) fail(EDHIP_ERR_ARGUMENT, "bit offsets must start at 0");
    if (n_bits < 0 || n_bits > 63) fail(EDHIP_ERR_ARGUMENT, "type UInt64 not enough to represent the hilbert space");
    std::vector<int> first_state(n_sites + 1, 0);
    for (int j = 0; j < n_sites; ++j) {
      const int wdt = bitoffsets[j + 1] - bitoffsets[j];
      if (wdt < 0 || n_states[j] < 1 || (wdt < 31 && n_states[j] > (1 << wdt)))
        fail(EDHIP_ERR_ARGUMENT, "site states do not fit the site's bits");
      first_state[j + 1] = first_state[j] + n_states[j];
    }
    // Sums of quantum-number tuples as vectors of int64.
    using QN = std::vector<int64_t>;
    auto qn_of = [&](int j, int s) {
      QN q(qn_dim);
      for (int d = 0; d < qn_dim; ++d) q[d] = state_qns[(size_t)(first_state[j] + s) * qn_dim + d];
      return q;
    };
    auto add = [](QN a, const QN& b) {
      for (size_t d = 0; d < a.size(); ++d) a[d] += b[d];
      return a;
    };
    std::vector<QN> allow;
    for (int a = 0; a < n_allowed; ++a) {
      QN q(qn_dim);
      for (int d = 0; d < qn_dim; ++d) q[d] = allowed_qns[(size_t)a * qn_dim + d];
      if (std::find(allow.begin(), allow.end(), q) == allow.end()) allow.push_back(q);
    }
    // fw[i][q]: number of configurations of sites 0..i-1 with total q
    // (hilbert_space_representation.jl:137-151 builds the same reachable sets).
    std::vector<std::map<QN, uint64_t>> fw(n_sites + 1);
    fw[0][QN(qn_dim, 0)] = 1;
    for (int j = 0; j < n_sites; ++j)
      for (const auto& [q, c] : fw[j])
        for (int s = 0; s < n_states[j]; ++s) fw[j + 1][add(q, qn_of(j, s))] += c;
    // completions(i, u): words of sites 0..i-1 that bring upper sum u to an allowed total
    auto completions = [&](int i, const QN& u) {
      uint64_t c = 0;
      for (const QN& Q : allow) {
        QN need(qn_dim);
        for (int d = 0; d < qn_dim; ++d) need[d] = Q[d] - u[d];
        auto it = fw[i].find(need);
        if (it != fw[i].end()) c += it->second;
      }
      return c;
    };
    const uint64_t dim = completions(n_sites, QN(qn_dim, 0));
    if (dim >= 0xffffffffull) fail(EDHIP_ERR_ARGUMENT, "more than 2^32-1 basis states");
    // Levels from the top site down; upper sums with no completion are dropped.
    std::vector<std::vector<QN>> upper(n_sites + 1);
    std::vector<std::map<QN, int>> upper_idx(n_sites + 1);
    if (dim) upper[n_sites].push_back(QN(qn_dim, 0));
    std::vector<QnSite> sites(n_sites);
    std::vector<uint64_t> cnt;
    std::vector<int32_t> nxt;
    for (int j = n_sites - 1; j >= 0; --j) {
      sites[j] = QnSite{(int32_t)cnt.size(), n_states[j], bitoffsets[j], std::max<int32_t>(1, (int32_t)upper[j + 1].size())};
      for (const QN& u : upper[j + 1])
        for (int s = 0; s < n_states[j]; ++s) {
          const QN v = add(u, qn_of(j, s));
          const uint64_t c = completions(j, v);
          if (!c) {
            cnt.push_back(0);
            nxt.push_back(-1);
            continue;
          }
          auto it = upper_idx[j].find(v);
          if (it == upper_idx[j].end()) {
            it = upper_idx[j].emplace(v, (int)upper[j].size()).first;
            upper[j].push_back(v);
          }
          cnt.push_back(c);
          nxt.push_back(it->second);
        }
      if (upper[j + 1].empty())  // empty sector: keep every level's entry range valid
        for (int s = 0; s < n_states[j]; ++s) {
          cnt.push_back(0);
          nxt.push_back(-1);
        }
      if (cnt.size() > (size_t)INT32_MAX / 2) fail(EDHIP_ERR_ARGUMENT, "quantum-number tables too large");
    }
    sp->ctx = ctx;
    sp->n_bits = n_bits;
    sp->dim = (int64_t)dim;
    sp->keys.alloc((size_t)std::max<uint64_t>(dim, 1) * 8);
    if (dim) {
      DevBuf d_sites, d_cnt, d_nxt;
      upload(d_sites, sites, ctx->stream);
      upload(d_cnt, cnt, ctx->stream);
      upload(d_nxt, nxt, ctx->stream);
      HIP_OR_THROW(launch_enumerate_qn(dim, n_sites, d_sites.as<QnSite>(), (int)cnt.size(), d_cnt.as<uint64_t>(),
                                       d_nxt.as<int32_t>(), sp->keys.as<uint64_t>(), ctx->stream));
      ctx->sync();
    }
    sp->materialized = true;
    sp->build_index();
    ctx->sync();
  });
  if (rc == EDHIP_OK) *out = sp.release();
  return rc;
}

int edhip_space_reduce(edhip_ctx* ctx, const edhip_space* parent, int n_group, const int32_t* bit_perm,
                       const double* chi_re, const double* chi_im, double tol, edhip_space** out) {
  return edhip_space_reduce_ex(ctx, parent, n_group, bit_perm, nullptr, chi_re, chi_im, tol, out);
}

int edhip_space_reduce_ex(edhip_ctx* ctx, const edhip_space* parent, int n_group, const int32_t* bit_perm,
                          const uint64_t* flip_host, const double* chi_re, const double* chi_im, double tol,
                          edhip_space** out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  *out = nullptr;
  auto sp = std::make_unique<edhip_space>();
  int rc = guarded(ctx, [&] {
    check_space(parent);
    if (parent->group) fail(EDHIP_ERR_ARGUMENT, "the parent space is already reduced");
    if (n_group < 1 || !bit_perm || !chi_re || !chi_im) fail(EDHIP_ERR_ARGUMENT, "empty symmetry group");
    if (n_group > 4096) fail(EDHIP_ERR_ARGUMENT, "symmetry group too large");
    const int nb = parent->n_bits;
    auto gh = std::make_unique<GroupHost>();
    gh->G = n_group;
    gh->perm.assign(bit_perm, bit_perm + (size_t)n_group * nb);
    if (flip_host) gh->xr.assign(flip_host, flip_host + n_group);
    gh->chi.resize(n_group);
    for (int g = 0; g < n_group; ++g) gh->chi[g] = {chi_re[g], chi_im[g]};
    build_group(*gh, nb, tol, ctx->stream);
    if (gh->has_xor) check_flip_closed(parent);
    {
      uint32_t need = gh->lay.total + (uint32_t)((nb + 1) * (nb + 1) * 8);
      if (need > 64 * 1024)
        fail(EDHIP_ERR_ARGUMENT, "symmetry tables exceed the 64 KiB LDS budget (" + std::to_string(need) + " bytes)");
    }
    ReduceSource src = parent->source();
    // chunking: ~4M threads
    const uint64_t NW = src.n_words;
    uint64_t n_threads = std::min<uint64_t>(std::max<uint64_t>((NW + 15) / 16, 1), 1ull << 22);
    uint64_t chunk = (NW + n_threads - 1) / n_threads;
    if (chunk == 0) chunk = 1;
    n_threads = (NW + chunk - 1) / chunk;
    sp->ctx = ctx;
    sp->n_bits = nb;
    sp->parent = parent;
    if (n_threads == 0) {
      sp->dim = 0;
    } else {
      DevBuf counts, offsets, tmp;
      counts.alloc(n_threads * 4);
      offsets.alloc(n_threads * 8);
      HIP_OR_THROW(launch_reduce_count(src, chunk, n_threads, gh->dev, gh->lay, nullptr, counts.as<uint32_t>(),
                                       ctx->stream));
      size_t tb = 0;
      HIP_OR_THROW(exclusive_scan_u32_to_u64(counts.as<uint32_t>(), offsets.as<uint64_t>(), n_threads, nullptr, tb,
                                             ctx->stream));
      tmp.alloc(tb);
      HIP_OR_THROW(exclusive_scan_u32_to_u64(counts.as<uint32_t>(), offsets.as<uint64_t>(), n_threads, tmp.p, tb,
                                             ctx->stream));
      uint64_t last_off = 0;
      uint32_t last_cnt = 0;
      HIP_OR_THROW(hipMemcpyAsync(&last_off, offsets.as<uint64_t>() + (n_threads - 1), 8, hipMemcpyDeviceToHost,
                                  ctx->stream));
      HIP_OR_THROW(hipMemcpyAsync(&last_cnt, counts.as<uint32_t>() + (n_threads - 1), 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
      ctx->sync();
      uint64_t D = last_off + last_cnt;
      sp->dim = (int64_t)D;
      sp->keys.alloc(std::max<uint64_t>(D, 1) * 8);
      sp->orbit.alloc(std::max<uint64_t>(D, 1) * 2);
      HIP_OR_THROW(launch_reduce_write(src, chunk, n_threads, gh->dev, gh->lay, offsets.as<uint64_t>(),
                                       sp->keys.as<uint64_t>(), sp->orbit.as<uint16_t>(), ctx->stream));
      ctx->sync();
    }
    sp->materialized = true;
    sp->group = std::move(gh);
    sp->build_index();
    ctx->sync();
  });
  if (rc == EDHIP_OK) *out = sp.release();
  return rc;
}

void edhip_space_destroy(edhip_space* s) {
  if (!s) return;
  if (s->ctx) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
  }
  delete s;
}

int edhip_space_dim(const edhip_space* s, int64_t* dim) {
  if (!s || !dim) return EDHIP_ERR_ARGUMENT;
  *dim = s->dim;
  return EDHIP_OK;
}

int edhip_space_copy_basis(const edhip_space* s, uint64_t* out) {
  if (!s || (!out && s->dim)) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    s->materialize();
    if (s->dim)
      HIP_OR_THROW(hipMemcpyAsync(out, s->keys.p, (size_t)s->dim * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    s->ctx->sync();
  });
}

int edhip_space_device_basis(const edhip_space* s, const uint64_t** out) {
  if (!s || !out) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    s->materialize();
    *out = s->keys.as<uint64_t>();
  });
}

int edhip_space_copy_orbit(const edhip_space* s, uint16_t* out) {
  if (!s || (!out && s->dim)) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    if (!s->group) {
      for (int64_t i = 0; i < s->dim; ++i) out[i] = 1;
      return;
    }
    if (s->dim)
      HIP_OR_THROW(hipMemcpyAsync(out, s->orbit.p, (size_t)s->dim * 2, hipMemcpyDeviceToHost, s->ctx->stream));
    s->ctx->sync();
  });
}

int edhip_space_lookup(const edhip_space* s, int64_t n, const uint64_t* keys, int64_t* idx) {
  if (!s || n < 0 || (n && (!keys || !idx))) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    if (n == 0) return;
    DevBuf dk, di;
    dk.alloc((size_t)n * 8);
    di.alloc((size_t)n * 8);
    HIP_OR_THROW(hipMemcpyAsync(dk.p, keys, (size_t)n * 8, hipMemcpyHostToDevice, s->ctx->stream));
    HIP_OR_THROW(launch_lookup(s->index_dev(), n, dk.as<uint64_t>(), di.as<int64_t>(), s->ctx->stream));
    HIP_OR_THROW(hipMemcpyAsync(idx, di.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    s->ctx->sync();
  });
}

int edhip_space_export_mapping(const edhip_space* s, int64_t* idx, double* are, double* aim) {
  if (!s) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    if (!s->group || !s->parent) fail(EDHIP_ERR_ARGUMENT, "space is not symmetry reduced");
    const edhip_space* p = s->parent;
    const int64_t np = p->dim;
    if (np == 0) return;
    if (!idx || !are || !aim) fail(EDHIP_ERR_ARGUMENT, "null output");
    DevBuf di, dr, dm;
    di.alloc((size_t)np * 8);
    dr.alloc((size_t)np * 8);
    dm.alloc((size_t)np * 8);
    ReduceSource src = p->source();
    HIP_OR_THROW(launch_export_mapping(src, s->group->dev, s->group->lay, s->index_dev(), s->orbit.as<uint16_t>(),
                                       di.as<int64_t>(), dr.as<double>(), dm.as<double>(), s->ctx->stream));
    HIP_OR_THROW(hipMemcpyAsync(idx, di.p, (size_t)np * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_OR_THROW(hipMemcpyAsync(are, dr.p, (size_t)np * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_OR_THROW(hipMemcpyAsync(aim, dm.p, (size_t)np * 8, hipMemcpyDeviceToHost, s->ctx->stream));
    s->ctx->sync();
  });
}

namespace {
void check_vector_map(const edhip_space* s, int prec, int64_t n_large, const void* large, int64_t n_small,
                      const void* small) {
  if (!s->group || !s->parent) fail(EDHIP_ERR_ARGUMENT, "space is not symmetry reduced");
  if (elem_size(prec) == 0) fail(EDHIP_ERR_ARGUMENT, "unknown precision code");
  if (!prec_is_complex(prec) && s->group->complex_phase)
    fail(EDHIP_ERR_ARGUMENT, "complex characters need a complex vector type");
  if (n_large != s->parent->dim)
    fail(EDHIP_ERR_DIMENSION, "Dimension of the input vector should match the larger representation");
  if (n_small != s->dim)
    fail(EDHIP_ERR_DIMENSION, "Dimension of the input vector should match the reduced representation");
  if ((n_large > 0 && !large) || (n_small > 0 && !small)) fail(EDHIP_ERR_ARGUMENT, "null vector");
}
}  // namespace

int edhip_space_vector_reduce(const edhip_space* s, int prec, int64_t n_large, const void* large, int64_t n_small,
                              void* small) {
  if (!s) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    check_vector_map(s, prec, n_large, large, n_small, small);
    if (s->dim == 0) return;
    s->materialize();
    HIP_OR_THROW(launch_vec_reduce(s->keys.as<uint64_t>(), s->orbit.as<uint16_t>(), s->dim, s->group->dev,
                                   s->group->lay, s->parent->index_dev(), prec, large, small, s->ctx->stream));
  });
}

int edhip_space_vector_unreduce(const edhip_space* s, int prec, int64_t n_small, const void* small, int64_t n_large,
                                void* large) {
  if (!s) return EDHIP_ERR_ARGUMENT;
  return guarded(s->ctx, [&] {
    check_vector_map(s, prec, n_large, large, n_small, small);
    if (n_large == 0) return;
    HIP_OR_THROW(launch_vec_unreduce(s->parent->source(), s->group->dev, s->group->lay, s->index_dev(1),
                                     s->orbit.as<uint16_t>(), prec, small, large, s->ctx->stream));
  });
}

int edhip_space_info(const edhip_space* s, int* reduced, int* group_order, int* canon_kind) {
  if (!s) return EDHIP_ERR_ARGUMENT;
  if (reduced) *reduced = s->group ? 1 : 0;
  if (group_order) *group_order = s->group ? s->group->G : 1;
  if (canon_kind) *canon_kind = s->group ? s->group->mode : 0;
  return EDHIP_OK;
}

// ---- operators --------------------------------------------------------------
int edhip_op_create(edhip_ctx* ctx, const edhip_space* space, int64_t n_terms, const uint64_t* mask,
                    const uint64_t* row, const uint64_t* col, const double* are, const double* aim, edhip_op** out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  *out = nullptr;
  auto op = std::make_unique<edhip_op>();
  int rc = guarded(ctx, [&] {
    check_space(space);
    if (n_terms < 0 || (n_terms > 0 && (!mask || !row || !col || !are || !aim)))
      fail(EDHIP_ERR_ARGUMENT, "bad term arrays");
    op->ctx = ctx;
    op->space = space;
    op->n_terms = n_terms;
    op->mask.assign(mask, mask + n_terms);
    op->row.assign(row, row + n_terms);
    op->col.assign(col, col + n_terms);
    op->are.assign(are, are + n_terms);
    op->aim.assign(aim, aim + n_terms);
    std::vector<OffTerm> offr, offc, allr, allc;
    std::vector<DiagTerm> dg;
    for (int64_t t = 0; t < n_terms; ++t) {
      if ((~mask[t]) & row[t]) fail(EDHIP_ERR_ARGUMENT, "every bit of bitrow not in bitmask should be set to zero");
      if ((~mask[t]) & col[t]) fail(EDHIP_ERR_ARGUMENT, "every bit of bitcol not in bitmask should be set to zero");
      if (aim[t] != 0.0) op->complex_amp = true;
      allr.push_back(OffTerm{mask[t], row[t], col[t], are[t], aim[t]});
      allc.push_back(OffTerm{mask[t], col[t], row[t], are[t], aim[t]});
      if (row[t] == col[t]) {
        dg.push_back(DiagTerm{mask[t], row[t], are[t], aim[t]});
        op->diag_ids.push_back(t);
      } else {
        if (__builtin_popcountll(row[t]) != __builtin_popcountll(col[t])) op->keeps_popcount = false;
        offr.push_back(OffTerm{mask[t], row[t], col[t], are[t], aim[t]});
        offc.push_back(OffTerm{mask[t], col[t], row[t], are[t], aim[t]});
        op->off_ids.push_back(t);
      }
    }
    op->n_off = (int)offr.size();
    op->n_diag = (int)dg.size();
    upload(op->off_row, offr, ctx->stream);
    upload(op->off_col, offc, ctx->stream);
    // runs of equal masks padded to even lengths (both forms share the masks,
    // so they pad alike); the padding term keeps the mask, matches nothing
    // (a match outside the mask) and flips nothing
    auto pad_runs = [](const std::vector<OffTerm>& v) {
      std::vector<OffTerm> out;
      for (size_t i = 0; i < v.size();) {
        size_t j = i;
        while (j < v.size() && v[j].mask == v[i].mask) ++j;
        out.insert(out.end(), v.begin() + i, v.begin() + j);
        if ((j - i) & 1) out.push_back(OffTerm{v[i].mask, ~v[i].mask, ~v[i].mask, 0.0, 0.0});
        i = j;
      }
      return out;
    };
    const std::vector<OffTerm> mvr = pad_runs(offr), mvc = pad_runs(offc);
    op->n_mv = (int)mvr.size();
    upload(op->mv_row, mvr, ctx->stream);
    upload(op->mv_col, mvc, ctx->stream);
    upload(op->diag, dg, ctx->stream);
    upload(op->all_row, allr, ctx->stream);
    upload(op->all_col, allc, ctx->stream);
    op->dev_row = make_opdev(op.get(), false);
    op->dev_col = make_opdev(op.get(), true);
    op->pairs_ok = build_pair_groups(op.get(), 0, ctx->stream) && build_pair_groups(op.get(), 1, ctx->stream);
    ctx->sync();
  });
  if (rc == EDHIP_OK) *out = op.release();
  return rc;
}

void edhip_op_destroy(edhip_op* op) {
  if (!op) return;
  (void)hipSetDevice(op->ctx->device);
  (void)hipStreamSynchronize(op->ctx->stream);
  delete op;
}

int edhip_op_is_complex(const edhip_op* op, int* c) {
  if (!op || !c) return EDHIP_ERR_ARGUMENT;
  *c = op->is_complex() ? 1 : 0;
  return EDHIP_OK;
}

int edhip_apply(edhip_op* op, int prec, int transpose, int accumulate, int64_t n_x, const void* x, int64_t n_y,
                void* y) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    check_prec_for_op(op, prec);
    const int64_t d = op->space->dim;
    if (n_y != d) fail(EDHIP_ERR_DIMENSION, "out has length " + std::to_string(n_y) + " != dimension " + std::to_string(d));
    if (n_x != d)
      fail(EDHIP_ERR_DIMENSION, "state has length " + std::to_string(n_x) + " != dimension " + std::to_string(d));
    op_matvec(op, prec, transpose ? 1 : 0, accumulate ? 1 : 0, x, y);
  });
}

int edhip_apply_rows(edhip_op* op, int prec, int transpose, int64_t row_begin, int64_t row_end, int64_t n_x,
                     const void* x, int64_t n_y, void* y) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    check_prec_for_op(op, prec);
    const int64_t d = op->space->dim;
    if (row_begin < 0 || row_end < row_begin || row_end > d)
      fail(EDHIP_ERR_ARGUMENT, "rows [" + std::to_string(row_begin) + ", " + std::to_string(row_end) +
                                   ") outside [0, " + std::to_string(d) + ")");
    if (n_x != d)
      fail(EDHIP_ERR_DIMENSION, "state has length " + std::to_string(n_x) + " != dimension " + std::to_string(d));
    if (n_y != row_end - row_begin)
      fail(EDHIP_ERR_DIMENSION, "out has length " + std::to_string(n_y) + " != " + std::to_string(row_end - row_begin) +
                                    " rows");
    term_list_matvec(op, prec, transpose ? 1 : 0, 0, x, y, row_begin, row_end);
  });
}

int edhip_apply_host(edhip_op* op, int prec, int transpose, int accumulate, int64_t n_x, const void* x, int64_t n_y,
                     void* y) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    check_prec_for_op(op, prec);
    const int64_t d = op->space->dim;
    if (n_y != d) fail(EDHIP_ERR_DIMENSION, "out has length " + std::to_string(n_y) + " != dimension " + std::to_string(d));
    if (n_x != d)
      fail(EDHIP_ERR_DIMENSION, "state has length " + std::to_string(n_x) + " != dimension " + std::to_string(d));
    const size_t bytes = (size_t)d * elem_size(prec);
    op->hx.ensure(bytes);
    op->hy.ensure(bytes);
    hipStream_t st = op->ctx->stream;
    if (d) {
      HIP_OR_THROW(hipMemcpyAsync(op->hx.p, x, bytes, hipMemcpyHostToDevice, st));
      if (accumulate) HIP_OR_THROW(hipMemcpyAsync(op->hy.p, y, bytes, hipMemcpyHostToDevice, st));
      op_matvec(op, prec, transpose ? 1 : 0, accumulate ? 1 : 0, op->hx.p, op->hy.p);
      HIP_OR_THROW(hipMemcpyAsync(y, op->hy.p, bytes, hipMemcpyDeviceToHost, st));
    }
    op->ctx->sync();
  });
}

int edhip_op_set_strategy(edhip_op* op, int kernel, int lookup) {
  if (!op || kernel < 0 || kernel > 3 || lookup < 0 || lookup > 4) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    if (kernel == 2 && !fast_path_usable(op))
      fail(EDHIP_ERR_ARGUMENT, "the pair-group kernel does not cover this operator or symmetry group");
    if (lookup == 3) (void)op->space->index_dev(3);  // raises when no bitmap index exists
    if (kernel == 3) {
      ensure_assembled(op);
    } else {
      op->release_assembled();
    }
    op->strat_kernel = kernel;
    op->strat_lookup = lookup;
  });
}

int edhip_op_assembled_info(const edhip_op* op, uint64_t* entries, uint64_t* bytes, double* build_seconds) {
  if (!op || !entries || !bytes || !build_seconds) return EDHIP_ERR_ARGUMENT;
  *entries = op->asm_built ? op->asm_nnz : 0;
  *bytes = op->asm_built ? (uint64_t)(op->asm_ent.bytes + op->asm_rowptr.bytes + op->asm_w.bytes +
                                      op->asm_acls.bytes + op->asm_pcls.bytes)
                         : 0;
  *build_seconds = op->asm_built ? op->asm_build_s : 0.0;
  return EDHIP_OK;
}

int edhip_op_last_strategy(const edhip_op* op, int* kernel, int* lookup) {
  if (!op || !kernel || !lookup) return EDHIP_ERR_ARGUMENT;
  *kernel = op->used_kernel;
  *lookup = op->used_lookup;
  return EDHIP_OK;
}

int edhip_op_count_images(edhip_op* op, int64_t* n_images) {
  if (!op || !n_images) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    DevBuf t;
    t.alloc(8);
    HIP_OR_THROW(hipMemsetAsync(t.p, 0, 8, op->ctx->stream));
    HIP_OR_THROW(launch_count_images(op->dev_row, t.as<unsigned long long>(), op->ctx->stream));
    unsigned long long h = 0;
    HIP_OR_THROW(hipMemcpyAsync(&h, t.p, 8, hipMemcpyDeviceToHost, op->ctx->stream));
    op->ctx->sync();
    *n_images = (int64_t)h;
  });
}

int edhip_iterate(edhip_op* op, int transpose, int64_t n_rows, const int64_t* rows, int64_t* idx, double* re,
                  double* im) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    if (n_rows < 0 || (n_rows && (!rows || !idx || !re || !im))) fail(EDHIP_ERR_ARGUMENT, "bad arguments");
    for (int64_t k = 0; k < n_rows; ++k)
      if (rows[k] < 0 || rows[k] >= op->space->dim) fail(EDHIP_ERR_BOUNDS, "row index out of bounds");
    if (n_rows == 0 || op->n_terms == 0) {
      for (int64_t k = 0; k < n_rows * op->n_terms; ++k) idx[k] = -2;
      return;
    }
    hipStream_t st = op->ctx->stream;
    const size_t nt = (size_t)n_rows * op->n_terms;
    op->iter_rows.ensure((size_t)n_rows * 8);
    op->iter_idx.ensure(nt * 8);
    op->iter_re.ensure(nt * 8);
    op->iter_im.ensure(nt * 8);
    HIP_OR_THROW(hipMemcpyAsync(op->iter_rows.p, rows, (size_t)n_rows * 8, hipMemcpyHostToDevice, st));
    HIP_OR_THROW(launch_iterate(transpose ? op->dev_col : op->dev_row,
                                transpose ? op->all_col.as<OffTerm>() : op->all_row.as<OffTerm>(), (int)op->n_terms,
                                transpose, n_rows, op->iter_rows.as<int64_t>(), op->iter_idx.as<int64_t>(),
                                op->iter_re.as<double>(), op->iter_im.as<double>(), st));
    HIP_OR_THROW(hipMemcpyAsync(idx, op->iter_idx.p, nt * 8, hipMemcpyDeviceToHost, st));
    HIP_OR_THROW(hipMemcpyAsync(re, op->iter_re.p, nt * 8, hipMemcpyDeviceToHost, st));
    HIP_OR_THROW(hipMemcpyAsync(im, op->iter_im.p, nt * 8, hipMemcpyDeviceToHost, st));
    op->ctx->sync();
  });
}

int edhip_sparse_build(edhip_op* op, double tol, int64_t* nnz) {
  if (!op || !nnz) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    hipStream_t st = op->ctx->stream;
    const int64_t d = op->space->dim;
    const int slot = 1 + op->n_off;
    DevBuf sidx, sre, sim, cnt, tmp;
    sidx.alloc((size_t)d * slot * 8);
    sre.alloc((size_t)d * slot * 8);
    sim.alloc((size_t)d * slot * 8);
    cnt.alloc((size_t)(d + 1) * 8);
    HIP_OR_THROW(hipMemsetAsync(cnt.p, 0, (size_t)(d + 1) * 8, st));
    HIP_OR_THROW(launch_sparse_columns(op->dev_col, 1, tol, slot, d, nullptr, sidx.as<int64_t>(), sre.as<double>(),
                                       sim.as<double>(), cnt.as<int64_t>(), st));
    op->sp_colptr.alloc((size_t)(d + 1) * 8);
    size_t tb = 0;
    HIP_OR_THROW(exclusive_scan_i64(cnt.as<int64_t>(), op->sp_colptr.as<int64_t>(), (uint64_t)(d + 1), nullptr, tb, st));
    tmp.alloc(tb);
    HIP_OR_THROW(exclusive_scan_i64(cnt.as<int64_t>(), op->sp_colptr.as<int64_t>(), (uint64_t)(d + 1), tmp.p, tb, st));
    int64_t total = 0;
    HIP_OR_THROW(hipMemcpyAsync(&total, op->sp_colptr.as<int64_t>() + d, 8, hipMemcpyDeviceToHost, st));
    op->ctx->sync();
    op->sp_rowval.alloc((size_t)std::max<int64_t>(total, 1) * 8);
    op->sp_re.alloc((size_t)std::max<int64_t>(total, 1) * 8);
    op->sp_im.alloc((size_t)std::max<int64_t>(total, 1) * 8);
    HIP_OR_THROW(launch_sparse_compact(d, slot, op->sp_colptr.as<int64_t>(), sidx.as<int64_t>(), sre.as<double>(),
                                       sim.as<double>(), op->sp_rowval.as<int64_t>(), op->sp_re.as<double>(),
                                       op->sp_im.as<double>(), st));
    op->ctx->sync();
    op->sp_nnz = total;
    *nnz = total;
  });
}

int edhip_get_vectors(edhip_op* op, int transpose, int64_t n, const int64_t* which, double tol, int64_t* counts,
                      int64_t* idx, double* re, double* im) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    if (n < 0 || (n && (!which || !counts || !idx || !re || !im))) fail(EDHIP_ERR_ARGUMENT, "bad arguments");
    const int64_t d = op->space->dim;
    for (int64_t k = 0; k < n; ++k)
      if (which[k] < 0 || which[k] >= d) fail(EDHIP_ERR_BOUNDS, "index out of bounds");
    if (n == 0) return;
    hipStream_t st = op->ctx->stream;
    const int slot = (int)(1 + op->n_terms);  // the documented stride; 1 + n_off entries are used at most
    const size_t ne = (size_t)n * slot;
    DevBuf sel, sidx, sre, sim, cnt;
    sel.alloc((size_t)n * 8);
    sidx.alloc(ne * 8);
    sre.alloc(ne * 8);
    sim.alloc(ne * 8);
    cnt.alloc((size_t)n * 8);
    HIP_OR_THROW(hipMemcpyAsync(sel.p, which, (size_t)n * 8, hipMemcpyHostToDevice, st));
    HIP_OR_THROW(launch_sparse_columns(transpose ? op->dev_col : op->dev_row, transpose ? 1 : 0, tol, slot, n,
                                       sel.as<int64_t>(), sidx.as<int64_t>(), sre.as<double>(), sim.as<double>(),
                                       cnt.as<int64_t>(), st));
    HIP_OR_THROW(hipMemcpyAsync(counts, cnt.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIP_OR_THROW(hipMemcpyAsync(idx, sidx.p, ne * 8, hipMemcpyDeviceToHost, st));
    HIP_OR_THROW(hipMemcpyAsync(re, sre.p, ne * 8, hipMemcpyDeviceToHost, st));
    HIP_OR_THROW(hipMemcpyAsync(im, sim.p, ne * 8, hipMemcpyDeviceToHost, st));
    op->ctx->sync();
  });
}

int edhip_matrix(edhip_op* op, void* out) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    const int64_t d = op->space->dim;
    if (d && !out) fail(EDHIP_ERR_ARGUMENT, "null output");
    if (d > (int64_t)1 << 20) fail(EDHIP_ERR_ARGUMENT, "dimension too large for a dense matrix");
    if (d == 0) return;
    hipStream_t st = op->ctx->stream;
    const int cx = op->is_complex() ? 1 : 0;
    const size_t bytes = (size_t)d * (size_t)d * (cx ? 16 : 8);
    HIP_OR_THROW(hipMemsetAsync(out, 0, bytes, st));
    HIP_OR_THROW(launch_dense_columns(op->dev_col, cx, (double*)out, st));
    op->ctx->sync();
  });
}

int edhip_sparse_export(edhip_op* op, int64_t* colptr, int64_t* rowval, double* re, double* im) {
  if (!op) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    if (op->sp_nnz < 0) fail(EDHIP_ERR_ARGUMENT, "call edhip_sparse_build first");
    hipStream_t st = op->ctx->stream;
    const int64_t d = op->space->dim;
    HIP_OR_THROW(hipMemcpyAsync(colptr, op->sp_colptr.p, (size_t)(d + 1) * 8, hipMemcpyDeviceToHost, st));
    if (op->sp_nnz) {
      HIP_OR_THROW(hipMemcpyAsync(rowval, op->sp_rowval.p, (size_t)op->sp_nnz * 8, hipMemcpyDeviceToHost, st));
      HIP_OR_THROW(hipMemcpyAsync(re, op->sp_re.p, (size_t)op->sp_nnz * 8, hipMemcpyDeviceToHost, st));
      HIP_OR_THROW(hipMemcpyAsync(im, op->sp_im.p, (size_t)op->sp_nnz * 8, hipMemcpyDeviceToHost, st));
    }
    op->ctx->sync();
  });
}

// ---- eigensolver -----------------------------------------------------------------
void edhip_lanczos_default_opts(edhip_lanczos_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->nev = 1;
  o->max_basis = 120;
  o->max_restarts = 50;
  o->mode = EDHIP_LANCZOS_FP64;
  o->want_vectors = 0;
  o->refine_basis = 12;
  o->max_refine = 20;
  o->check_every = 5;
  o->criterion = 0;
  o->seed = 20241202ull;
  o->tol = 1e-10;
}

int edhip_lanczos(edhip_op* op, const edhip_lanczos_opts* opts, double* evals, void* evecs, edhip_lanczos_stats* stats) {
  if (!op || !opts || !evals) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    Timer talloc;
    edhip_ctx* ctx = op->ctx;
    const int64_t n = op->space->dim;
    const bool cx = op->is_complex();
    const int nev = opts->nev;
    if (nev < 1 || nev > 16) fail(EDHIP_ERR_ARGUMENT, "nev must be in 1..16");
    if (n < nev) fail(EDHIP_ERR_ARGUMENT, "nev exceeds the dimension");
    if (opts->max_basis < nev + 2 && opts->max_basis < n) fail(EDHIP_ERR_ARGUMENT, "max_basis too small for nev");
    const int prec64 = cx ? EDHIP_C128 : EDHIP_F64;
    const int prec_store = (opts->mode == EDHIP_LANCZOS_MIXED) ? (cx ? EDHIP_C64 : EDHIP_F32) : prec64;
    edhip_lanczos_stats st{};
    std::vector<double> theta, res;

    if (opts->reorth < 0 || opts->reorth > 3) fail(EDHIP_ERR_ARGUMENT, "reorth must be 0, 1, 2 or 3");
    if (opts->assemble < 0 || opts->assemble > 1) fail(EDHIP_ERR_ARGUMENT, "assemble must be 0 or 1");
    TRLanczos L;
    L.op = op;
    L.reorth = opts->reorth == 1 ? 1 : 0;
    // auto: partial reorthogonalisation on bases over 2^22 elements
    L.partial = opts->reorth == 2 || (opts->reorth == 0 && n > (int64_t)1 << 22);
    L.seed = opts->seed;
    L.prec = prec_store;
    L.set_size(n);
    L.m = (int)std::min<int64_t>(opts->max_basis, n);
    const bool mixed = opts->mode == EDHIP_LANCZOS_MIXED;
    // One workspace: the FP64 Ritz vectors, then the main basis and its restart
    // scratch, later reused by the refinement cycles' basis and scratch.
    const size_t es_s = elem_size(prec_store), es64 = elem_size(prec64);
    const int m_ref = (int)std::min<int64_t>(std::max(opts->refine_basis, 2), n);
    const size_t ritz_b = align256((size_t)nev * n * es64);
    const size_t main_v = align256((size_t)(L.m + 1) * L.ld * es_s);
    // restart scratch: the kept vectors for a chunk of at most 2^24 rows
    const int64_t rows_t = std::min<int64_t>(L.ld, (int64_t)1 << 24);
    const size_t main_t = align256((size_t)keep_count(L.m, nev + 1) * rows_t * es_s);
    const size_t ref_v = mixed ? align256((size_t)(m_ref + 1) * L.ld * es64) : 0;
    const size_t ref_t = mixed ? align256((size_t)keep_count(m_ref, 2) * rows_t * es64) : 0;
    char* arena = krylov_arena_for(op, ritz_b + std::max(main_v + main_t, ref_v + ref_t));
    ctx->sync();
    st.t_alloc = talloc.sec();
    // the time to E0 starts once the workspace is there (a first call
    // allocates it, later calls of the same size reuse it)
    Timer total;
    // assembled rows for the solve (kernel auto, opts->assemble 0), released
    // when it ends (also on errors)
    SolveAssembly solve_asm;
    if (opts->assemble == 0) st.assembled = solve_asm.begin(op) ? 1 : 0;
    void* ritz = arena;
    L.Vp = arena + ritz_b;
    L.tmpp = arena + ritz_b + main_v;
    L.tmp_bytes = main_t;
    HIP_OR_THROW(vec_gaussian(prec_store, n, opts->seed, L.vec(0), ctx->stream));
    double nrm = std::sqrt(L.dot(L.vec(0), L.vec(0)).real());
    L.scale(L.vec(0), 1.0 / nrm);

    // The FP32 phase stops on the eigenvalue criterion (its residual cannot go
    // below single-precision rounding) at max(tol, tol_fp32), tol_fp32 = 1e-10
    // by default: converged that far, the FP32 Ritz vector needs one FP64
    // product (its Rayleigh quotient and residual) to settle the eigenvalue
    // (6x6 J1-J2: 65 FP32 + 1 FP64 products, against 57 + 10 when the FP32
    // phase stops at 1e-8).  The FP64 phase uses the caller's rule.
    L.criterion = mixed ? 1 : opts->criterion;
    if (opts->tol_fp32 < 0.0) fail(EDHIP_ERR_ARGUMENT, "tol_fp32 must be >= 0");
    const double tol_main = mixed ? std::max(opts->tol, opts->tol_fp32 > 0.0 ? opts->tol_fp32 : 1e-10) : opts->tol;
    // Large bases check convergence at every step: the host eigensolve of the
    // small projected matrix costs far less than the product it may save.
    const bool large = n > (int64_t)1 << 22;
    const int check_main = large ? 1 : opts->check_every;
    bool conv = L.run(nev, tol_main, opts->max_restarts, check_main, theta, res, ritz, prec64, nev + 1);
    st.iterations = L.iterations;
    st.restarts = L.restarts;
    st.matvecs = L.matvecs;
    st.second_passes = L.second_passes;
    st.full_reorth_steps = L.full_steps;
    st.t_matvec = L.t_matvec;
    st.t_ortho = L.t_ortho;
    st.ortho_bytes = L.ortho_bytes;
    for (int i = 0; i < nev && i < 16; ++i) st.unrefined[i] = theta[i];

    if (mixed) {
      Timer tr;
      conv = true;
      // FP64 residual correction with locking: eigenpair i restarts an FP64
      // Lanczos from its mixed-precision Ritz vector y, orthogonal to the
      // already refined ones.  Its first step is the FP64 residual: theta =
      // y^T H y (Rayleigh quotient) and beta = ||H y - theta y||, which
      // bounds the eigenvalue error by beta^2 / gap; when that bound (or
      // the residual rule) holds, one FP64 product settles the eigenvalue.
      // Otherwise the FP64 basis grows from [y, r] (Rayleigh-Ritz on it at
      // every step) and thick-restarts until it holds.
      const std::vector<double> theta32 = theta;
      for (int i = 0; i < nev; ++i) {
        TRLanczos R;
        R.op = op;
        R.reorth = opts->reorth == 1 ? 1 : 0;
        R.seed = opts->seed + 0x632BE59BD9B4E019ull * (uint64_t)(i + 1);
        R.prec = prec64;
        R.set_size(n);
        R.m = (int)std::min<int64_t>(std::max(opts->refine_basis, 2), n - i);
        R.criterion = opts->criterion;
        R.stationary = false;
        R.accept_invariant = true;
        R.gap_hint = (theta32.size() > (size_t)i + 1) ? (theta32[i + 1] - theta32[i]) : 0.0;
        R.Vp = arena + ritz_b;
        R.tmpp = arena + ritz_b + ref_v;
        R.tmp_bytes = ref_t;
        R.locked = ritz;
        R.n_locked = i;
        void* start = (char*)ritz + (size_t)i * n * elem_size(prec64);
        HIP_OR_THROW(hipMemcpyAsync(R.vec(0), start, (size_t)n * elem_size(prec64), hipMemcpyDeviceToDevice,
                                    ctx->stream));
        if (i) R.orthogonalize(ritz, n, i, R.vec(0));
        double nr = std::sqrt(R.dot(R.vec(0), R.vec(0)).real());
        R.scale(R.vec(0), 1.0 / nr);
        std::vector<double> th1, rs1;
        bool ok = R.run(1, opts->tol, std::max(opts->max_refine - 1, 0), 1, th1, rs1, start, prec64, 2);
        theta[i] = th1[0];
        res[i] = rs1.empty() ? 0.0 : rs1[0];
        st.matvecs_fp64 += R.matvecs;
        st.second_passes += R.second_passes;
        st.full_reorth_steps += R.full_steps;
        st.t_ortho += R.t_ortho;
        st.ortho_bytes += R.ortho_bytes;
        conv = conv && ok;
      }
      st.t_refine = tr.sec();
    }
    for (int i = 0; i < nev; ++i) {
      evals[i] = theta[i];
      if (i < 16) st.residual[i] = res.size() > (size_t)i ? res[i] / std::max(1.0, std::fabs(theta[i])) : 0.0;
    }
    if (evecs && opts->want_vectors)
      HIP_OR_THROW(hipMemcpyAsync(evecs, ritz, (size_t)nev * n * elem_size(prec64), hipMemcpyDeviceToDevice,
                                  ctx->stream));
    ctx->sync();
    st.converged = conv ? 1 : 0;
    st.t_total = total.sec();
    if (stats) *stats = st;
    if (!conv) fail(EDHIP_ERR_NOCONV, "Lanczos did not converge within max_restarts");
  });
}

// ---- differentiation ------------------------------------------------------------------
int edhip_term_expect(edhip_op* op, int prec, const void* phi, const void* psi, double* out) {
  if (!op || !out) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    check_prec_for_op(op, prec);
    const int nt = op->n_diag + op->n_off;
    if (nt == 0) return;
    int64_t nb = term_expect_blocks(op->dev_row);
    op->ex_partial.ensure((size_t)term_expect_scratch_rows(op->dev_row) * nt * 2 * 8);
    op->ex_out.ensure((size_t)nt * 2 * 8);
    const int nc = prec_is_complex(prec) ? 2 : 1;
    if (term_expect_window(op->dev_row, nc) < 1)
      fail(EDHIP_ERR_ARGUMENT, "operator and symmetry tables leave no LDS room for the fused expectation kernel");
    // psi of the images from its spread, as the products read x, where its
    // code table leaves LDS room for the bins
    OpDev d = op->dev_row;
    if (attach_spread(op, prec, d, psi, op->ctx->stream) && term_expect_window(d, nc) < 1) d.xm = XMapDev{};
    edhip_ctx* ctx = op->ctx;
    if (ctx->kernel_timing) HIP_OR_THROW(hipEventRecord(ctx->ev[14], ctx->stream));
    HIP_OR_THROW(launch_term_expect(d, prec, phi, psi, op->ex_partial.as<double>(), nb, op->ex_out.as<double>(),
                                    ctx->stream));
    if (ctx->kernel_timing) {
      HIP_OR_THROW(hipEventRecord(ctx->ev[15], ctx->stream));
      HIP_OR_THROW(hipEventSynchronize(ctx->ev[15]));
      float f = 0.f;
      HIP_OR_THROW(hipEventElapsedTime(&f, ctx->ev[14], ctx->ev[15]));
      op->kt_ms += f;
      ++op->kt_n;
    }
    std::vector<double> h((size_t)nt * 2);
    HIP_OR_THROW(hipMemcpyAsync(h.data(), op->ex_out.p, h.size() * 8, hipMemcpyDeviceToHost, op->ctx->stream));
    op->ctx->sync();
    for (int64_t t = 0; t < op->n_terms; ++t) out[2 * t] = out[2 * t + 1] = 0.0;
    for (int i = 0; i < op->n_diag; ++i) {
      out[2 * op->diag_ids[i]] = h[2 * (size_t)i];
      out[2 * op->diag_ids[i] + 1] = h[2 * (size_t)i + 1];
    }
    for (int i = 0; i < op->n_off; ++i) {
      out[2 * op->off_ids[i]] = h[2 * (size_t)(op->n_diag + i)];
      out[2 * op->off_ids[i] + 1] = h[2 * (size_t)(op->n_diag + i) + 1];
    }
  });
}

int edhip_term_expect_noalloc(edhip_op* op, int prec, const void* phi, const void* psi, double* out) {
  if (!op || !out) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    check_prec_for_op(op, prec);
    const int nt = op->n_diag + op->n_off;
    if (nt > kNoAllocMaxTerms)
      fail(EDHIP_ERR_ARGUMENT, "edhip_term_expect_noalloc: " + std::to_string(nt) + " terms exceed its maximum of " +
                                   std::to_string(kNoAllocMaxTerms));
    for (int64_t t = 0; t < op->n_terms; ++t) out[2 * t] = out[2 * t + 1] = 0.0;
    if (nt == 0) return;
    double* scr = term_expect_noalloc_scratch(op->ctx->device);
    if (!scr) fail(EDHIP_ERR_DEVICE, "the scratch of this device was not set up");
    // the operator's lookup as built at its creation (no spread of psi: that
    // would allocate), partial rows of this operator's width in the scratch
    const OpDev& d = op->dev_row;
    const int nc = prec_is_complex(prec) ? 2 : 1;
    if (term_expect_window(d, nc) < 1)
      fail(EDHIP_ERR_ARGUMENT, "operator and symmetry tables leave no LDS room for the fused expectation kernel");
    const int64_t nb = term_expect_blocks(d);
    const size_t width = (size_t)nt * 2;
    double* res = scr + (size_t)term_expect_scratch_rows(d) * width;
    if ((size_t)(res - scr) + width > term_expect_noalloc_doubles())
      fail(EDHIP_ERR_ARGUMENT, "edhip_term_expect_noalloc: scratch too small for this operator");
    HIP_OR_THROW(launch_term_expect(d, prec, phi, psi, scr, nb, res, op->ctx->stream));
    double h[2 * kNoAllocMaxTerms];
    HIP_OR_THROW(hipMemcpyAsync(h, res, width * 8, hipMemcpyDeviceToHost, op->ctx->stream));
    op->ctx->sync();
    for (int i = 0; i < op->n_diag; ++i) {
      out[2 * op->diag_ids[i]] = h[2 * (size_t)i];
      out[2 * op->diag_ids[i] + 1] = h[2 * (size_t)i + 1];
    }
    for (int i = 0; i < op->n_off; ++i) {
      out[2 * op->off_ids[i]] = h[2 * (size_t)(op->n_diag + i)];
      out[2 * op->off_ids[i] + 1] = h[2 * (size_t)(op->n_diag + i) + 1];
    }
  });
}

int edhip_minres(edhip_op* op, int prec, double shift, int n_deflate, const void* deflate, const void* b, void* x,
                 double tol, int max_iter, edhip_minres_stats* stats) {
  if (!op || !b || !x) return EDHIP_ERR_ARGUMENT;
  return guarded(op->ctx, [&] {
    if (prec != EDHIP_F64 && prec != EDHIP_C128) fail(EDHIP_ERR_ARGUMENT, "MINRES runs in FP64 or ComplexF64");
    check_prec_for_op(op, prec);
    if (n_deflate < 0 || (n_deflate > 0 && !deflate)) fail(EDHIP_ERR_ARGUMENT, "bad deflation vectors");
    edhip_ctx* ctx = op->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = op->space->dim;
    const size_t bytes = (size_t)n * elem_size(prec);
    TRLanczos H;  // reuse the vector helpers
    H.op = op;
    H.prec = prec;
    H.set_size(n);
    const size_t vb = align256(bytes);
    char* buf = krylov_arena_for(op, vb * 7);
    SolveAssembly solve_asm;  // assembled rows for the solve, as the eigensolver takes them
    const bool assembled = solve_asm.begin(op);
    auto V = [&](int i) { return (void*)(buf + (size_t)i * vb); };
    void *r1 = V(0), *r2 = V(1), *y = V(2), *v = V(3), *w = V(4), *w1 = V(5), *w2 = V(6);
    auto project = [&](void* z) {
      if (n_deflate) H.orthogonalize(deflate, n, n_deflate, z);
    };
    auto apply_A = [&](const void* in, void* out) {
      op_matvec(op, prec, 0, 0, in, out);
      if (shift != 0.0) HIP_OR_THROW(vec_axpby(prec, n, -shift, 0.0, in, 1.0, 0.0, out, s));
      project(out);
    };
    // x = 0; r1 = P b
    HIP_OR_THROW(vec_fill_zero(prec, n, x, s));
    HIP_OR_THROW(hipMemcpyAsync(r1, b, bytes, hipMemcpyDeviceToDevice, s));
    project(r1);
    HIP_OR_THROW(hipMemcpyAsync(y, r1, bytes, hipMemcpyDeviceToDevice, s));
    HIP_OR_THROW(hipMemcpyAsync(r2, r1, bytes, hipMemcpyDeviceToDevice, s));
    HIP_OR_THROW(vec_fill_zero(prec, n, w, s));
    HIP_OR_THROW(vec_fill_zero(prec, n, w2, s));
    double beta1 = std::sqrt(std::max(H.dot(r1, r1).real(), 0.0));
    edhip_minres_stats out{};
    out.assembled = assembled ? 1 : 0;
    if (beta1 == 0.0) {
      out.converged = 1;
      if (stats) *stats = out;
      return;
    }
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
    int itn = 0;
    bool conv = false;
    for (itn = 1; itn <= max_iter; ++itn) {
      // v = y / beta ; y = A v - (beta/oldb) r1
      HIP_OR_THROW(vec_axpby(prec, n, 1.0 / beta, 0.0, y, 0.0, 0.0, v, s));
      apply_A(v, y);
      if (itn >= 2) HIP_OR_THROW(vec_axpby(prec, n, -beta / oldb, 0.0, r1, 1.0, 0.0, y, s));
      double alfa = H.dot(v, y).real();
      HIP_OR_THROW(vec_axpby(prec, n, -alfa / beta, 0.0, r2, 1.0, 0.0, y, s));
      // r1 = r2 ; r2 = y
      std::swap(r1, r2);
      HIP_OR_THROW(hipMemcpyAsync(r2, y, bytes, hipMemcpyDeviceToDevice, s));
      oldb = beta;
      beta = std::sqrt(std::max(H.dot(y, y).real(), 0.0));
      // plane rotation
      double oldeps = epsln;
      double delta = cs * dbar + sn * alfa;
      double gbar = sn * dbar - cs * alfa;
      epsln = sn * beta;
      dbar = -cs * beta;
      double gamma = std::max(std::hypot(gbar, beta), 1e-300);
      cs = gbar / gamma;
      sn = beta / gamma;
      double phi = cs * phibar;
      phibar = sn * phibar;
      // w1 = w2 ; w2 = w ; w = (v - oldeps w1 - delta w2) / gamma ; x += phi w
      std::swap(w1, w2);  // w1 <- old w2
      std::swap(w2, w);   // w2 <- old w, w <- old w1 (scratch)
      HIP_OR_THROW(vec_axpby(prec, n, 1.0 / gamma, 0.0, v, 0.0, 0.0, w, s));
      HIP_OR_THROW(vec_axpby(prec, n, -oldeps / gamma, 0.0, w1, 1.0, 0.0, w, s));
      HIP_OR_THROW(vec_axpby(prec, n, -delta / gamma, 0.0, w2, 1.0, 0.0, w, s));
      HIP_OR_THROW(vec_axpby(prec, n, phi, 0.0, w, 1.0, 0.0, x, s));
      out.residual = phibar / beta1;
      if (phibar <= tol * beta1) {
        conv = true;
        break;
      }
      if (beta == 0.0) {
        conv = true;
        break;
      }
    }
    project(x);
    ctx->sync();
    out.iterations = std::min(itn, max_iter);
    out.converged = conv ? 1 : 0;
    if (stats) *stats = out;
    if (!conv) fail(EDHIP_ERR_NOCONV, "MINRES did not converge");
  });
}

// ---- vector helpers ------------------------------------------------------------
int edhip_vec_dot(edhip_ctx* ctx, int prec, int64_t n, const void* x, const void* y, double* out) {
  if (!ctx || !out) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    if (elem_size(prec) == 0) fail(EDHIP_ERR_ARGUMENT, "unknown precision code");
    double* part = ctx->partial_buf(2 * (size_t)vec_partial_blocks(n));
    double* o = ctx->small_buf(2);
    HIP_OR_THROW(vec_dot(prec, n, x, y, part, o, ctx->stream));
    HIP_OR_THROW(hipMemcpyAsync(out, o, 16, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int edhip_vec_gaussian(edhip_ctx* ctx, int prec, int64_t n, uint64_t seed, void* x) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    if (elem_size(prec) == 0) fail(EDHIP_ERR_ARGUMENT, "unknown precision code");
    HIP_OR_THROW(vec_gaussian(prec, n, seed, x, ctx->stream));
  });
}

int edhip_vec_convert(edhip_ctx* ctx, int pi, int po, int64_t n, const void* x, void* y) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    if (elem_size(pi) == 0 || elem_size(po) == 0) fail(EDHIP_ERR_ARGUMENT, "unknown precision code");
    HIP_OR_THROW(vec_convert(pi, po, n, x, y, ctx->stream));
  });
}

int edhip_event_record(edhip_ctx* ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 16) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] { HIP_OR_THROW(hipEventRecord(ctx->ev[slot], ctx->stream)); });
}

int edhip_debug_fail_allocs(int kinds, int count) {
  if (kinds < 0 || kinds > (kAllocSpread | kAllocWorkspace | kAllocAssembled) || count < 0) return EDHIP_ERR_ARGUMENT;
  g_fail_count.store(0);
  g_fail_kinds.store(kinds);
  g_fail_count.store(count);
  return EDHIP_OK;
}

int edhip_debug_asm_shape(int lanes, int batch) {
  if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32) return EDHIP_ERR_ARGUMENT;
  if (batch != 0 && batch != 4 && batch != 8 && batch != 16) return EDHIP_ERR_ARGUMENT;
  g_asm_lanes.store(lanes);
  g_asm_batch.store(batch);
  return EDHIP_OK;
}

int edhip_debug_block_order(int chunk) {
  if (chunk < 0 || chunk > 64) return EDHIP_ERR_ARGUMENT;
  g_xcd_chunk.store(chunk);
  return EDHIP_OK;
}

int edhip_set_kernel_timing(edhip_ctx* ctx, int on) {
  if (!ctx) return EDHIP_ERR_ARGUMENT;
  ctx->kernel_timing = on != 0;
  return EDHIP_OK;
}

int edhip_op_kernel_time(edhip_op* op, int reset, double* ms_total, int64_t* launches) {
  if (!op || !ms_total || !launches) return EDHIP_ERR_ARGUMENT;
  *ms_total = op->kt_ms;
  *launches = op->kt_n;
  if (reset) {
    op->kt_ms = 0.0;
    op->kt_n = 0;
  }
  return EDHIP_OK;
}

int edhip_event_elapsed(edhip_ctx* ctx, int s0, int s1, double* ms) {
  if (!ctx || !ms || s0 < 0 || s0 >= 16 || s1 < 0 || s1 >= 16) return EDHIP_ERR_ARGUMENT;
  return guarded(ctx, [&] {
    HIP_OR_THROW(hipEventSynchronize(ctx->ev[s1]));
    float f = 0.f;
    HIP_OR_THROW(hipEventElapsedTime(&f, ctx->ev[s0], ctx->ev[s1]));
    *ms = f;
  });
}

}  // extern "C"
