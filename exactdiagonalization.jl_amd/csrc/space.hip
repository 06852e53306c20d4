// Space kernels: sector enumeration, prefix table, symmetry reduction,
// mapping export and batched lookup.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace edhip {

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_enumerate_sector(uint64_t first, uint64_t count, int n_bits,
                                                             const uint64_t* __restrict__ W, uint64_t* __restrict__ out) {
  extern __shared__ unsigned char lds[];
  uint64_t* Wl = (uint64_t*)lds;
  const int nw = (n_bits + 1) * (n_bits + 1);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) Wl[i] = W[i];
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = unrank_word(first + i, n_bits, Wl);
}

// Multi-quantum-number sectors (sites of any number of states, QN tuples):
// words[i] = the i-th word of the sector in ascending order.  The most
// significant site is the last one, so ascending words are the lexicographic
// order of (state of site n-1, ..., state of site 0); the walk fixes the sites
// from the top, skipping the counted completions of every smaller state.
// Table entry sites[j].off + u*sites[j].ns + s: the number of words whose sites
// above j sum to upper-sum u (an index into the sums reachable at that level)
// and whose site j is in state s, and the index of the new upper sum (-1 when
// no word has it; such an entry has count 0 and is never taken).
template <bool STAGE>
__global__ __launch_bounds__(kBlock) void k_enumerate_qn(uint64_t count, int n_sites, const QnSite* __restrict__ sites,
                                                         int n_entries, const uint64_t* __restrict__ cnt_g,
                                                         const int32_t* __restrict__ nxt_g, uint64_t* __restrict__ out) {
  extern __shared__ unsigned char lds[];
  const uint64_t* cnt = cnt_g;
  const int32_t* nxt = nxt_g;
  if (STAGE) {
    uint64_t* c = (uint64_t*)lds;
    int32_t* x = (int32_t*)(c + n_entries);
    for (int i = threadIdx.x; i < n_entries; i += blockDim.x) {
      c[i] = cnt_g[i];
      x[i] = nxt_g[i];
    }
    __syncthreads();
    cnt = c;
    nxt = x;
  }
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t r = i, w = 0;
  int u = 0;
  for (int j = n_sites - 1; j >= 0; --j) {
    const QnSite st = sites[j];
    const int e = st.off + min(u, st.n_upper - 1) * st.ns;
    int pick = st.ns - 1;
    for (int s = 0; s < st.ns - 1; ++s) {
      const uint64_t c = cnt[e + s];
      if (r < c) {
        pick = s;
        break;
      }
      r -= c;
    }
    w |= (uint64_t)pick << st.shift;
    u = max(nxt[e + pick], 0);
  }
  out[i] = w;
}

__global__ __launch_bounds__(kBlock) void k_build_prefix(const uint64_t* __restrict__ keys, int64_t dim, int shift,
                                                         uint64_t n_buckets, uint32_t* __restrict__ prefix) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n_buckets) return;
  uint64_t key = j << shift;
  int64_t lo = 0, hi = dim;
  while (lo < hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  prefix[j] = (uint32_t)lo;
}

// Walk the words of chunk [r0, r1) of the source in ascending order.
template <typename F>
__device__ __forceinline__ void for_chunk(const ReduceSource& src, const uint64_t* Wl, uint64_t r0, uint64_t r1, F&& f) {
  if (r0 >= r1) return;
  if (src.list) {
    for (uint64_t r = r0; r < r1; ++r) f(src.list[r]);
    return;
  }
  uint64_t w = unrank_word(r0, src.n_bits, Wl);
  for (uint64_t r = r0; r < r1; ++r) {
    if (r != r0) w = src.single_popcount ? next_same_popcount(w) : unrank_word(r, src.n_bits, Wl);
    f(w);
  }
}

__device__ __forceinline__ const uint64_t* stage_W(const ReduceSource& src, unsigned char* lds, uint32_t off) {
  uint64_t* Wl = (uint64_t*)(lds + off);
  if (!src.list) {
    const int nw = (src.n_bits + 1) * (src.n_bits + 1);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) Wl[i] = src.W[i];
  }
  return Wl;
}

__global__ __launch_bounds__(kBlock) void k_reduce_count(ReduceSource src, uint64_t chunk, uint64_t n_threads, GroupDev g,
                                                         GroupLdsLayout lay, uint32_t* __restrict__ counts) {
  extern __shared__ unsigned char lds[];
  GroupLds L = stage_group(g, lay, lds);
  const uint64_t* Wl = stage_W(src, lds, lay.total);
  __syncthreads();
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_threads) return;
  uint64_t r0 = t * chunk;
  uint64_t r1 = r0 + chunk < src.n_words ? r0 + chunk : src.n_words;
  uint32_t c = 0;
  for_chunk(src, Wl, r0, r1, [&](uint64_t w) {
    int orbit;
    if (is_representative(g, L, w, orbit)) ++c;
  });
  counts[t] = c;
}

__global__ __launch_bounds__(kBlock) void k_reduce_write(ReduceSource src, uint64_t chunk, uint64_t n_threads, GroupDev g,
                                                         GroupLdsLayout lay, const uint64_t* __restrict__ offsets,
                                                         uint64_t* __restrict__ reps, uint16_t* __restrict__ orbits) {
  extern __shared__ unsigned char lds[];
  GroupLds L = stage_group(g, lay, lds);
  const uint64_t* Wl = stage_W(src, lds, lay.total);
  __syncthreads();
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_threads) return;
  uint64_t r0 = t * chunk;
  uint64_t r1 = r0 + chunk < src.n_words ? r0 + chunk : src.n_words;
  uint64_t pos = offsets[t];
  for_chunk(src, Wl, r0, r1, [&](uint64_t w) {
    int orbit;
    if (is_representative(g, L, w, orbit)) {
      reps[pos] = w;
      orbits[pos] = (uint16_t)orbit;
      ++pos;
    }
  });
}

__global__ __launch_bounds__(kBlock) void k_export_mapping(ReduceSource src, GroupDev g, GroupLdsLayout lay, IndexDev ix,
                                                           const uint16_t* __restrict__ orbit,
                                                           int64_t* __restrict__ idx_out, double* __restrict__ re_out,
                                                           double* __restrict__ im_out) {
  extern __shared__ unsigned char lds[];
  GroupLds L = stage_group(g, lay, lds);
  const uint64_t* Wl = stage_W(src, lds, lay.total);
  __syncthreads();
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= src.n_words) return;
  uint64_t w = src.list ? src.list[p] : unrank_word(p, src.n_bits, Wl);
  uint64_t s;
  int h, cnt;
  canonicalize(g, L, w, s, h, cnt);
  int64_t r = search_sorted(ix.keys, ix.prefix, ix.n_buckets, ix.shift, s);
  if (r < 0) {
    idx_out[p] = -1;
    re_out[p] = 0.0;
    im_out[p] = 0.0;
    return;
  }
  double scale = g.invsqrt[orbit[r]];
  idx_out[p] = r;
  re_out[p] = g.ph_re[h] * scale;
  im_out[p] = g.ph_im[h] * scale;
}

// symmetry_unreduce (symmetry_reduce.jl:208-225): large[p] = U[p] * small[r(p)]
// for every parent word p, 0 where p lies outside the reduced space.  The
// mapping (r, U) is recomputed as in k_export_mapping; one lane per parent
// entry, coalesced writes.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_vec_unreduce(ReduceSource src, GroupDev g, GroupLdsLayout lay, IndexDev ix,
                                                         const uint16_t* __restrict__ orbit, const V* __restrict__ small,
                                                         V* __restrict__ large) {
  extern __shared__ unsigned char lds[];
  GroupLds L = stage_group(g, lay, lds);
  const uint64_t* Wl = stage_W(src, lds, lay.total);
  __syncthreads();
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= src.n_words) return;
  uint64_t w = src.list ? src.list[p] : unrank_word(p, src.n_bits, Wl);
  uint64_t s;
  int h, cnt;
  canonicalize(g, L, w, s, h, cnt);
  int64_t r = search_sorted(ix.keys, ix.prefix, ix.n_buckets, ix.shift, s);
  V out = vzero<V>();
  if (r >= 0) {
    using R = typename VT<V>::R;
    const double scale = g.invsqrt[orbit[r]];
    vfma(out, (R)(L.ph_re[h] * scale), (R)(L.ph_im[h] * scale), small[r]);
  }
  large[p] = out;
}

// symmetry_reduce of a vector (symmetry_reduce.jl:38-56):
//   small[r] = sum_{p in orbit(r)} conj(U[p]) * large[p].
// Gather form, one lane per representative s: every element g gives the
// orbit member g(s) with conj(U[g(s)]) = chi(g) / sqrt(N_s); a member reached
// by |Stab(s)| elements (all with the same character, since s is compatible)
// is counted once by dividing by |Stab(s)| = G / N_s.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_vec_reduce(const uint64_t* __restrict__ reps, const uint16_t* __restrict__ orbit,
                                                       int64_t dim, GroupDev g, GroupLdsLayout lay, IndexDev pix,
                                                       const V* __restrict__ large, V* __restrict__ small) {
  extern __shared__ unsigned char lds[];
  GroupLds L = stage_group(g, lay, lds);
  uint64_t* Wl = (uint64_t*)(lds + lay.total);
  if (pix.rank_mode) {
    const int nw = (pix.n_bits + 1) * (pix.n_bits + 1);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) Wl[i] = pix.W[i];
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= dim) return;
  using R = typename VT<V>::R;
  const uint64_t s = reps[r];
  const int n_orb = orbit[r];
  const double scale = g.invsqrt[n_orb] * (double)n_orb / (double)g.G;
  V acc = vzero<V>();
  auto add = [&](int e, uint64_t img) {
    const int64_t p = index_of(pix, Wl, img);
    if (p < 0) return;
    const int ie = L.inv[e];  // chi(e) = conj(phase[e^-1])
    vfma(acc, (R)(L.ph_re[ie] * scale), (R)(-L.ph_im[ie] * scale), large[p]);
  };
  if (g.mode == kGroupTorus) {
    const int nT = g.Lx * g.Ly;
    for (int p = 0; p < g.n_cosets; ++p) {  // elements T(pos) p
      const uint64_t sp = apply_coset(g, L, p, s);
      for (int bb = 0; bb < g.Ly; ++bb) {
        const uint64_t shifted = rows_shift(sp, bb, g.Lx, g.nb, g.full_mask);
        for (int a = 0; a < g.Lx; ++a)
          add(L.coset_elem[p * nT + a + g.Lx * bb], rows_rotl(shifted, a, g.Lx, L.rot_hi, L.rot_lo));
      }
    }
  } else {
    const int stride = g.n_nib << 4;
    with_nib_tables(g, L, [&](const uint64_t* nib) {
      for (int e = 0; e < g.G; ++e) add(e, perm_apply_nib(nib + e * stride, g.n_nib, s) ^ (g.has_xor ? L.xr[e] : 0ull));
    });
  }
  small[r] = acc;
}

__global__ __launch_bounds__(kBlock) void k_lookup(IndexDev ix, int64_t n, const uint64_t* __restrict__ keys,
                                                   int64_t* __restrict__ out) {
  extern __shared__ unsigned char lds[];
  uint64_t* Wl = (uint64_t*)lds;
  if (ix.rank_mode) {
    const int nw = (ix.n_bits + 1) * (ix.n_bits + 1);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) Wl[i] = ix.W[i];
  }
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = index_of(ix, Wl, keys[i]);
}

__global__ __launch_bounds__(kBlock) void k_bitmap_set(const uint64_t* __restrict__ keys, int64_t dim,
                                                       unsigned long long* __restrict__ blocks) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const uint64_t k = keys[i];
  const uint64_t j = k / kRankBlockWords;
  const uint32_t within = (uint32_t)(k - j * kRankBlockWords);
  atomicOr(blocks + j * 8 + (within >> 6), 1ull << (within & 63u));
}

// per block: total members (for the scan) and the packed counts of u64 0..w-1
__global__ __launch_bounds__(kBlock) void k_bitmap_block_counts(uint64_t* __restrict__ blocks, uint64_t n_blocks,
                                                                uint32_t* __restrict__ totals) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_blocks) return;
  uint64_t* blk = blocks + j * 8;
  uint32_t c = 0;
  uint64_t packed = 0;
  for (int w = 0; w < 6; ++w) {
    if (w > 0) packed |= (uint64_t)c << (9 * (w - 1));
    c += (uint32_t)__builtin_popcountll(blk[w]);
  }
  blk[7] = packed;
  totals[j] = c;
}

__global__ __launch_bounds__(kBlock) void k_bitmap_set_base(uint64_t* __restrict__ blocks, uint64_t n_blocks,
                                                            const uint64_t* __restrict__ base) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_blocks) return;
  blocks[j * 8 + 6] = base[j];
}

// flags[key >> shift] = 1 for the top value of every key (idempotent stores)
__global__ __launch_bounds__(kBlock) void k_mark_tops(const uint64_t* __restrict__ keys, int64_t dim, int shift,
                                                      uint32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  flags[keys[i] >> shift] = 1u;
}

// Before a product: x spread over the word positions (XMapDev).
template <typename V>
__global__ __launch_bounds__(kBlock) void k_xmap_fill(XMapDev X, const uint64_t* __restrict__ keys, int64_t dim,
                                                      const V* __restrict__ x, const uint16_t* __restrict__ orbit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const uint64_t k = keys[i];
  const uint64_t pos = xmap_pos(X, X.colex, X.topcode[k >> X.top_shift], k);
  V v = x[i];
  if (orbit) {  // x / sqrt(N), scaled in double
    using R = typename VT<V>::R;
    const double sc = 1.0 / sqrt((double)orbit[i]);
    if constexpr (VT<V>::cx) {
      v.re = (R)((double)v.re * sc);
      v.im = (R)((double)v.im * sc);
    } else {
      v = (R)((double)v * sc);
    }
  }
  ((V*)X.values)[pos] = v;
}

__global__ void k_u32_to_u64(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

inline uint32_t w_bytes(int n_bits) { return (uint32_t)((n_bits + 1) * (n_bits + 1) * sizeof(uint64_t)); }

}  // namespace

hipError_t launch_enumerate_sector(uint64_t first, uint64_t count, int n_bits, const uint64_t* d_W, uint64_t* d_out,
                                   hipStream_t st) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_enumerate_sector, dim3(grid_for(count)), dim3(kBlock), w_bytes(n_bits), st, first, count, n_bits,
                     d_W, d_out);
  return hipGetLastError();
}

hipError_t launch_enumerate_qn(uint64_t count, int n_sites, const QnSite* d_sites, int n_entries, const uint64_t* d_cnt,
                               const int32_t* d_nxt, uint64_t* d_out, hipStream_t st) {
  if (count == 0) return hipSuccess;
  const size_t lds = (size_t)n_entries * (sizeof(uint64_t) + sizeof(int32_t));
  if (lds <= kQnStageBytes)
    hipLaunchKernelGGL(k_enumerate_qn<true>, dim3(grid_for(count)), dim3(kBlock), (uint32_t)lds, st, count, n_sites,
                       d_sites, n_entries, d_cnt, d_nxt, d_out);
  else
    hipLaunchKernelGGL(k_enumerate_qn<false>, dim3(grid_for(count)), dim3(kBlock), 0, st, count, n_sites, d_sites,
                       n_entries, d_cnt, d_nxt, d_out);
  return hipGetLastError();
}

hipError_t launch_build_prefix(const uint64_t* d_keys, int64_t dim, int shift, uint64_t n_buckets, uint32_t* d_prefix,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_build_prefix, dim3(grid_for(n_buckets + 1)), dim3(kBlock), 0, st, d_keys, dim, shift, n_buckets,
                     d_prefix);
  return hipGetLastError();
}

hipError_t launch_reduce_count(const ReduceSource& src, uint64_t chunk, uint64_t n_threads, const GroupDev& g,
                               const GroupLdsLayout& lay, const uint64_t*, uint32_t* d_counts, hipStream_t st) {
  if (n_threads == 0) return hipSuccess;
  uint32_t lds = lay.total + (src.list ? 0u : w_bytes(src.n_bits));
  hipLaunchKernelGGL(k_reduce_count, dim3(grid_for(n_threads)), dim3(kBlock), lds, st, src, chunk, n_threads, g, lay,
                     d_counts);
  return hipGetLastError();
}

hipError_t launch_reduce_write(const ReduceSource& src, uint64_t chunk, uint64_t n_threads, const GroupDev& g,
                               const GroupLdsLayout& lay, const uint64_t* d_offsets, uint64_t* d_reps,
                               uint16_t* d_orbit, hipStream_t st) {
  if (n_threads == 0) return hipSuccess;
  uint32_t lds = lay.total + (src.list ? 0u : w_bytes(src.n_bits));
  hipLaunchKernelGGL(k_reduce_write, dim3(grid_for(n_threads)), dim3(kBlock), lds, st, src, chunk, n_threads, g, lay,
                     d_offsets, d_reps, d_orbit);
  return hipGetLastError();
}

hipError_t launch_export_mapping(const ReduceSource& src, const GroupDev& g, const GroupLdsLayout& lay,
                                 const IndexDev& reduced_index, const uint16_t* d_orbit, int64_t* d_idx, double* d_re,
                                 double* d_im, hipStream_t st) {
  if (src.n_words == 0) return hipSuccess;
  uint32_t lds = lay.total + (src.list ? 0u : w_bytes(src.n_bits));
  hipLaunchKernelGGL(k_export_mapping, dim3(grid_for(src.n_words)), dim3(kBlock), lds, st, src, g, lay, reduced_index,
                     d_orbit, d_idx, d_re, d_im);
  return hipGetLastError();
}

hipError_t launch_lookup(const IndexDev& ix, int64_t n, const uint64_t* d_keys, int64_t* d_idx, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint32_t lds = ix.rank_mode ? w_bytes(ix.n_bits) : 0u;
  hipLaunchKernelGGL(k_lookup, dim3(grid_for((uint64_t)n)), dim3(kBlock), lds, st, ix, n, d_keys, d_idx);
  return hipGetLastError();
}

hipError_t launch_bitmap_build(const uint64_t* d_keys, int64_t dim, uint64_t* d_blocks, uint64_t n_blocks,
                               uint32_t* d_totals, uint64_t* d_base, void* d_tmp, size_t& tmp_bytes, hipStream_t st) {
  if (d_tmp == nullptr) {
    return hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_totals, d_base, (int64_t)n_blocks, st);
  }
  if (dim > 0) {
    hipLaunchKernelGGL(k_bitmap_set, dim3(grid_for((uint64_t)dim)), dim3(kBlock), 0, st, d_keys, dim,
                       (unsigned long long*)d_blocks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_bitmap_block_counts, dim3(grid_for(n_blocks)), dim3(kBlock), 0, st, d_blocks, n_blocks, d_totals);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_totals, d_base, (int64_t)n_blocks, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bitmap_set_base, dim3(grid_for(n_blocks)), dim3(kBlock), 0, st, d_blocks, n_blocks, d_base);
  return hipGetLastError();
}

hipError_t launch_mark_tops(const uint64_t* d_keys, int64_t dim, int shift, uint32_t* d_flags, hipStream_t st) {
  if (dim == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_tops, dim3(grid_for((uint64_t)dim)), dim3(kBlock), 0, st, d_keys, dim, shift, d_flags);
  return hipGetLastError();
}

hipError_t launch_xmap_fill(int prec, const XMapDev& X, const uint64_t* d_keys, int64_t dim, const void* x,
                            const uint16_t* d_orbit, hipStream_t st) {
  if (dim == 0) return hipSuccess;
  dim3 grid(grid_for((uint64_t)dim));
  switch (prec) {
    case kF64:
      hipLaunchKernelGGL(k_xmap_fill<double>, grid, dim3(kBlock), 0, st, X, d_keys, dim, (const double*)x, d_orbit);
      break;
    case kC128:
      hipLaunchKernelGGL(k_xmap_fill<Cx<double>>, grid, dim3(kBlock), 0, st, X, d_keys, dim, (const Cx<double>*)x,
                         d_orbit);
      break;
    case kF32:
      hipLaunchKernelGGL(k_xmap_fill<float>, grid, dim3(kBlock), 0, st, X, d_keys, dim, (const float*)x, d_orbit);
      break;
    case kC64:
      hipLaunchKernelGGL(k_xmap_fill<Cx<float>>, grid, dim3(kBlock), 0, st, X, d_keys, dim, (const Cx<float>*)x,
                         d_orbit);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t exclusive_scan_u32_to_u64(const uint32_t* d_in, uint64_t* d_out, uint64_t n, void* d_tmp, size_t& tmp_bytes,
                                     hipStream_t st) {
  // d_out doubles as the widened input; the scan runs in place.
  if (d_tmp == nullptr) {
    return hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_out, d_out, (int64_t)n, st);
  }
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(n)), dim3(kBlock), 0, st, d_in, d_out, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_out, d_out, (int64_t)n, st);
}

hipError_t exclusive_scan_i64(const int64_t* d_in, int64_t* d_out, uint64_t n, void* d_tmp, size_t& tmp_bytes,
                              hipStream_t st) {
  if (d_tmp == nullptr) {
    return hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_in, d_out, (int64_t)n, st);
  }
  if (n == 0) return hipSuccess;
  return hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, d_in, d_out, (int64_t)n, st);
}

namespace {
template <typename V>
hipError_t vec_unreduce_t(const ReduceSource& src, const GroupDev& g, const GroupLdsLayout& lay, const IndexDev& ix,
                          const uint16_t* d_orbit, const void* small, void* large, hipStream_t st) {
  uint32_t lds = lay.total + (src.list ? 0u : w_bytes(src.n_bits));
  hipLaunchKernelGGL(k_vec_unreduce<V>, dim3(grid_for(src.n_words)), dim3(kBlock), lds, st, src, g, lay, ix, d_orbit,
                     (const V*)small, (V*)large);
  return hipGetLastError();
}

template <typename V>
hipError_t vec_reduce_t(const uint64_t* d_reps, const uint16_t* d_orbit, int64_t dim, const GroupDev& g,
                        const GroupLdsLayout& lay, const IndexDev& pix, const void* large, void* small, hipStream_t st) {
  uint32_t lds = lay.total + (pix.rank_mode ? w_bytes(pix.n_bits) : 0u);
  hipLaunchKernelGGL(k_vec_reduce<V>, dim3(grid_for((uint64_t)dim)), dim3(kBlock), lds, st, d_reps, d_orbit, dim, g, lay,
                     pix, (const V*)large, (V*)small);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_vec_unreduce(const ReduceSource& src, const GroupDev& g, const GroupLdsLayout& lay,
                               const IndexDev& reduced_index, const uint16_t* d_orbit, int prec, const void* d_small,
                               void* d_large, hipStream_t st) {
  if (src.n_words == 0) return hipSuccess;
  switch (prec) {
    case kF64:
      return vec_unreduce_t<double>(src, g, lay, reduced_index, d_orbit, d_small, d_large, st);
    case kC128:
      return vec_unreduce_t<Cx<double>>(src, g, lay, reduced_index, d_orbit, d_small, d_large, st);
    case kF32:
      return vec_unreduce_t<float>(src, g, lay, reduced_index, d_orbit, d_small, d_large, st);
    case kC64:
      return vec_unreduce_t<Cx<float>>(src, g, lay, reduced_index, d_orbit, d_small, d_large, st);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_vec_reduce(const uint64_t* d_reps, const uint16_t* d_orbit, int64_t dim, const GroupDev& g,
                             const GroupLdsLayout& lay, const IndexDev& parent_index, int prec, const void* d_large,
                             void* d_small, hipStream_t st) {
  if (dim == 0) return hipSuccess;
  switch (prec) {
    case kF64:
      return vec_reduce_t<double>(d_reps, d_orbit, dim, g, lay, parent_index, d_large, d_small, st);
    case kC128:
      return vec_reduce_t<Cx<double>>(d_reps, d_orbit, dim, g, lay, parent_index, d_large, d_small, st);
    case kF32:
      return vec_reduce_t<float>(d_reps, d_orbit, dim, g, lay, parent_index, d_large, d_small, st);
    case kC64:
      return vec_reduce_t<Cx<float>>(d_reps, d_orbit, dim, g, lay, parent_index, d_large, d_small, st);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace edhip
