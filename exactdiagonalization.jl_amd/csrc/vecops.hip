// Vector kernels of the Lanczos / MINRES drivers.  Vectors are stored in one
// of four precisions (f64, c128, f32, c64); every reduction accumulates in
// double and is deterministic: a fixed grid of blocks, a fixed in-block tree,
// then a fixed-order sum of the per-block partials.
#include "kernels.hpp"

namespace edhip {

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;  // fixed grid for deterministic reductions
constexpr int kTile = 8;          // vectors per multi-dot tile

template <typename V>
__device__ __forceinline__ void ld(const V* p, int64_t i, double& re, double& im) {
  if constexpr (VT<V>::cx) {
    re = (double)p[i].re;
    im = (double)p[i].im;
  } else {
    re = (double)p[i];
    im = 0.0;
  }
}

template <typename V>
__device__ __forceinline__ void st(V* p, int64_t i, double re, double im) {
  using R = typename VT<V>::R;
  if constexpr (VT<V>::cx) {
    p[i].re = (R)re;
    p[i].im = (R)im;
  } else {
    (void)im;
    p[i] = (R)re;
  }
}

__device__ __forceinline__ void block_sum2(double& a, double& b, double* sh) {
  // sh: 2*kBlock doubles
  sh[threadIdx.x] = a;
  sh[kBlock + threadIdx.x] = b;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[threadIdx.x] += sh[threadIdx.x + s];
      sh[kBlock + threadIdx.x] += sh[kBlock + threadIdx.x + s];
    }
    __syncthreads();
  }
  a = sh[0];
  b = sh[kBlock];
}

// partial[blk] = sum_i conj(x_i) y_i over the block's grid-stride slice
template <typename V>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const V* __restrict__ x, const V* __restrict__ y,
                                                double* __restrict__ partial) {
  __shared__ double sh[2 * kBlock];
  double sr = 0.0, si = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double xr, xi, yr, yi;
    ld(x, i, xr, xi);
    ld(y, i, yr, yi);
    sr += xr * yr + xi * yi;
    si += xr * yi - xi * yr;
  }
  block_sum2(sr, si, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr;
    partial[2 * blockIdx.x + 1] = si;
  }
}

// out[j] = sum_b partial[b*width + j] in a fixed order: block j; thread t
// adds the partials b = t, t + kBlock, ... in turn, then a fixed tree in LDS.
// (One thread per column, adding all nb partials in a chain of dependent
// loads, took ~150 us per reduction.)
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ partial, int nb, int width,
                                                         double* __restrict__ out) {
  __shared__ double sh[kBlock];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) s += partial[(int64_t)b * width + j];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sh[0];
}

// Tile of kTile vectors: partial[(blk)*(2k) + 2j..] = sum_i conj(V_j,i) w_i
template <typename V>
__global__ __launch_bounds__(kBlock) void k_multi_dot(int64_t n, int k, const V* __restrict__ Vm, int64_t ldv,
                                                      const V* __restrict__ w, double* __restrict__ partial) {
  __shared__ double sh[2 * kBlock];
  const int j0 = blockIdx.y * kTile;
  double ar[kTile], ai[kTile];
#pragma unroll
  for (int q = 0; q < kTile; ++q) ar[q] = ai[q] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double wr, wi;
    ld(w, i, wr, wi);
#pragma unroll
    for (int q = 0; q < kTile; ++q) {
      if (j0 + q < k) {
        double vr, vi;
        ld(Vm + (int64_t)(j0 + q) * ldv, i, vr, vi);
        ar[q] += vr * wr + vi * wi;
        ai[q] += vr * wi - vi * wr;
      }
    }
  }
  for (int q = 0; q < kTile; ++q) {
    double a = ar[q], b = ai[q];
    block_sum2(a, b, sh);
    if (threadIdx.x == 0 && j0 + q < k) {
      partial[(int64_t)blockIdx.x * 2 * k + 2 * (j0 + q)] = a;
      partial[(int64_t)blockIdx.x * 2 * k + 2 * (j0 + q) + 1] = b;
    }
    __syncthreads();
  }
}

// w -= sum_j coef_j V_j
template <typename V>
__global__ __launch_bounds__(kBlock) void k_multi_axpy(int64_t n, int k, const V* __restrict__ Vm, int64_t ldv,
                                                       const double* __restrict__ coef, V* __restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double wr, wi;
    ld(w, i, wr, wi);
    for (int j = 0; j < k; ++j) {
      double vr, vi;
      ld(Vm + (int64_t)j * ldv, i, vr, vi);
      const double cr = coef[2 * j], ci = coef[2 * j + 1];
      wr -= cr * vr - ci * vi;
      wi -= cr * vi + ci * vr;
    }
    st(w, i, wr, wi);
  }
}

// ---- 16-byte-load variants of the multi-dot / multi-axpy ----------------
// Used when every vector starts on a 16-byte boundary (the Krylov basis is
// stored with a padded stride): each lane moves 16 bytes (E elements) per
// load instead of one element.
template <typename V>
struct Pack16 {
  static constexpr int E = 16 / (int)sizeof(V);
  __device__ __forceinline__ static void load(const V* p, double re[E], double im[E]) {
    const uint4 u = *(const uint4*)p;
    const V* v = (const V*)&u;
#pragma unroll
    for (int e = 0; e < E; ++e) ld(v, e, re[e], im[e]);
  }
  __device__ __forceinline__ static void store(V* p, const double re[E], const double im[E]) {
    uint4 u;
    V* v = (V*)&u;
#pragma unroll
    for (int e = 0; e < E; ++e) st(v, e, re[e], im[e]);
    *(uint4*)p = u;
  }
};

__device__ __forceinline__ double wave_reduce(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// partial[blk * 2k + 2 (j0 + q) + {0,1}] = sum over this block's elements of
// conj(V_{j0+q}) w, for a tile of KT vectors (blockIdx.y); fixed grid, fixed
// order: deterministic.
template <typename V, int KT>
__global__ __launch_bounds__(kBlock) void k_multi_dot16(int64_t n, int k, const V* __restrict__ Vm, int64_t ldv,
                                                        const V* __restrict__ w, double* __restrict__ partial) {
  constexpr int E = Pack16<V>::E;
  constexpr bool CX = VT<V>::cx;
  __shared__ double sh[kBlock / 64][KT][2];
  const int j0 = blockIdx.y * KT;
  double ar[KT], ai[KT];
#pragma unroll
  for (int q = 0; q < KT; ++q) ar[q] = ai[q] = 0.0;
  const int64_t ng = n / E;
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < ng; g += (int64_t)gridDim.x * kBlock) {
    double wr[E], wi[E];
    Pack16<V>::load(w + g * E, wr, wi);
#pragma unroll
    for (int q = 0; q < KT; ++q) {
      if (j0 + q < k) {
        double vr[E], vi[E];
        Pack16<V>::load(Vm + (int64_t)(j0 + q) * ldv + g * E, vr, vi);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          ar[q] += vr[e] * wr[e] + vi[e] * wi[e];
          if constexpr (CX) ai[q] += vr[e] * wi[e] - vi[e] * wr[e];
        }
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < n - ng * E) {  // the last n mod E elements
    const int64_t i = ng * E + threadIdx.x;
    double wr, wi;
    ld(w, i, wr, wi);
#pragma unroll
    for (int q = 0; q < KT; ++q) {
      if (j0 + q < k) {
        double vr, vi;
        ld(Vm + (int64_t)(j0 + q) * ldv, i, vr, vi);
        ar[q] += vr * wr + vi * wi;
        if constexpr (CX) ai[q] += vr * wi - vi * wr;
      }
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < KT; ++q) {
    const double a = wave_reduce(ar[q]);
    const double b = CX ? wave_reduce(ai[q]) : 0.0;
    if (lane == 0) {
      sh[wave][q][0] = a;
      sh[wave][q][1] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * KT) {
    const int q = threadIdx.x >> 1, c = threadIdx.x & 1;
    if (j0 + q < k) {
      double v = 0.0;
      for (int wv = 0; wv < kBlock / 64; ++wv) v += sh[wv][q][c];
      partial[(int64_t)blockIdx.x * 2 * k + 2 * (j0 + q) + c] = v;
    }
  }
}

// w -= sum_j coef_j V_j, 16 bytes per lane and vector
template <typename V>
__global__ __launch_bounds__(kBlock) void k_multi_axpy16(int64_t n, int k, const V* __restrict__ Vm, int64_t ldv,
                                                         const double* __restrict__ coef, V* __restrict__ w) {
  constexpr int E = Pack16<V>::E;
  constexpr bool CX = VT<V>::cx;
  const int64_t ng = n / E;
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g < ng) {
    double wr[E], wi[E];
    Pack16<V>::load(w + g * E, wr, wi);
#pragma unroll 4
    for (int j = 0; j < k; ++j) {
      double vr[E], vi[E];
      Pack16<V>::load(Vm + (int64_t)j * ldv + g * E, vr, vi);
      const double cr = coef[2 * j], ci = CX ? coef[2 * j + 1] : 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        wr[e] -= cr * vr[e] - ci * vi[e];
        if constexpr (CX) wi[e] -= cr * vi[e] + ci * vr[e];
      }
    }
    Pack16<V>::store(w + g * E, wr, wi);
  } else if (g < ng + (n - ng * E)) {  // the last n mod E elements, one per lane
    const int64_t i = ng * E + (g - ng);
    double wr, wi;
    ld(w, i, wr, wi);
    for (int j = 0; j < k; ++j) {
      double vr, vi;
      ld(Vm + (int64_t)j * ldv, i, vr, vi);
      const double cr = coef[2 * j], ci = coef[2 * j + 1];
      wr -= cr * vr - ci * vi;
      wi -= cr * vi + ci * vr;
    }
    st(w, i, wr, wi);
  }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double ar, double ai, const V* __restrict__ x, double br,
                                                  double bi, V* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double xr = 0.0, xi = 0.0, yr = 0.0, yi = 0.0;
    if (x) ld(x, i, xr, xi);
    if (br != 0.0 || bi != 0.0) ld(y, i, yr, yi);
    const double r = ar * xr - ai * xi + br * yr - bi * yi;
    const double m = ar * xi + ai * xr + br * yi + bi * yr;
    st(y, i, r, m);
  }
}

template <typename VI, typename VO>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t n, const VI* __restrict__ x, VO* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double r, m;
    ld(x, i, r, m);
    st(y, i, r, m);
  }
}

// out_l[i] = sum_j coef[j*k + l] V_j[i], a tile of kTile outputs per blockIdx.y
template <typename VI, typename VO>
__global__ __launch_bounds__(kBlock) void k_combine(int64_t n, int m, const VI* __restrict__ Vm, int64_t ldv, int k,
                                                    const double* __restrict__ coef, VO* __restrict__ out, int64_t ldo) {
  const int l0 = blockIdx.y * kTile;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double ar[kTile], ai[kTile];
#pragma unroll
    for (int q = 0; q < kTile; ++q) ar[q] = ai[q] = 0.0;
    for (int j = 0; j < m; ++j) {
      double vr, vi;
      ld(Vm + (int64_t)j * ldv, i, vr, vi);
#pragma unroll
      for (int q = 0; q < kTile; ++q) {
        if (l0 + q < k) {
          const double c = coef[j * k + l0 + q];
          ar[q] += c * vr;
          ai[q] += c * vi;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kTile; ++q)
      if (l0 + q < k) st(out + (int64_t)(l0 + q) * ldo, i, ar[q], ai[q]);
  }
}

// SplitMix64 (Steele, Lea, Flood 2014) and Box-Muller: a counter-based
// Gaussian stream, identical for every launch shape.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t z) {  // (0, 1]
  return ((double)(z >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_gaussian(int64_t n, uint64_t seed, V* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint64_t base = seed ^ ((uint64_t)i * 0xD1B54A32D192ED03ull);
    const double u1 = u01(splitmix64(base));
    const double u2 = u01(splitmix64(base + 1));
    const double r = sqrt(-2.0 * log(u1));
    const double a = 6.283185307179586 * u2;
    st(x, i, r * cos(a), r * sin(a));
  }
}

inline unsigned grid_stride_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

template <template <typename> class F, typename... A>
hipError_t dispatch(int prec, A&&... a) {
  switch (prec) {
    case kF64:
      return F<double>::run(a...);
    case kC128:
      return F<Cx<double>>::run(a...);
    case kF32:
      return F<float>::run(a...);
    case kC64:
      return F<Cx<float>>::run(a...);
    default:
      return hipErrorInvalidValue;
  }
}

template <typename V>
struct DotOp {
  static hipError_t run(int64_t n, const void* x, const void* y, double* partial, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dot<V>, dim3(kRedBlocks), dim3(kBlock), 0, s, n, (const V*)x, (const V*)y, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(2), dim3(kBlock), 0, s, partial, kRedBlocks, 2, out);
    return hipGetLastError();
  }
};

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

template <typename V>
struct MultiDotOp {
  static hipError_t run(int64_t n, int k, const void* Vm, int64_t ldv, const void* w, double* partial, double* out,
                        hipStream_t s) {
    if (k <= 0) return hipSuccess;
    if (aligned16(Vm) && aligned16(w) && ((ldv * (int64_t)sizeof(V)) & 15) == 0) {
      constexpr int KT = VT<V>::cx ? 8 : 16;
      const int nb = kRedBlocks;
      dim3 grid(nb, (k + KT - 1) / KT);
      hipLaunchKernelGGL((k_multi_dot16<V, KT>), grid, dim3(kBlock), 0, s, n, k, (const V*)Vm, ldv, (const V*)w,
                         partial);
      const int width = 2 * k;
      hipLaunchKernelGGL(k_sum_partials, dim3(width), dim3(kBlock), 0, s, partial, nb, width,
                         out);
      return hipGetLastError();
    }
    int nb = kRedBlocks / 4;
    dim3 grid(nb, (k + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_multi_dot<V>, grid, dim3(kBlock), 0, s, n, k, (const V*)Vm, ldv, (const V*)w, partial);
    int width = 2 * k;
    hipLaunchKernelGGL(k_sum_partials, dim3(width), dim3(kBlock), 0, s, partial, nb, width, out);
    return hipGetLastError();
  }
};

template <typename V>
struct MultiAxpyOp {
  static hipError_t run(int64_t n, int k, const void* Vm, int64_t ldv, const double* coef, void* w, hipStream_t s) {
    if (aligned16(Vm) && aligned16(w) && ((ldv * (int64_t)sizeof(V)) & 15) == 0) {
      constexpr int E = Pack16<V>::E;
      const int64_t ng = n / E;
      const int64_t threads = ng + (n - ng * E);
      hipLaunchKernelGGL(k_multi_axpy16<V>, dim3((unsigned)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, n, k,
                         (const V*)Vm, ldv, coef, (V*)w);
      return hipGetLastError();
    }
    hipLaunchKernelGGL(k_multi_axpy<V>, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, n, k, (const V*)Vm, ldv, coef,
                       (V*)w);
    return hipGetLastError();
  }
};

template <typename V>
struct AxpbyOp {
  static hipError_t run(int64_t n, double ar, double ai, const void* x, double br, double bi, void* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby<V>, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, n, ar, ai, (const V*)x, br, bi,
                       (V*)y);
    return hipGetLastError();
  }
};

template <typename V>
struct GaussOp {
  static hipError_t run(int64_t n, uint64_t seed, void* x, hipStream_t s) {
    hipLaunchKernelGGL(k_gaussian<V>, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, n, seed, (V*)x);
    return hipGetLastError();
  }
};

template <typename VI>
struct ConvertFrom {
  template <typename VO>
  struct To {
    static hipError_t run(int64_t n, const void* x, void* y, hipStream_t s) {
      hipLaunchKernelGGL((k_convert<VI, VO>), dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, n, (const VI*)x, (VO*)y);
      return hipGetLastError();
    }
  };
};

template <typename VI>
struct CombineFrom {
  template <typename VO>
  struct To {
    static hipError_t run(int64_t n, int m, const void* V, int64_t ldv, int k, const double* coef, void* out,
                          int64_t ldo, hipStream_t s) {
      dim3 grid(grid_stride_blocks(n), (k + kTile - 1) / kTile);
      hipLaunchKernelGGL((k_combine<VI, VO>), grid, dim3(kBlock), 0, s, n, m, (const VI*)V, ldv, k, coef, (VO*)out, ldo);
      return hipGetLastError();
    }
  };
};

template <typename VI>
struct CombOp {
  static hipError_t run(int prec_out, int64_t n, int m, const void* V, int64_t ldv, int k, const double* coef,
                        void* out, int64_t ldo, hipStream_t s) {
    return dispatch<CombineFrom<VI>::template To>(prec_out, n, m, V, ldv, k, coef, out, ldo, s);
  }
};

template <typename VI>
struct ConvOp {
  static hipError_t run(int prec_out, int64_t n, const void* x, void* y, hipStream_t s) {
    return dispatch<ConvertFrom<VI>::template To>(prec_out, n, x, y, s);
  }
};

}  // namespace

int64_t vec_partial_blocks(int64_t) { return kRedBlocks; }

hipError_t vec_dot(int prec, int64_t n, const void* x, const void* y, double* d_partial, double* d_out2,
                   hipStream_t s) {
  return dispatch<DotOp>(prec, n, x, y, d_partial, d_out2, s);
}

hipError_t vec_multi_dot(int prec, int64_t n, int k, const void* Vm, int64_t ldv, const void* w, double* d_partial,
                         double* d_out, hipStream_t s) {
  if (k <= 0) return hipSuccess;
  return dispatch<MultiDotOp>(prec, n, k, Vm, ldv, w, d_partial, d_out, s);
}

hipError_t vec_multi_axpy(int prec, int64_t n, int k, const void* Vm, int64_t ldv, const double* d_coef, void* w,
                          hipStream_t s) {
  if (k <= 0) return hipSuccess;
  return dispatch<MultiAxpyOp>(prec, n, k, Vm, ldv, d_coef, w, s);
}

hipError_t vec_axpby(int prec, int64_t n, double ar, double ai, const void* x, double br, double bi, void* y,
                     hipStream_t s) {
  return dispatch<AxpbyOp>(prec, n, ar, ai, x, br, bi, y, s);
}

hipError_t vec_convert(int prec_in, int prec_out, int64_t n, const void* x, void* y, hipStream_t s) {
  return dispatch<ConvOp>(prec_in, prec_out, n, x, y, s);
}

hipError_t vec_combine(int prec_v, int prec_out, int64_t n, int m, const void* V, int64_t ldv, int k,
                       const double* d_coef, void* out, int64_t ldo, hipStream_t s) {
  if (k <= 0) return hipSuccess;
  return dispatch<CombOp>(prec_v, prec_out, n, m, V, ldv, k, d_coef, out, ldo, s);
}

hipError_t vec_gaussian(int prec, int64_t n, uint64_t seed, void* x, hipStream_t s) {
  return dispatch<GaussOp>(prec, n, seed, x, s);
}

hipError_t vec_fill_zero(int prec, int64_t n, void* x, hipStream_t s) {
  return dispatch<AxpbyOp>(prec, n, 0.0, 0.0, (const void*)nullptr, 0.0, 0.0, x, s);
}

}  // namespace edhip
