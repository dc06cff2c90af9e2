// Deterministic device-wide reductions for wave64 (gfx950).
//
// Every reduction kernel runs a fixed grid; each block writes one partial in
// slot order and a one-block finaliser kernel folds the partials in a fixed
// order and applies the scalar logic (CG step lengths, convergence flags,
// bisection decisions).  Results are bit-identical from run to run for a
// given problem size, which the "same alone and in the mix" serving check
// relies on.
#pragma once

#include <hip/hip_runtime.h>

namespace t3x {

constexpr int kBlock = 256;
constexpr int kMaxRedBlocks = 2048;   // grid-stride reductions: blocks (x up to 8 values each)
constexpr int kMaxPartials = 65536;   // partials buffer (doubles): one-value reductions over bigger grids

// XCD-aware order of a 3-D grid of tiles.  Work-groups are dispatched
// round-robin over the 8 XCDs (linear dispatch slot b runs on XCD b % 8), each
// with its own L2; this maps slot b to tile t so that XCD x gets the
// contiguous range of tiles [x q + min(x, r), ...) in (x, y, z) order, and
// tiles that share halo cache lines share an L2.  A bijection on the grid.
__device__ __forceinline__ dim3 xcd_tile_order() {
  const int nb = gridDim.x * gridDim.y * gridDim.z;
  const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int q = nb / 8, r = nb % 8, x = b % 8;
  const int t = x * q + min(x, r) + b / 8;
  return dim3(t % gridDim.x, (t / gridDim.x) % gridDim.y, t / (gridDim.x * gridDim.y));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Block-wide sum of NV values per thread (result valid in thread 0); NT:
// threads per block.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[i][w] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int k = 0; k < NT / 64; ++k) s += sh[i][k];
      v[i] = s;
    }
  }
  __syncthreads();
}

// Producer side: block sums of NV values to partials[i*gridDim.x + blockIdx.x].
// No fence and no atomics: the consumer is the next kernel on the stream
// (device-scope fences write back the per-XCD L2 and cost tens of
// microseconds on a streaming kernel; a kernel boundary costs ~3 us).
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* partials) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) partials[(size_t)i * gridDim.x + blockIdx.x] = v[i];
  }
}

__device__ __forceinline__ void write_partial_max(double v, double* partials) {
  __shared__ double sh[kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_max(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int k = 1; k < kBlock / 64; ++k) m = fmax(m, sh[k]);
    partials[blockIdx.x] = m;
  }
}

// Consumer side (one block of NT threads): fixed-order fold of nb partials
// per value; totals valid in thread 0.  For one value each thread keeps 16
// loads in flight (the partials come from other XCDs' L2 through memory,
// ~1 us a round): 65536 partials fold in 4 rounds with 1024 threads.  Folds
// of many values keep one (their registers).
template <int NV, int NT = kBlock>
__device__ __forceinline__ void fold_partials(const double* partials, int nb, double (&out)[NV]) {
  constexpr int U = NV == 1 ? 16 : (NV == 2 ? 8 : 1);
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double* p = partials + (size_t)i * nb;
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = 0.0;
    int b = threadIdx.x;
    for (; b + (U - 1) * NT < nb; b += U * NT) {
      double t[U];
#pragma unroll
      for (int u = 0; u < U; ++u) t[u] = p[b + u * NT];
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] += t[u];
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u)  // fewer than U left
      if (b + u * NT < nb) a[u] += p[b + u * NT];
#pragma unroll
    for (int w = U / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int u = 0; u < w; ++u) a[u] += a[u + w];
    out[i] = a[0];
  }
  block_sum<NV, NT>(out);
}

__device__ __forceinline__ double fold_max(const double* partials, int nb) {
  __shared__ double sh[kBlock / 64];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) m = fmax(m, partials[b]);
  m = wave_max(m);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < kBlock / 64; ++k) m = fmax(m, sh[k]);
  if (threadIdx.x == 0) m = fmax(m, sh[0]);
  return m;
}

inline int red_blocks(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > kMaxRedBlocks) b = kMaxRedBlocks;
  if (b < 1) b = 1;
  return (int)b;
}
// One-value reductions over vectors (the PCG and filter-CG passes, dot
// products): up to kMaxPartials blocks; a reduction that may write into a
// caller's smaller partials buffer keeps red_blocks.
// Measured on MI355X at 101.7M elements (r -= l q with r.r and an FP32
// Jacobi step): 0.61 ms with 65536 blocks against 0.70 ms with 2048.
inline int red1_blocks(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > kMaxPartials) b = kMaxPartials;
  if (b < 1) b = 1;
  return (int)b;
}
// Element-wise kernels (grid-stride loops): one thread per element.  An
// HBM stream with many short-lived blocks runs faster than a grid that loops:
// y += l p, p = z + b p over 101.7M elements 0.62 ms with one element per
// thread, 0.78 ms with 8192 blocks looping.
inline int stream_blocks(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > (1LL << 30)) b = 1LL << 30;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace t3x
