// Coarsest-level direct solver of libt3x (Solving_AssembleFEAstencil chol,
// TOP3D_XL.m:1052-1065, and the triangular solves of Solving_Vcycle, 767).
//
// The coarsest Galerkin matrix is banded after an axis-permuted node order
// (bandwidth bw).  The reference factors it with CHOLMOD; here:
//
//  1. right-looking blocked banded Cholesky, 64-column panels, one launch per
//     panel: every work-group of the trailing update solves the panel tiles
//     it needs (X = A L_JJ^{-T}, blocked by 16 columns on the matrix cores
//     with the inverses of L_JJ's 16x16 diagonal blocks) and applies its
//     X_I X_K^T update on the matrix cores; the group owning the next
//     diagonal tile then factors it (register/readlane POTRF), so the only
//     serial chain per panel is that tile's pivots.  Solved panel tiles go
//     to a scratch panel and reach the band in the next launch (the band's
//     A(I, J) is still being read during the current one).  The 64x64 tile
//     inverses for the solves are formed afterwards, in parallel.
//  2. k_trsm_blocks / k_h_blocks: the band is cut into solve blocks of SB >=
//     bw rows, so each block couples only to its predecessor:
//        forward   y_b = Dinv_b v_b - G_b y_{b-1},   G_b = Dinv_b L_{b,b-1}
//        backward  x_b = Dinv_b^T y_b - H_b x_{b+1}, H_b = Dinv_b^T L_{b+1,b}^T
//     Dinv_b, G_b come from one batched blocked TRSM on [I | L_{b,b-1}],
//     H_b from a batched GEMM.
//  3. the two recurrences as a chain of 2*nblk+1 small launches instead of
//     one dependent step per 64 rows.
//
// Kernel boundaries are the only cross-work-group synchronisation: on MI355X
// a device-wide barrier inside a kernel (cooperative grid sync, or fences that
// write back the per-XCD L2) costs more than a dependent launch (~3 us).
//
// All arithmetic is FP64.  Rows beyond the true size are identity rows.
#include <hip/hip_runtime.h>

#include "t3x_coarse.h"
#include "t3x_kernels.h"
#include "t3x_reduce.h"

#include <algorithm>

namespace t3x {

constexpr int TB = 64;      // tile size
constexpr int CT = 256;     // threads per work-group

struct BandView {
  double* B;  // row-major band: L(i, i-k) at B[i*W + k]
  int n;      // padded order
  int bw;
  __device__ __forceinline__ int W() const { return bw + 1; }
  __device__ __forceinline__ bool in(int i, int k) const { return i >= k && i - k <= bw && i < n && k >= 0; }
  // branch-free: the address is clamped into the band so loads can be batched
  __device__ __forceinline__ double get(int i, int k) const {
    const int d = i - k;
    const bool ok = (d >= 0) & (d <= bw) & (i < n) & (k >= 0);
    const size_t idx = (size_t)min(max(i, 0), n - 1) * W() + min(max(d, 0), bw);
    const double v = B[idx];
    return ok ? v : 0.0;
  }
  __device__ __forceinline__ void set(int i, int k, double v) const {
    if (in(i, k)) B[(size_t)i * W() + (i - k)] = v;
  }
};

// ---- 64x64 tile helpers in LDS -----------------------------------------
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(x) for a positive normal x: the hardware estimate and one Newton
// step (relative error below 1e-16 after the step; a factored 64x64 tile
// reproduces its matrix to 1.4e-15)
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double h = fma(-x * y, y, 1.0);  // 1 - x y^2
  return fma(y * h, 0.5, y);
}

typedef double cd4 __attribute__((ext_vector_type(4)));

// In-place Cholesky of the lower triangle of sA, blocked in 4 panels of 16
// columns.  The block's first wave factors a 64x16 panel in registers (lane i
// owns row i): per column the pivot and the column entries are broadcast with
// readlane (compile-time lanes), the pivot's reciprocal square root scales
// the column, and the rank-1 update stays in registers.  Every lane applies
// the same operations branch-free: a lane's entries right of its diagonal
// only ever hold values that are neither broadcast nor stored.  Then the
// panel's rank-16 update of the trailing triangle runs on the matrix cores,
// its 16x16 tiles round-robin over the 4 waves (the upper halves of the
// diagonal tiles are updated too and never read).  One 64x64 tile: 8.3 us
// against 12.8 us with the library rsqrt and 2x2 register-blocked updates.
__device__ __forceinline__ void tile_potrf(double (*sA)[TB + 1], int* status) {
  constexpr int PB = 16;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int b = 0; b < TB / PB; ++b) {
    const int b0 = b * PB;
    if (threadIdx.x < 64) {
      double p[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) p[k] = sA[lane][b0 + k];
      bool bad = false;
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        const double piv = readlane_d(p[k], b0 + k);
        bad |= !(piv > 0.0);
        p[k] *= rsqrt_nr(piv);  // the pivot's lane gets sqrt(piv), lanes below L(i, c)
#pragma unroll
        for (int kk = k + 1; kk < PB; ++kk) p[kk] -= p[k] * readlane_d(p[k], b0 + kk);
      }
      if (lane >= b0) {
#pragma unroll
        for (int k = 0; k < PB; ++k)
          if (b0 + k <= lane) sA[lane][b0 + k] = p[k];
      }
      if (bad && lane == 0) *status = 1;
    }
    __syncthreads();
    if (b0 + PB < TB) {  // trailing triangle: 16x16 tiles (I, J), J <= I, at rows/columns >= b0 + PB
      const int t0 = b0 + PB, m = (TB - t0) / 16;
      for (int t = w; t < m * (m + 1) / 2; t += CT / 64) {
        int I = 0;
        while ((I + 1) * (I + 2) / 2 <= t) ++I;
        const int J = t - I * (I + 1) / 2;
        const int R0 = t0 + 16 * I, C0 = t0 + 16 * J;
        cd4 acc;
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) acc[r4] = sA[R0 + kq + 4 * r4][C0 + li];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-sA[R0 + li][b0 + 4 * kb + kq], sA[C0 + li][b0 + 4 * kb + kq],
                                                     acc, 0, 0, 0);
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) sA[R0 + kq + 4 * r4][C0 + li] = acc[r4];
      }
      __syncthreads();
    }
  }
}

// Inverse of the lower-triangular sA (64x64) into sI, blocked in 16x16:
// wave w inverts diagonal block w (lane = column, forward substitution with
// reciprocal diagonals), then the off-diagonal blocks follow by block
// diagonals d = 1, 2, 3:  X_IJ = -X_II (sum_{J<=K<I} L_IK X_KJ), I = J + d.
__device__ __forceinline__ void tile_trinv(const double (*sA)[TB + 1], double (*sI)[TB + 1], double (*sT)[TB + 1]) {
  constexpr int PB = 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < TB * TB; t += CT) sI[t / TB][t % TB] = 0.0;
  __syncthreads();
  if (lane < PB) {
    const int o = wave * PB, c = lane;
    double x[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) s -= sA[o + i][o + k] * x[k];
      x[i] = (i >= c) ? s / sA[o + i][o + i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < PB; ++i) sI[o + i][o + c] = x[i];
  }
  __syncthreads();
  for (int d = 1; d < TB / PB; ++d) {
    const int nb = TB / PB - d;  // blocks (J + d, J), J = 0 .. nb-1
    for (int t = threadIdx.x; t < nb * PB * PB; t += CT) {  // T = sum_K L_IK X_KJ
      const int J = t / (PB * PB), a = (t / PB) % PB, c = t % PB, I = J + d;
      double s = 0.0;
      for (int k = J * PB; k < I * PB; ++k) s += sA[I * PB + a][k] * sI[k][J * PB + c];
      sT[I * PB + a][J * PB + c] = s;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nb * PB * PB; t += CT) {  // X_IJ = -X_II T
      const int J = t / (PB * PB), a = (t / PB) % PB, c = t % PB, I = J + d;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < PB; ++k) s += sI[I * PB + a][I * PB + k] * sT[I * PB + k][J * PB + c];
      sI[I * PB + a][J * PB + c] = -s;
    }
    __syncthreads();
  }
}


// acc = A B^T on v_mfma_f64_16x16x4 (A, B: 64x64 row-major in LDS); wave w
// owns the 32x32 quarter (32 (w / 2), 32 (w % 2)) as 2x2 tiles of 16x16.
// D[r][c] of a tile at lane c + 16 (r % 4), register r / 4.
__device__ __forceinline__ void tile_gemm_nt_mfma(const double (*sA)[TB + 1], const double (*sB)[TB + 1],
                                                  cd4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, kq = lane >> 4;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
#pragma unroll
  for (int kb = 0; kb < TB / 4; ++kb) {
    const double a0 = sA[r0 + i][4 * kb + kq], a1 = sA[r0 + 16 + i][4 * kb + kq];
    const double b0 = sB[c0 + i][4 * kb + kq], b1 = sB[c0 + 16 + i][4 * kb + kq];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}
// (row, col) in the 64x64 product of element (a, b, r4) of tile_gemm_nt_mfma's acc
__device__ __forceinline__ int nt_row(int a, int r4) {
  return 32 * ((threadIdx.x >> 6) >> 1) + 16 * a + ((threadIdx.x & 63) >> 4) + 4 * r4;
}
__device__ __forceinline__ int nt_col(int b) { return 32 * ((threadIdx.x >> 6) & 1) + 16 * b + (threadIdx.x & 15); }
// acc += A B (A, B: 64x64 row-major in LDS), same output layout as above.
__device__ __forceinline__ void tile_gemm_nn_mfma(const double (*sA)[TB + 1], const double (*sB)[TB + 1],
                                                  cd4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, kq = lane >> 4;
  const int r0 = 32 * (w >> 1), c0 = 32 * (w & 1);
#pragma unroll
  for (int kb = 0; kb < TB / 4; ++kb) {
    const double a0 = sA[r0 + i][4 * kb + kq], a1 = sA[r0 + 16 + i][4 * kb + kq];
    const double b0 = sB[4 * kb + kq][c0 + i], b1 = sB[4 * kb + kq][c0 + 16 + i];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// Inverses of the four 16x16 diagonal blocks of the lower-triangular sL into
// sD[4][16][16]: wave w inverts block w, lane c (< 16) one column, by
// forward substitution with independent updates after each step; the
// reciprocal diagonal is formed first (lane k), off the dependency chain.
__device__ __forceinline__ void diag16_inverses(const double (*sL)[TB + 1], double (*sD)[16][17]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, o = 16 * w;
  const double rdl = lane < 16 ? 1.0 / sL[o + lane][o + lane] : 0.0;
  if (lane < 16) {
    const int c = lane;
    double bvec[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) bvec[i] = (i == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const double xk = bvec[k] * readlane_d(rdl, k);  // zero above the column's diagonal
      bvec[k] = xk;
#pragma unroll
      for (int i = k + 1; i < 16; ++i) bvec[i] -= sL[o + i][o + k] * xk;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) sD[w][i][c] = bvec[i];
  }
}

// X = A L^{-T} in place (sX: A on entry, X on exit), blocked by 16 columns on
// the matrix cores:  R_b = A_b - X_{<b} L_{b,<b}^T,  X_b = R_b Dinv_b^T,
// Dinv_b the inverses of L's diagonal blocks.  Wave w owns rows 16w..16w+15
// throughout, so the waves need no barrier between the steps.
__device__ __forceinline__ void tile_trsm_rt(double (*sX)[TB + 1], const double (*sL)[TB + 1],
                                             const double (*sD)[16][17]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, kq = lane >> 4;
  const int r0 = 16 * w;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    cd4 acc;
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) acc[r4] = sX[r0 + kq + 4 * r4][16 * b + i];  // D layout: row kq + 4 r4, col i
#pragma unroll
    for (int kb = 0; kb < 4 * b; ++kb)  // -= X_{<b} L_{b,<b}^T
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-sX[r0 + i][4 * kb + kq], sL[16 * b + i][4 * kb + kq], acc, 0, 0, 0);
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) sX[r0 + kq + 4 * r4][16 * b + i] = acc[r4];  // R_b
    cd4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)  // X_b = R_b Dinv_b^T
      x = __builtin_amdgcn_mfma_f64_16x16x4f64(sX[r0 + i][16 * b + 4 * kb + kq], sD[b][i][4 * kb + kq], x, 0, 0, 0);
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) sX[r0 + kq + 4 * r4][16 * b + i] = x[r4];
  }
}

// Stages a 64x64 tile of the band (rows i0.., columns k0..) into LDS: all 16
// loads of a thread are issued before the stores.
__device__ __forceinline__ void stage_tile(BandView V, int i0, int k0, double (*sT)[TB + 1]) {
  double v[TB * TB / CT];
#pragma unroll
  for (int u = 0; u < TB * TB / CT; ++u) {
    const int t = threadIdx.x + u * CT;
    v[u] = V.get(i0 + t / TB, k0 + t % TB);
  }
#pragma unroll
  for (int u = 0; u < TB * TB / CT; ++u) {
    const int t = threadIdx.x + u * CT;
    sT[t / TB][t % TB] = v[u];
  }
}
// Two or three tiles at once (one memory round trip): (i0, k0) -> s0,
// (i1, k1) -> s1 and, with THREE, (i2, k2) -> s2.  THREE is a compile-time
// choice: a load under a run-time condition makes the compiler wait for the
// loads before it at every branch.
template <bool THREE>
__device__ __forceinline__ void stage_tiles(BandView V, int i0, int k0, double (*s0)[TB + 1], int i1, int k1,
                                            double (*s1)[TB + 1], int i2, int k2, double (*s2)[TB + 1]) {
  constexpr int U = TB * TB / CT;
  double v0[U], v1[U], v2[THREE ? U : 1];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = threadIdx.x + u * CT, r = t / TB, c = t % TB;
    v0[u] = V.get(i0 + r, k0 + c);
    v1[u] = V.get(i1 + r, k1 + c);
    if constexpr (THREE) v2[u] = V.get(i2 + r, k2 + c);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = threadIdx.x + u * CT, r = t / TB, c = t % TB;
    s0[r][c] = v0[u];
    s1[r][c] = v1[u];
    if constexpr (THREE) s2[r][c] = v2[u];
  }
}

// ---- 1. banded Cholesky ---------------------------------------------------
__global__ __launch_bounds__(CT) void k_chol_diag(BandView V, int d0, int* status) {
  __shared__ double sA[TB][TB + 1];
  for (int t = threadIdx.x; t < TB * TB; t += CT) sA[t / TB][t % TB] = V.get(d0 + t / TB, d0 + t % TB);
  __syncthreads();
  tile_potrf(sA, status);
  for (int t = threadIdx.x; t < TB * TB; t += CT) {
    const int i = t / TB, k = t % TB;
    if (k <= i) V.set(d0 + i, d0 + k, sA[i][k]);
  }
}

// One launch per 64-column panel J (right-looking, the panel solve folded into
// the trailing update).  Work-group (I, K), K <= I, over the row tiles below
// diagonal tile J that the band reaches:
//   X_I = A(I, J) L_JJ^{-T}, X_K = A(K, J) L_JJ^{-T}   (substitution in LDS)
//   A(I, K) -= X_I X_K^T                               (matrix cores)
// The (I, I) groups keep L(I, J) = X_I in a scratch panel (the band's copy of
// A(I, J) is still being read by the other groups); the next launch's extra
// groups copy it into the band.  Group (0, 0) updates the next diagonal tile
// and factors it.
__global__ __launch_bounds__(CT) void k_chol_panel(BandView V, int J, int nrt, int nrtPrev,
                                                   double* __restrict__ pan, int panStride, int* status) {
  __shared__ double sL[TB][TB + 1];  // L_JJ (zeros above the diagonal)
  __shared__ double sA[TB][TB + 1];  // A(I, J) -> X_I
  __shared__ double sB[TB][TB + 1];  // A(K, J) -> X_K, or the next diagonal tile
  __shared__ double sD[4][16][17];   // inverses of L_JJ's 16x16 diagonal blocks
  const int bw = V.bw, n = V.n;
  const int j0 = J * TB, r0 = j0 + TB;
  const int npairs = nrt * (nrt + 1) / 2;
  if ((int)blockIdx.x >= npairs) {  // copy row tile t of panel J - 1 into the band
    const int t = blockIdx.x - npairs;
    if (t >= nrtPrev) return;
    const double* src = pan + (size_t)((J - 1) & 1) * panStride + (size_t)t * TB * TB;
    const int i0 = j0 + t * TB, k0 = j0 - TB;
    for (int e = threadIdx.x; e < TB * TB; e += CT) V.set(i0 + e / TB, k0 + e % TB, src[e]);
    return;
  }
  const int t0 = blockIdx.x;
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= t0) ++I;
  const int K = t0 - I * (I + 1) / 2;
  const int i0 = r0 + I * TB, k0 = r0 + K * TB;
  if (i0 - (k0 + TB - 1) > bw) return;  // tile outside the band
  // L_JJ (zeros above the diagonal), A(I, J), and A(K, J) or, for group
  // (0, 0), the next diagonal tile (no other group of this launch writes it)
  if (K != I)
    stage_tiles<true>(V, j0, j0, sL, i0, j0, sA, k0, j0, sB);
  else if (I == 0)
    stage_tiles<true>(V, j0, j0, sL, i0, j0, sA, r0, r0, sB);
  else
    stage_tiles<false>(V, j0, j0, sL, i0, j0, sA, 0, 0, nullptr);
  __syncthreads();
  diag16_inverses(sL, sD);
  __syncthreads();
  tile_trsm_rt(sA, sL, sD);
  if (K != I) tile_trsm_rt(sB, sL, sD);
  __syncthreads();
  if (K == I) {  // L(I, J) to the scratch panel
    double* dst = pan + (size_t)(J & 1) * panStride + (size_t)I * TB * TB;
    for (int t = threadIdx.x; t < TB * TB; t += CT) dst[t] = sA[t / TB][t % TB];
  }
  cd4 acc[2][2] = {};
  tile_gemm_nt_mfma(sA, K == I ? sA : sB, acc);  // X_I X_K^T
  if (I == 0) {  // the next diagonal tile (staged in sB): update in LDS, then factor
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) sB[nt_row(a, r4)][nt_col(b)] -= acc[a][b][r4];
    __syncthreads();
    tile_potrf(sB, status);
    for (int t = threadIdx.x; t < TB * TB; t += CT) {
      const int i = t / TB, k = t % TB;
      if (k <= i) V.set(r0 + i, r0 + k, sB[i][k]);
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const int i = i0 + nt_row(a, r4), k = k0 + nt_col(b);
        if (k <= i && i - k <= bw && i < n) V.B[(size_t)i * V.W() + (i - k)] -= acc[a][b][r4];
      }
}

// Inverses of all 64x64 diagonal tiles (after the factorization, in parallel).
__global__ __launch_bounds__(CT) void k_tile_inverses(BandView V, double* __restrict__ tinv) {
  __shared__ double sA[TB][TB + 1];
  __shared__ double sI[TB][TB + 1];
  __shared__ double sT[TB][TB + 1];
  const int d0 = blockIdx.x * TB;
  for (int t = threadIdx.x; t < TB * TB; t += CT) sA[t / TB][t % TB] = V.get(d0 + t / TB, d0 + t % TB);
  __syncthreads();
  tile_trinv(sA, sI, sT);
  double* inv = tinv + (size_t)blockIdx.x * TB * TB;
  for (int t = threadIdx.x; t < TB * TB; t += CT) inv[t] = sI[t / TB][t % TB];
}

// ---- 2. solve-block operators -------------------------------------------
// WG (b, p): columns p*64.. of X = L_bb^{-1} [I | L_{b,b-1}], SB x 2SB.
// Output: Dinv_b (SB x SB row-major) and its transpose DT_b, G_b (SB x SB).
__global__ __launch_bounds__(CT) void k_trsm_blocks(BandView V, int SB, const double* __restrict__ tinv,
                                                    double* __restrict__ Dinv, double* __restrict__ DT,
                                                    double* __restrict__ Gm) {
  // two tiles of LDS (66.6 KB: two work-groups per CU); the right-hand side
  // minus the accumulated products is formed in sX once the products are done
  __shared__ double sL[TB][TB + 1];
  __shared__ double sX[TB][TB + 1];
  // (dispatch order: the strips' work differs, from one tile pair to st(st+1)/2;
  // an XCD-contiguous order measured 2.44 against 2.34 ms at 256x512x256)
  const int npan = 2 * SB / TB;
  const int b = blockIdx.x / npan, p = blockIdx.x % npan;
  const int st = SB / TB;
  const bool ident = p < st;
  if (!ident && b == 0) return;  // G_0 is unused
  const int base = b * SB;
  double* Out = ident ? Dinv + (size_t)b * SB * SB : Gm + (size_t)b * SB * SB;
  const int col0 = ident ? p * TB : (p - st) * TB;
  const int Ist = ident ? p : 0;  // rows above the diagonal panel of L^{-1} are zero
  for (int I = 0; I < st; ++I) {
    if (I < Ist) {
      for (int t = threadIdx.x; t < TB * TB; t += CT) {
        const int a = t / TB, c = t % TB;
        Out[(size_t)(I * TB + a) * SB + col0 + c] = 0.0;
      }
      continue;
    }
    // T = R_I - sum_{K<I} L_bb(I,K) X_K
    cd4 acc[2][2] = {};
    for (int K = Ist; K < I; ++K) {
      if ((I * TB) - (K * TB + TB - 1) > V.bw) continue;  // L tile outside the band
      stage_tile(V, base + I * TB, base + K * TB, sL);
      for (int t = threadIdx.x; t < TB * TB; t += CT) sX[t / TB][t % TB] = Out[(size_t)(K * TB + t / TB) * SB + col0 + t % TB];
      __syncthreads();
      tile_gemm_nn_mfma(sL, sX, acc);
      __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
          const int ra = nt_row(a, r4), cc = nt_col(c);
          const double r = ident ? ((I * TB + ra == col0 + cc) ? 1.0 : 0.0)
                                 : V.get(base + I * TB + ra, base - SB + col0 + cc);
          sX[ra][cc] = r - acc[a][c][r4];
        }
    // X_I = inv(L_II) T
    const double* inv = tinv + (size_t)(base / TB + I) * TB * TB;
    for (int t = threadIdx.x; t < TB * TB; t += CT) sL[t / TB][t % TB] = inv[t];
    __syncthreads();
    {
      cd4 acc2[2][2] = {};
      tile_gemm_nn_mfma(sL, sX, acc2);  // inv(L_II) is lower triangular: the zeros add nothing
#pragma unroll
      for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
          for (int r4 = 0; r4 < 4; ++r4) {
            const int a = nt_row(a2, r4), c = nt_col(c2);
            Out[(size_t)(I * TB + a) * SB + col0 + c] = acc2[a2][c2][r4];
            if (ident) DT[(size_t)b * SB * SB + (size_t)(col0 + c) * SB + I * TB + a] = acc2[a2][c2][r4];
          }
    }
    __syncthreads();
  }
}

// H_b = Dinv_b^T L_{b+1,b}^T, i.e. H_b(i,j) = sum_k DT_b(i,k) L((b+1)SB + j, b SB + k).
// One work-group per (b, output tile I, J).
__global__ __launch_bounds__(CT) void k_h_blocks(BandView V, int SB, int nblk, const double* __restrict__ DT,
                                                 double* __restrict__ H) {
  __shared__ double sA[TB][TB + 1];
  __shared__ double sB[TB][TB + 1];
  const int st = SB / TB;
  const int per = st * st;
  // XCD-contiguous order: one solve block's output tiles, which read the same
  // DT_b and band tiles, run on one XCD and share its L2 (0.90 -> 0.75 ms at
  // 256x512x256)
  const int wg = xcd_tile_order().x;
  const int b = wg / per, rem = wg % per;
  const int I = rem / st, J = rem % st;
  if (b >= nblk - 1) return;
  const double* D = DT + (size_t)b * SB * SB;
  cd4 acc[2][2] = {};
  for (int K = I; K < st; ++K) {  // DT_b(i,k) = Dinv_b(k,i) is zero for k < i
    const int rowL0 = (b + 1) * SB + J * TB, colL0 = b * SB + K * TB;
    if (rowL0 - (colL0 + TB - 1) > V.bw) continue;
    for (int t = threadIdx.x; t < TB * TB; t += CT) sA[t / TB][t % TB] = D[(size_t)(I * TB + t / TB) * SB + K * TB + t % TB];
    stage_tile(V, rowL0, colL0, sB);  // row j, col k
    __syncthreads();
    tile_gemm_nt_mfma(sA, sB, acc);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4)
        H[(size_t)b * SB * SB + (size_t)(I * TB + nt_row(a, r4)) * SB + J * TB + nt_col(c)] = acc[a][c][r4];
}

// ---- 3. the solve --------------------------------------------------------
// One wave per row dot product: out[row] = base - sum_j M[row][j] v[j].
__device__ __forceinline__ double wave_dot(const double* __restrict__ m, const double* __restrict__ v, int len) {
  const int lane = threadIdx.x & 63;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int j = lane;
  for (; j + 192 < len; j += 256) {  // four independent loads in flight per lane
    s0 += m[j] * v[j];
    s1 += m[j + 64] * v[j + 64];
    s2 += m[j + 128] * v[j + 128];
    s3 += m[j + 192] * v[j + 192];
  }
  for (; j < len; j += 64) s0 += m[j] * v[j];
  double s = (s0 + s1) + (s2 + s3);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// forward phase 0: gather the permuted right-hand side and va_b = Dinv_b v_b
template <typename TR>
__global__ __launch_bounds__(CT) void k_solve_f0(int n, int np, int SB, const long long* __restrict__ map,
                                                 const TR* __restrict__ rdense, const double* __restrict__ Dinv,
                                                 double* __restrict__ va, const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int row = blockIdx.x * (CT / 64) + (threadIdx.x >> 6);
  if (row >= np) return;
  const int b = row / SB, i = row % SB, lane = threadIdx.x & 63;
  const double* m = Dinv + (size_t)b * SB * SB + (size_t)i * SB;
  double s = 0.0;
  for (int j = lane; j <= i; j += 64) {
    const int g = b * SB + j;
    s += m[j] * (g < n ? (double)rdense[map[g]] : 0.0);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) va[row] = s;
}
// v[b] -= M_b v[b + d] for the SB rows of block b (d = -1 forward, +1 backward)
__global__ __launch_bounds__(CT) void k_solve_step(int SB, int b, int d, const double* __restrict__ M,
                                                   double* __restrict__ v, const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int i = blockIdx.x * (CT / 64) + (threadIdx.x >> 6);
  if (i >= SB) return;
  const double s = wave_dot(M + (size_t)b * SB * SB + (size_t)i * SB, v + (size_t)(b + d) * SB, SB);
  if ((threadIdx.x & 63) == 0) v[(size_t)b * SB + i] -= s;
}
// backward phase 0: vb_b = Dinv_b^T y_b
__global__ __launch_bounds__(CT) void k_solve_b0(int np, int SB, const double* __restrict__ DT,
                                                 const double* __restrict__ va, double* __restrict__ vb,
                                                 const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int row = blockIdx.x * (CT / 64) + (threadIdx.x >> 6);
  if (row >= np) return;
  const int b = row / SB, i = row % SB;
  const double s = wave_dot(DT + (size_t)b * SB * SB + (size_t)i * SB + i, va + (size_t)b * SB + i, SB - i);
  if ((threadIdx.x & 63) == 0) vb[row] = s;
}
template <typename TX>
__global__ void k_solve_scatter(int n, const long long* __restrict__ map, const double* __restrict__ vb,
                                TX* __restrict__ xdense, const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int i = blockIdx.x * CT + threadIdx.x;
  if (i < n) xdense[map[i]] = (TX)vb[i];
}

// ---- 3b. the recurrences without a launch per block ------------------------
// A launch per block costs ~7.5 us of dispatch, ramp and drain against
// ~1.5 us of data.  Here one launch per recurrence (forward after
// k_solve_f0, backward after k_solve_b0) walks all its blocks: G work-groups
// of 4 waves, wave w of the grid owning row w of every block.  A block's rows
// carry their own readiness: each FP64 entry travels as two 64-bit words (32
// bits of the value, 32 bits of tag), written with agent-scope stores
// (coherent across the XCDs' L2s), and a group that needs block b - 1
// (forward) or b + 1 (backward) polls those words until every entry shows the
// tag of that block in this launch.  The row of G_b (or H_b) and the row's
// own term for the next block are loaded right after a block is written, so
// they arrive while the group waits; a step is then one poll, one 64-lane dot
// product and one store.  Tags: a launch pair takes its first tag from
// ChainState (the backward launch advances every group's word by nblk),
// block b carries first + b + 1 forward and first + nblk - b backward, so
// words left by an earlier launch never match.  All G = ceil(SB / 4) groups
// must be resident together: 4 waves and SB doubles of LDS each, far below
// the device's capacity.  Each wait is bounded: after kChainSpin polls the
// launch sets fail, every group stops waiting (the solve is then garbage) and
// the caller reports the error.  Measured alternatives (same operators):
// one launch for everything with a published count per block, 7.6 us per
// block; the same with an agent-scope release/acquire pair, which writes back
// and invalidates the L2, 30 us per block.
constexpr unsigned kChainSpin = 1u << 20;  // polls per wait (~1 s) before the launch gives up

__device__ __forceinline__ double chain_wave_sum(double s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}
// entry i of a tagged vector: words 2i (high half of the value) and 2i + 1 (low half)
__device__ __forceinline__ void chain_put(unsigned long long* w, int i, double v, unsigned tag) {
  const unsigned long long d = (unsigned long long)__double_as_longlong(v);
  __hip_atomic_store(w + 2 * i, (d & 0xFFFFFFFF00000000ull) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 2 * i + 1, (d << 32) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FWD: y_b = a_b - G_b y_{b-1} (a = Dinv_b v_b from k_solve_f0, overwritten
// by y for k_solve_b0); backward: x_b = a_b - H_b x_{b+1} (a = Dinv_b^T y_b
// from k_solve_b0), scattered to xdense as formed.  PT: block entries per
// thread (SB <= PT * CT).
template <typename TV, int PT, bool FWD>
__global__ __launch_bounds__(CT) void k_solve_chain(int n, int SB, int nblk, const long long* __restrict__ map,
                                                    TV* __restrict__ xdense, const double* __restrict__ M,
                                                    double* __restrict__ a, unsigned long long* W, ChainState* st,
                                                    const int* fl) {
  if (fl && fl[FL_DONE]) return;
  constexpr int NCH = PT * CT / 64;  // 64-column chunks of a row
  extern __shared__ double chainNbr[];  // the neighbour block
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (CT / 64) + (threadIdx.x >> 6);
  const bool act = row < SB;
  const unsigned tag0 = st->tag[blockIdx.x];  // written by this group's previous backward launch
  auto tag_of = [&](int b) { return FWD ? tag0 + b + 1 : tag0 + (nblk - b); };
  const int step = FWD ? 1 : -1;
  auto in_range = [&](int b) { return b >= 0 && b < nblk; };
  double rg[NCH], ab = 0.0;
  auto prefetch = [&](int b) {  // row of M_b and this row's own term, all loads in flight together
    const double* m = M + ((size_t)b * SB + row) * SB;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int j = lane + 64 * c;
      rg[c] = j < SB ? m[j] : 0.0;
    }
    ab = a[(size_t)b * SB + row];
  };
  auto emit = [&](int b, double v) {
    const int g = b * SB + row;
    chain_put(W, g, v, tag_of(b));
    if (FWD) a[g] = v;
    else if (g < n) xdense[map[g]] = (TV)v;
  };
  // the first block is its own term
  int b = FWD ? 0 : nblk - 1;
  if (act) {
    ab = a[(size_t)b * SB + row];
    if (lane == 0) emit(b, ab);
  }
  b += step;
  if (in_range(b) && act) prefetch(b);
  for (; in_range(b); b += step) {
    {  // block b - step into LDS once every entry carries its tag
      const int nb = b - step;
      const unsigned tag = tag_of(nb);
      unsigned pending = 0;
#pragma unroll
      for (int k = 0; k < PT; ++k)
        if (threadIdx.x + k * CT < SB) pending |= 1u << k;
      double v[PT];
      unsigned polls = 0;
      while (pending) {
        unsigned long long hi[PT], lo[PT];
#pragma unroll
        for (int k = 0; k < PT; ++k)
          if (pending & (1u << k)) {
            const int i = nb * SB + threadIdx.x + k * CT;
            hi[k] = __hip_atomic_load(W + 2 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            lo[k] = __hip_atomic_load(W + 2 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
#pragma unroll
        for (int k = 0; k < PT; ++k)
          if ((pending & (1u << k)) && (unsigned)hi[k] == tag && (unsigned)lo[k] == tag) {
            v[k] = __longlong_as_double((long long)((hi[k] & 0xFFFFFFFF00000000ull) | (lo[k] >> 32)));
            pending &= ~(1u << k);
          }
        if (!pending) break;
        if (__hip_atomic_load(&st->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (++polls > kChainSpin) {
          __hip_atomic_store(&st->fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
#pragma unroll
      for (int k = 0; k < PT; ++k)
        if (threadIdx.x + k * CT < SB) chainNbr[threadIdx.x + k * CT] = (pending & (1u << k)) ? 0.0 : v[k];
    }
    __syncthreads();
    if (act) {
      // wave_dot's order over the registered row (four running sums over
      // chunk quadruples while the whole quadruple is in the row, then the
      // remaining chunks into the first), so the result is bitwise that of
      // k_solve_step
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      bool quad = true;
#pragma unroll
      for (int q = 0; q < NCH / 4; ++q) {
        const int j = lane + 256 * q;
        quad = quad && j + 192 < SB;
        if (quad) {
          s0 += rg[4 * q] * chainNbr[j];
          s1 += rg[4 * q + 1] * chainNbr[j + 64];
          s2 += rg[4 * q + 2] * chainNbr[j + 128];
          s3 += rg[4 * q + 3] * chainNbr[j + 192];
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (j + 64 * r < SB) s0 += rg[4 * q + r] * chainNbr[j + 64 * r];
        }
      }
      const double s = ab - chain_wave_sum((s0 + s1) + (s2 + s3));
      if (lane == 0) emit(b, s);
      if (in_range(b + step)) prefetch(b + step);
    }
    __syncthreads();  // chainNbr free
  }
  if (!FWD && threadIdx.x == 0) st->tag[blockIdx.x] = tag0 + nblk;
}

// ---- host side -----------------------------------------------------------
long long CoarseSolver::chain_share() const {
  constexpr long long kNever = 2000000;
  if (SB > kChainMaxSB) return kNever;
  const int G = (SB + 3) / 4;
  const size_t lds = (size_t)SB * sizeof(double);
  int dev = 0, cus = 0, perCU = 0, perCU2 = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kNever;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return kNever;
  auto fwd = SB <= 4 * CT ? k_solve_chain<double, 4, true> : k_solve_chain<double, kChainMaxSB / CT, true>;
  auto bwd = SB <= 4 * CT ? k_solve_chain<float, 4, false> : k_solve_chain<float, kChainMaxSB / CT, false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fwd, CT, lds) != hipSuccess) return kNever;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU2, bwd, CT, lds) != hipSuccess) return kNever;
  const long long slots = (long long)std::min(perCU, perCU2) * cus;
  if (slots <= 0) return kNever;
  return (1000000LL * G + slots - 1) / slots;
}

void CoarseSolver::plan(int n_, int bw_) {
  n = n_;
  bw = bw_;
  SB = ((bw + 1 + TB - 1) / TB) * TB;
  if (SB < TB) SB = TB;
  np = ((n + SB - 1) / SB) * SB;
  nblk = np / SB;
}

size_t CoarseSolver::band_len() const { return (size_t)np * (bw + 1); }

void CoarseSolver::factor(hipStream_t s, double* band, double* tinv, double* pan, double* Dinv, double* DT,
                          double* Gm, double* H, int* status) {
  BandView V{band, np, bw};
  const int nt = np / TB;
  const int panStride = (int)(panel_len() / 2);
  hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(CT), 0, s, V, 0, status);
  int nrtPrev = 0;
  for (int J = 0; J < nt; ++J) {
    const int j0 = J * TB, r0 = j0 + TB;
    const int rlast = std::min(np - 1, j0 + TB - 1 + bw);
    const int nrt = rlast >= r0 ? (rlast - r0) / TB + 1 : 0;
    const int grid = nrt * (nrt + 1) / 2 + nrtPrev;
    if (nrt > 0 || nrtPrev > 0)
      hipLaunchKernelGGL(k_chol_panel, dim3(grid), dim3(CT), 0, s, V, J, nrt, nrtPrev, pan, panStride, status);
    if (nrt == 0 && J + 1 < nt) hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(CT), 0, s, V, r0, status);
    nrtPrev = nrt;
  }
  if (nrtPrev > 0)  // copy the last panel
    hipLaunchKernelGGL(k_chol_panel, dim3(nrtPrev), dim3(CT), 0, s, V, nt, 0, nrtPrev, pan, panStride, status);
  hipLaunchKernelGGL(k_tile_inverses, dim3(nt), dim3(CT), 0, s, V, tinv);
  const int st = SB / TB;
  hipLaunchKernelGGL(k_trsm_blocks, dim3(nblk * 2 * st), dim3(CT), 0, s, V, SB, tinv, Dinv, DT, Gm);
  if (nblk > 1) hipLaunchKernelGGL(k_h_blocks, dim3(nblk * st * st), dim3(CT), 0, s, V, SB, nblk, DT, H);
}

template <typename TV>
static void solve_impl(const CoarseSolver& C, hipStream_t s, const long long* map, const TV* rdense, TV* xdense,
                       const double* Dinv, const double* DT, const double* Gm, const double* H, double* va,
                       double* vb, const int* fl) {
  const int wpb = CT / 64;
  const int n = C.n, np = C.np, SB = C.SB, nblk = C.nblk;
  if (C.one_launch()) {  // va, vb: np plain doubles, then 2 np tagged words each
    auto fwd = SB <= 4 * CT ? k_solve_chain<TV, 4, true> : k_solve_chain<TV, kChainMaxSB / CT, true>;
    auto bwd = SB <= 4 * CT ? k_solve_chain<TV, 4, false> : k_solve_chain<TV, kChainMaxSB / CT, false>;
    const dim3 grid((SB + wpb - 1) / wpb);
    hipLaunchKernelGGL(k_solve_f0<TV>, dim3((np + wpb - 1) / wpb), dim3(CT), 0, s, n, np, SB, map, rdense, Dinv, va,
                       fl);
    hipLaunchKernelGGL(fwd, grid, dim3(CT), SB * sizeof(double), s, n, SB, nblk, map, xdense, Gm, va,
                       (unsigned long long*)(va + np), C.chain, fl);
    hipLaunchKernelGGL(k_solve_b0, dim3((np + wpb - 1) / wpb), dim3(CT), 0, s, np, SB, DT, va, vb, fl);
    hipLaunchKernelGGL(bwd, grid, dim3(CT), SB * sizeof(double), s, n, SB, nblk, map, xdense, H, vb,
                       (unsigned long long*)(vb + np), C.chain, fl);
    return;
  }
  hipLaunchKernelGGL(k_solve_f0<TV>, dim3((np + wpb - 1) / wpb), dim3(CT), 0, s, n, np, SB, map, rdense, Dinv, va,
                     fl);
  for (int b = 1; b < nblk; ++b)
    hipLaunchKernelGGL(k_solve_step, dim3((SB + wpb - 1) / wpb), dim3(CT), 0, s, SB, b, -1, Gm, va, fl);
  hipLaunchKernelGGL(k_solve_b0, dim3((np + wpb - 1) / wpb), dim3(CT), 0, s, np, SB, DT, va, vb, fl);
  for (int b = nblk - 2; b >= 0; --b)
    hipLaunchKernelGGL(k_solve_step, dim3((SB + wpb - 1) / wpb), dim3(CT), 0, s, SB, b, +1, H, vb, fl);
  hipLaunchKernelGGL(k_solve_scatter<TV>, dim3((n + CT - 1) / CT), dim3(CT), 0, s, n, map, vb, xdense, fl);
}

void CoarseSolver::solve(hipStream_t s, const long long* map, const double* rdense, double* xdense, const double* Dinv,
                         const double* DT, const double* Gm, const double* H, double* va, double* vb,
                         const int* fl) const {
  solve_impl(*this, s, map, rdense, xdense, Dinv, DT, Gm, H, va, vb, fl);
}
void CoarseSolver::solve(hipStream_t s, const long long* map, const float* rdense, float* xdense, const double* Dinv,
                         const double* DT, const double* Gm, const double* H, double* va, double* vb,
                         const int* fl) const {
  solve_impl(*this, s, map, rdense, xdense, Dinv, DT, Gm, H, va, vb, fl);
}

// identity rows beyond the true order (after band assembly)
__global__ void k_band_pad(double* B, int n, int np, int W) {
  const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) B[(size_t)i * W] = 1.0;
}
void CoarseSolver::pad_identity(hipStream_t s, double* band) {
  if (np > n) hipLaunchKernelGGL(k_band_pad, dim3((np - n + 255) / 256), dim3(256), 0, s, band, n, np, bw + 1);
}

}  // namespace t3x
