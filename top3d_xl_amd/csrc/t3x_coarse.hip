// Coarsest-level direct solver of libt3x (Solving_AssembleFEAstencil chol,
// TOP3D_XL.m:1052-1065, and the triangular solves of Solving_Vcycle, 767).
//
// The coarsest Galerkin matrix is banded after an axis-permuted node order
// (bandwidth bw).  The reference factors it with CHOLMOD; here:
//
//  1. right-looking blocked banded Cholesky, 64-column panels, two launches
//     per panel (TRSM of the panel, SYRK of the trailing band).  The work-
//     group that finishes the update of the next diagonal tile factors it and
//     stores the tile's inverse, so no serial phase sits between launches.
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
// In-place Cholesky of the lower triangle of sA, blocked in 4 panels of 16
// columns: the block's first wave factors a panel in registers (lane i owns
// row i; the pivot and the column entries travel by cross-lane shuffles, no
// LDS round-trip per element), then all threads apply the panel's rank-16
// update to the trailing lower triangle.
__device__ void tile_potrf(double (*sA)[TB + 1], int* status) {
  constexpr int PB = 16;
  for (int b = 0; b < TB / PB; ++b) {
    const int b0 = b * PB;
    if (threadIdx.x < TB) {
      const int i = threadIdx.x;
      double p[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) p[k] = sA[i][b0 + k];
      bool bad = false;
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        const int c = b0 + k;
        const double piv = __shfl(p[k], c, 64);
        const double d = sqrt(piv);
        bad |= !(piv > 0.0);
        if (i == c) p[k] = d;
        else if (i > c) p[k] = p[k] / d;
#pragma unroll
        for (int kk = k + 1; kk < PB; ++kk) {
          const double ljc = __shfl(p[k], b0 + kk, 64);  // L(b0 + kk, c)
          if (i >= b0 + kk) p[kk] -= p[k] * ljc;
        }
      }
      if (i >= b0) {
#pragma unroll
        for (int k = 0; k < PB; ++k)
          if (b0 + k <= i) sA[i][b0 + k] = p[k];
      }
      if (bad && i == b0) *status = 1;
    }
    __syncthreads();
    // trailing update: rows and columns >= b0 + PB, lower triangle
    const int t0 = b0 + PB, m = TB - t0;
    for (int t = threadIdx.x; t < m * m; t += CT) {
      const int i = t0 + t / m, j = t0 + t % m;
      if (j > i) continue;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < PB; ++k) acc += sA[i][b0 + k] * sA[j][b0 + k];
      sA[i][j] -= acc;
    }
    __syncthreads();
  }
}

// Inverse of the lower-triangular sA into sI (one column per thread).
__device__ void tile_trinv(const double (*sA)[TB + 1], double (*sI)[TB + 1]) {
  if (threadIdx.x < TB) {
    const int c = threadIdx.x;
    for (int i = 0; i < c; ++i) sI[i][c] = 0.0;
    sI[c][c] = 1.0 / sA[c][c];
    for (int i = c + 1; i < TB; ++i) {
      double s = 0.0;
      for (int k = c; k < i; ++k) s += sA[i][k] * sI[k][c];
      sI[i][c] = -s / sA[i][i];
    }
  }
  __syncthreads();
}

// acc[4][4] += A(4ti.., :) * B(:, 4tj..) for thread (ti, tj) = (t/16, t%16);
// A given row-major sA[m][k], B given as sB[k][n] (transB = false) or
// sB[n][k] (transB = true).
template <bool transB>
__device__ __forceinline__ void tile_gemm_acc(const double (*sA)[TB + 1], const double (*sB)[TB + 1],
                                              double (&acc)[4][4]) {
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll 4
  for (int k = 0; k < TB; ++k) {
    double a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] = sA[4 * ti + r][k];
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = transB ? sB[4 * tj + q][k] : sB[k][4 * tj + q];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[r][q] += a[r] * b[q];
  }
}

// ---- 1. banded Cholesky ---------------------------------------------------
// Factor diagonal tile d0 (global row/col) in place.
__device__ void diag_tile(BandView V, int d0, double (*sA)[TB + 1], int* status) {
  for (int t = threadIdx.x; t < TB * TB; t += CT) {
    const int i = t / TB, k = t % TB;
    sA[i][k] = V.get(d0 + i, d0 + k);
  }
  __syncthreads();
  tile_potrf(sA, status);
  for (int t = threadIdx.x; t < TB * TB; t += CT) {
    const int i = t / TB, k = t % TB;
    if (k <= i) V.set(d0 + i, d0 + k, sA[i][k]);
  }
}

__global__ __launch_bounds__(CT) void k_chol_diag(BandView V, int d0, int* status) {
  __shared__ double sA[TB][TB + 1];
  diag_tile(V, d0, sA, status);
}

// L(i, J) = A(i, J) * L_JJ^{-T} for the row tiles below diagonal tile J, by
// forward substitution in 4 column blocks of 16: lane i solves its row's
// block in registers (right-looking: independent FMAs after each division;
// entries of L_JJ are uniform LDS broadcasts), then updates the row's later
// columns in LDS with the block (a rank-16 update).
__global__ __launch_bounds__(CT) void k_chol_trsm(BandView V, int J) {
  constexpr int PB = 16;
  __shared__ double sL[TB][TB + 1];
  __shared__ double sX[TB][TB + 1];
  const int j0 = J * TB, r0 = j0 + TB;
  const int i0 = r0 + blockIdx.x * TB;
  for (int t = threadIdx.x; t < TB * TB; t += CT) {
    const int i = t / TB, c = t % TB;
    sL[i][c] = V.get(j0 + i, j0 + c);
    sX[i][c] = V.get(i0 + i, j0 + c);
  }
  __syncthreads();
  if (threadIdx.x < TB) {
    const int i = threadIdx.x;
    for (int b = 0; b < TB / PB; ++b) {
      const int b0 = b * PB;
      double x[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) x[k] = sX[i][b0 + k];
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        x[k] = x[k] / sL[b0 + k][b0 + k];
#pragma unroll
        for (int kk = k + 1; kk < PB; ++kk) x[kk] -= x[k] * sL[b0 + kk][b0 + k];
      }
#pragma unroll
      for (int k = 0; k < PB; ++k) sX[i][b0 + k] = x[k];
      for (int j = b0 + PB; j < TB; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < PB; ++k) acc += x[k] * sL[j][b0 + k];
        sX[i][j] -= acc;
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < TB * TB; t += CT) V.set(i0 + t / TB, j0 + t % TB, sX[t / TB][t % TB]);
}

// Trailing update of the band by panel J; tile 0 is the next diagonal tile,
// which its work-group factors right after updating it.
__global__ __launch_bounds__(CT) void k_chol_syrk(BandView V, int J, int* status) {
  __shared__ double sA[TB][TB + 1];
  __shared__ double sB[TB][TB + 1];
  const int bw = V.bw, n = V.n;
  const int j0 = J * TB, r0 = j0 + TB;
  const int t0 = blockIdx.x;
  int I = 0;
  while ((I + 1) * (I + 2) / 2 <= t0) ++I;
  const int K = t0 - I * (I + 1) / 2;
  const int i0 = r0 + I * TB, k0 = r0 + K * TB;
  if (i0 - (k0 + TB - 1) > bw) return;  // tile outside the band
  for (int t = threadIdx.x; t < TB * TB; t += CT) {
    const int a = t / TB, c = t % TB;
    sA[a][c] = V.get(i0 + a, j0 + c);
    sB[a][c] = V.get(k0 + a, j0 + c);
  }
  __syncthreads();
  double acc[4][4] = {};
  tile_gemm_acc<true>(sA, sB, acc);  // A(i, J) * A(k, J)^T
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 4 * ti + r, k = k0 + 4 * tj + q;
      if (k <= i && i - k <= bw && i < n) V.B[(size_t)i * V.W() + (i - k)] -= acc[r][q];
    }
  if (t0 == 0) {
    __syncthreads();
    diag_tile(V, r0, sA, status);
  }
}

// Inverses of all 64x64 diagonal tiles (after the factorization, in parallel).
__global__ __launch_bounds__(CT) void k_tile_inverses(BandView V, double* __restrict__ tinv) {
  __shared__ double sA[TB][TB + 1];
  __shared__ double sI[TB][TB + 1];
  const int d0 = blockIdx.x * TB;
  for (int t = threadIdx.x; t < TB * TB; t += CT) sA[t / TB][t % TB] = V.get(d0 + t / TB, d0 + t % TB);
  __syncthreads();
  tile_trinv(sA, sI);
  double* inv = tinv + (size_t)blockIdx.x * TB * TB;
  for (int t = threadIdx.x; t < TB * TB; t += CT) inv[t] = sI[t / TB][t % TB];
}

// ---- 2. solve-block operators -------------------------------------------
// WG (b, p): columns p*64.. of X = L_bb^{-1} [I | L_{b,b-1}], SB x 2SB.
// Output: Dinv_b (SB x SB row-major) and its transpose DT_b, G_b (SB x SB).
__global__ __launch_bounds__(CT) void k_trsm_blocks(BandView V, int SB, const double* __restrict__ tinv,
                                                    double* __restrict__ Dinv, double* __restrict__ DT,
                                                    double* __restrict__ Gm) {
  __shared__ double sL[TB][TB + 1];
  __shared__ double sX[TB][TB + 1];
  __shared__ double sT[TB][TB + 1];
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
    for (int t = threadIdx.x; t < TB * TB; t += CT) {
      const int a = t / TB, c = t % TB;
      double r;
      if (ident) r = (I * TB + a == col0 + c) ? 1.0 : 0.0;
      else r = V.get(base + I * TB + a, base - SB + col0 + c);
      sT[a][c] = r;
    }
    __syncthreads();
    double acc[4][4] = {};
    for (int K = Ist; K < I; ++K) {
      if ((I * TB) - (K * TB + TB - 1) > V.bw) continue;  // L tile outside the band
      for (int t = threadIdx.x; t < TB * TB; t += CT) {
        const int a = t / TB, c = t % TB;
        sL[a][c] = V.get(base + I * TB + a, base + K * TB + c);
        sX[a][c] = Out[(size_t)(K * TB + a) * SB + col0 + c];
      }
      __syncthreads();
      tile_gemm_acc<false>(sL, sX, acc);
      __syncthreads();
    }
    {
      const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) sT[4 * ti + r][4 * tj + q] -= acc[r][q];
    }
    __syncthreads();
    // X_I = inv(L_II) T
    const double* inv = tinv + (size_t)(base / TB + I) * TB * TB;
    for (int t = threadIdx.x; t < TB * TB; t += CT) sL[t / TB][t % TB] = inv[t];
    __syncthreads();
    {
      double acc2[4][4] = {};
      tile_gemm_acc<false>(sL, sT, acc2);  // inv(L_II) is lower triangular: the zeros add nothing
      const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int a = 4 * ti + r, c = 4 * tj + q;
          Out[(size_t)(I * TB + a) * SB + col0 + c] = acc2[r][q];
          if (ident) DT[(size_t)b * SB * SB + (size_t)(col0 + c) * SB + I * TB + a] = acc2[r][q];
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
  const int b = blockIdx.x / per, rem = blockIdx.x % per;
  const int I = rem / st, J = rem % st;
  if (b >= nblk - 1) return;
  const double* D = DT + (size_t)b * SB * SB;
  double acc[4][4] = {};
  for (int K = I; K < st; ++K) {  // DT_b(i,k) = Dinv_b(k,i) is zero for k < i
    const int rowL0 = (b + 1) * SB + J * TB, colL0 = b * SB + K * TB;
    if (rowL0 - (colL0 + TB - 1) > V.bw) continue;
    for (int t = threadIdx.x; t < TB * TB; t += CT) {
      const int a = t / TB, c = t % TB;
      sA[a][c] = D[(size_t)(I * TB + a) * SB + K * TB + c];
      sB[a][c] = V.get(rowL0 + a, colL0 + c);  // row j, col k
    }
    __syncthreads();
    tile_gemm_acc<true>(sA, sB, acc);
    __syncthreads();
  }
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q)
      H[(size_t)b * SB * SB + (size_t)(I * TB + 4 * ti + r) * SB + J * TB + 4 * tj + q] = acc[r][q];
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

// ---- host side -----------------------------------------------------------
void CoarseSolver::plan(int n_, int bw_) {
  n = n_;
  bw = bw_;
  SB = ((bw + 1 + TB - 1) / TB) * TB;
  if (SB < TB) SB = TB;
  np = ((n + SB - 1) / SB) * SB;
  nblk = np / SB;
}

size_t CoarseSolver::band_len() const { return (size_t)np * (bw + 1); }

void CoarseSolver::factor(hipStream_t s, double* band, double* tinv, double* Dinv, double* DT, double* Gm,
                          double* H, int* status) {
  BandView V{band, np, bw};
  const int nt = np / TB;
  hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(CT), 0, s, V, 0, status);
  for (int J = 0; J < nt; ++J) {
    const int j0 = J * TB, r0 = j0 + TB;
    const int rlast = std::min(np - 1, j0 + TB - 1 + bw);
    const int nrt = rlast >= r0 ? (rlast - r0) / TB + 1 : 0;
    if (nrt > 0) {
      hipLaunchKernelGGL(k_chol_trsm, dim3(nrt), dim3(CT), 0, s, V, J);
      hipLaunchKernelGGL(k_chol_syrk, dim3(nrt * (nrt + 1) / 2), dim3(CT), 0, s, V, J, status);
    } else if (J + 1 < nt) {
      hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(CT), 0, s, V, r0, status);
    }
  }
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
