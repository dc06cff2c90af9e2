// Coarse-level K*u (TOP3D_XL.m:727-744).  The fine level's K*u, with the
// element matrix in sum-factorised form, is in t3x_applyk_sf.hip.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "t3x_kernels.h"
#include "t3x_reduce.h"

namespace t3x {

// Epilogues at a completed node (Ku: the BC operator's row, u on fixed rows):
//   EPI_RESID  y = r - Ku                     (the 'Standard' cycle's defect, 761)
//   EPI_SMOOTH y = u + omega (r - Ku) / d     (post-smoothing, 775)
enum { EPI_RESID = 1, EPI_SMOOTH = 2 };

// Coarse-level K*u with the stored Galerkin cell matrices (Solving_KbyU_MatrixFree
// for iLevel > 1, 727-744), in two passes so that every stored entry is read
// exactly once with fully coalesced loads:
//   1. one thread per cell forms Y_c = Ks_c u_c (24 values, upper triangle
//      entry-major: a wave reads one entry of 64 consecutive cells) and
//      writes Y_c entry-major to a scratch array;
//   2. one thread per node adds the 3 rows of Y_c that belong to it from
//      its <= 8 cells in a fixed order and applies the smoother epilogue.
// The scratch (24 floats per node slot) costs 8% of the matrix bytes.
// Coarse levels carry no fixed DOFs (the boundary conditions are inside the
// Galerkin matrices); vectors and matrices are zero on void nodes and cells.
// T = float (the mixed form's FP32 smoother) or double (the FP64 'Standard'
// cycle); TK, the matrices' storage: T, or FP16 with a per-cell power-of-two
// scale (the products formed unscaled, the cell's outputs scaled once,
// exactly).
// (The FP16 instance holds a cell's loads in 226 VGPRs, two waves per SIMD;
// bounded to four waves, 80 VGPRs, it ran 0.30 against 0.19 ms per V-cycle.)
template <typename T, typename TK = T>
__global__ __launch_bounds__(kBlock) void k_cell_Ks_u(Geo g, const TK* __restrict__ Ks, const T* __restrict__ u,
                                                      T* __restrict__ Yc, const int* fl,
                                                      const uint8_t* __restrict__ emask,
                                                      const float* __restrict__ scale = nullptr) {
  if (fl && fl[FL_DONE]) return;
  const int plane = g.NY * g.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= plane) return;
  const int cx = idx / g.NY, cy = idx - cx * g.NY;
  const int cz = blockIdx.y;
  if (cy >= g.ny || cx >= g.nx || cz >= g.nz) return;
  const long long NN = g.NN;
  const long long cell = idx + (long long)plane * cz;
  if (!emask[cell]) {  // void cell: zero product without reading its (zero) matrix
#pragma unroll
    for (int i = 0; i < 24; ++i) Yc[(long long)i * NN + cell] = T(0);
    return;
  }
  T ue[24], ye[24];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long m = cell + kLocDy[a] + (long long)g.NY * kLocDx[a] + (long long)plane * kLocDz[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ue[3 * a + c] = u[c * NN + m];
      ye[3 * a + c] = T(0);
    }
  }
#pragma unroll
  for (int q = 0; q < 24; ++q) {
#pragma unroll
    for (int p = 0; p <= q; ++p) {
      const T v = T(Ks[(long long)(q * (q + 1) / 2 + p) * NN + cell]);
      ye[p] += v * ue[q];
      if (p != q) ye[q] += v * ue[p];
    }
  }
  if constexpr (std::is_same<TK, T>::value) {
#pragma unroll
    for (int i = 0; i < 24; ++i) Yc[(long long)i * NN + cell] = ye[i];
  } else {
    const T sc = T(scale[cell]);
#pragma unroll
    for (int i = 0; i < 24; ++i) Yc[(long long)i * NN + cell] = ye[i] * sc;
  }
}

// Ks -> FP16 with a per-cell power-of-two scale (one thread per cell)
__global__ __launch_bounds__(kBlock) void k_scale_ks_f16(Geo g, const double* __restrict__ Ks,
                                                         __half* __restrict__ Ksh, float* __restrict__ scale) {
  const long long NN = g.NN;
  const long long cell = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (cell >= NN) return;
  double m = 0.0;
  for (int e = 0; e < kTri; ++e) m = fmax(m, fabs(Ks[(long long)e * NN + cell]));
  int ex = 0;
  if (m > 0.0) (void)frexp(m, &ex);  // m < 2^ex
  const double s = m > 0.0 ? ldexp(1.0, ex) : 1.0;
  scale[cell] = (float)s;
  for (int e = 0; e < kTri; ++e) Ksh[(long long)e * NN + cell] = __float2half((float)(Ks[(long long)e * NN + cell] / s));
}

template <typename T, int EPI>
__global__ __launch_bounds__(kBlock) void k_node_gather_Ks(Geo g, const T* __restrict__ Yc, const T* __restrict__ u,
                                                           T* __restrict__ y, const T* __restrict__ rv,
                                                           const T* __restrict__ dv, double omega, const int* fl,
                                                           const uint8_t* __restrict__ nmask) {
  if (fl && fl[FL_DONE]) return;
  const int plane = g.NY * g.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= plane) return;
  const int x = idx / g.NY, yy = idx - x * g.NY;
  const int z = blockIdx.y;
  const long long NN = g.NN;
  const long long n = idx + (long long)plane * z;
  if (!nmask[n]) {  // void node: r, u and the result are 0
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c * NN + n] = T(0);
    return;
  }
  T acc[3] = {0, 0, 0};
#pragma unroll
  for (int cc = 0; cc < 8; ++cc) {
    const int CY = cc & 1, CX = (cc >> 1) & 1, CZ = cc >> 2;
    const int cy = yy - 1 + CY, cx = x - 1 + CX, cz = z - 1 + CZ;
    if (cy < 0 || cx < 0 || cz < 0 || cy >= g.ny || cx >= g.nx || cz >= g.nz) continue;
    const long long cell = cy + (long long)g.NY * cx + (long long)plane * cz;
    const int jl = loc_index(1 - CY, 1 - CX, 1 - CZ);  // this node's local index in the cell
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] += Yc[(long long)(3 * jl + r) * NN + cell];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (EPI == EPI_RESID) {
      y[c * NN + n] = rv[c * NN + n] - acc[c];
    } else {
      const T dc = dv[c * NN + n];
      y[c * NN + n] = dc != T(0) ? u[c * NN + n] + T(omega) * (rv[c * NN + n] - acc[c]) / dc : T(0);
    }
  }
}

static dim3 node_grid(const Geo& g) { return dim3((g.NY * g.NX + kBlock - 1) / kBlock, g.NZ); }

long long coarse_ks_tmp_len(const Geo& g) { return 24LL * g.NN; }

template <typename T, typename TK>
static void coarse_resid(hipStream_t s, const Geo& g, const TK* Ks, const T* x, const T* r, T* d, T* tmp,
                         const int* fl, const uint8_t* emask, const uint8_t* nmask, const float* scale = nullptr) {
  hipLaunchKernelGGL((k_cell_Ks_u<T, TK>), node_grid(g), dim3(kBlock), 0, s, g, Ks, x, tmp, fl, emask, scale);
  hipLaunchKernelGGL((k_node_gather_Ks<T, EPI_RESID>), node_grid(g), dim3(kBlock), 0, s, g, (const T*)tmp, x, d, r,
                     (const T*)nullptr, 0.0, fl, nmask);
}

template <typename T, typename TK>
static void coarse_smooth(hipStream_t s, const Geo& g, const TK* Ks, const T* x, const T* r, const T* diag,
                          double omega, T* out, T* tmp, const int* fl, const uint8_t* emask, const uint8_t* nmask,
                          const float* scale = nullptr) {
  hipLaunchKernelGGL((k_cell_Ks_u<T, TK>), node_grid(g), dim3(kBlock), 0, s, g, Ks, x, tmp, fl, emask, scale);
  hipLaunchKernelGGL((k_node_gather_Ks<T, EPI_SMOOTH>), node_grid(g), dim3(kBlock), 0, s, g, (const T*)tmp, x, out,
                     r, diag, omega, fl, nmask);
}

void launch_coarse_resid32(hipStream_t s, const Geo& g, const float* Ks, const float* x, const float* r, float* d,
                           float* tmp, const int* fl, const uint8_t* emask, const uint8_t* nmask) {
  coarse_resid(s, g, Ks, x, r, d, tmp, fl, emask, nmask);
}
void launch_coarse_smooth32(hipStream_t s, const Geo& g, const float* Ks, const float* x, const float* r,
                            const float* diag, double omega, float* out, float* tmp, const int* fl,
                            const uint8_t* emask, const uint8_t* nmask) {
  coarse_smooth(s, g, Ks, x, r, diag, omega, out, tmp, fl, emask, nmask);
}
void launch_coarse_resid32(hipStream_t s, const Geo& g, const __half* Ks, const float* scale, const float* x,
                           const float* r, float* d, float* tmp, const int* fl, const uint8_t* emask,
                           const uint8_t* nmask) {
  coarse_resid(s, g, Ks, x, r, d, tmp, fl, emask, nmask, scale);
}
void launch_coarse_smooth32(hipStream_t s, const Geo& g, const __half* Ks, const float* scale, const float* x,
                            const float* r, const float* diag, double omega, float* out, float* tmp, const int* fl,
                            const uint8_t* emask, const uint8_t* nmask) {
  coarse_smooth(s, g, Ks, x, r, diag, omega, out, tmp, fl, emask, nmask, scale);
}
void launch_scale_ks_f16(hipStream_t s, const Geo& g, const double* Ks, __half* Ksh, float* scale) {
  hipLaunchKernelGGL(k_scale_ks_f16, dim3((unsigned)((g.NN + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, g, Ks, Ksh,
                     scale);
}
void launch_coarse_resid64(hipStream_t s, const Geo& g, const double* Ks, const double* x, const double* r,
                           double* d, double* tmp, const int* fl, const uint8_t* emask, const uint8_t* nmask) {
  coarse_resid(s, g, Ks, x, r, d, tmp, fl, emask, nmask);
}
void launch_coarse_smooth64(hipStream_t s, const Geo& g, const double* Ks, const double* x, const double* r,
                            const double* diag, double omega, double* out, double* tmp, const int* fl,
                            const uint8_t* emask, const uint8_t* nmask) {
  coarse_smooth(s, g, Ks, x, r, diag, omega, out, tmp, fl, emask, nmask);
}

}  // namespace t3x
