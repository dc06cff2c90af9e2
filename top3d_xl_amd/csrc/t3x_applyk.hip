// Tiled matrix-free K*u for gfx950 (Solving_KbyU_MatrixFree, TOP3D_XL.m:641-726).
//
// Node-centric and z-marching: a work-group owns a 32x8 tile of node columns
// (y, x) and walks a chunk of node layers in z.  The masked displacements of
// the node layers (tile + 1-node halo, SoA, y fastest) sit in three LDS
// buffers: while cell layer k is computed from node layers k and k+1, the
// global loads of layer k+2 are in flight into registers and land in the
// third buffer afterwards, so one barrier per layer suffices.  The moduli of
// the 4 cells around a thread's column come straight from global memory
// (cache hits), one layer ahead.
// Each thread adds, for the 4 cells of layer k around its column, the rows
// of E_e*Ke belonging to its node in layer k (which then is complete and is
// written) and to its node in layer k+1 (carried to the next step), reading
// the cells' displacements from LDS as it goes: few registers (about 100),
// so five waves per SIMD hide the latency of the coefficient loads.
// Ke offsets are compile-time constants, so the coefficients are wave-uniform
// scalar loads; every FMA reads one scalar and one vector register.
//
// Boundary conditions as in the reference (703-715, 1338-1437): fixed DOFs are
// zeroed when staged (P_free u) and the fixed rows return u itself.
#include <hip/hip_runtime.h>

#include "t3x_kernels.h"
#include "t3x_reduce.h"

namespace t3x {

constexpr int AK_TY = 32, AK_TX = 8, AK_T = AK_TY * AK_TX;  // 256 threads
constexpr int AK_HY = AK_TY + 2, AK_HX = AK_TX + 2;         // halo tile
constexpr int AK_ZC = 32;                                   // node layers per work-group (minimum)
constexpr int AK_NS = (AK_HX * AK_HY + AK_T - 1) / AK_T;    // staged halo nodes per thread (2)

// Ke through a constant-address-space pointer: its entries are wave-uniform
// scalar loads.
template <typename T>
using KeP = const __attribute__((address_space(4))) T*;

// One cell's contribution: cell at column offset (cy, cx) in {0,1} relative to
// (y-1, x-1); the thread's node sits at local (dy, dx) = (1-cy, 1-cx) in it.
// s0 / s1: node layers k / k+1 in LDS as [component][x][y]; (hx, hy) the
// thread's column in the halo tile minus one.
template <typename T, int CY, int CX>
__device__ __forceinline__ void cell_pair(KeP<T> Ke, const T* __restrict__ s0, const T* __restrict__ s1, int hx,
                                          int hy, T Ee, T (&a0)[3], T (&a1)[3]) {
  constexpr int dy = 1 - CY, dx = 1 - CX;
  constexpr int jb = loc_index(dy, dx, 0);  // node in the bottom layer (k)
  constexpr int jt = loc_index(dy, dx, 1);  // node in the top layer (k+1)
  T acc0[3] = {0, 0, 0}, acc1[3] = {0, 0, 0};
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int ky = CY + kLocDy[j], kx = CX + kLocDx[j];
    const T* src = kLocDz[j] ? s1 : s0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v = src[(c * AK_HX + hx + kx) * AK_HY + hy + ky];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        acc0[r] += Ke[(3 * jb + r) * 24 + 3 * j + c] * v;
        acc1[r] += Ke[(3 * jt + r) * 24 + 3 * j + c] * v;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    a0[r] += acc0[r] * Ee;
    a1[r] += acc1[r] * Ee;
  }
}

// Epilogues at a completed node (Ku: the BC operator's row, u on fixed rows):
//   EPI_KU     y = Ku                         partials: u.y (the PCG's p.q)
//   EPI_RESID  y = r - Ku                     (the 'Standard' cycle's defect, 761)
//   EPI_SMOOTH y = u + omega (r - Ku) / d     partials: y.r (post-smoothing, 775)
enum { EPI_KU = 0, EPI_RESID = 1, EPI_SMOOTH = 2 };

// T: arithmetic and LDS type; TE, TU, TY, TR, TD: types of E, u, y, r, d in HBM.
// zc: node layers per work-group (a multiple of AK_ZC).
template <typename T, typename TE, typename TU, typename TY, int EPI, typename TR, typename TD>
__global__ __launch_bounds__(AK_T) void k_apply_K_tiled(Geo g, int zc, const T* __restrict__ Ke,
                                                        const TE* __restrict__ E, const uint8_t* __restrict__ fix,
                                                        const TU* __restrict__ u, TY* __restrict__ y,
                                                        double* __restrict__ partials, const int* fl, int reduce,
                                                        const TR* __restrict__ rv, const TD* __restrict__ dv,
                                                        double omega) {
  if (fl && fl[FL_DONE]) return;
  __shared__ T sU[3][3 * AK_HX * AK_HY];
  const int tid = threadIdx.x;
  const int ty = tid % AK_TY, tx = tid / AK_TY;
  const int y0 = blockIdx.x * AK_TY, x0 = blockIdx.y * AK_TX;
  const int zc0 = blockIdx.z * zc;
  const int zc1 = min(g.NZ, zc0 + zc);  // node layers written: [zc0, zc1)
  const int yy = y0 + ty, xx = x0 + tx;
  const bool own = yy < g.NY && xx < g.NX;
  const long long NN = g.NN;
  const long long plane = (long long)g.NY * g.NX;

  // global loads of node layer z (masked u) into registers; LDS stores later
  T pu[AK_NS][3];
  auto load_stage = [&](int z) {
#pragma unroll
    for (int s2 = 0; s2 < AK_NS; ++s2) {
      const int t = tid + s2 * AK_T;
      const int hy = t % AK_HY, hx = t / AK_HY;
      const int gy = y0 - 1 + hy, gx = x0 - 1 + hx;
      pu[s2][0] = pu[s2][1] = pu[s2][2] = T(0);
      if (t < AK_HX * AK_HY && z >= 0 && z < g.NZ && gy >= 0 && gx >= 0 && gy < g.NY && gx < g.NX) {
        const long long n = gy + (long long)g.NY * gx + plane * z;
        const uint8_t f = fix[n];
        pu[s2][0] = (f & 1) ? T(0) : T(u[n]);
        pu[s2][1] = (f & 2) ? T(0) : T(u[NN + n]);
        pu[s2][2] = (f & 4) ? T(0) : T(u[2 * NN + n]);
      }
    }
  };
  // LDS stores of node layer z (from registers)
  auto store_stage = [&](int b) {
#pragma unroll
    for (int s2 = 0; s2 < AK_NS; ++s2) {
      const int t = tid + s2 * AK_T;
      if (t < AK_HX * AK_HY) {
        const int hy = t % AK_HY, hx = t / AK_HY;
#pragma unroll
        for (int c = 0; c < 3; ++c) sU[b][(c * AK_HX + hx) * AK_HY + hy] = pu[s2][c];
      }
    }
  };
  // moduli of the 4 cells of layer z around this column: (y-1+CY, x-1+CX)
  T ep[4];
  auto load_moduli = [&](int z) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int cy = yy - 1 + (c & 1), cx = xx - 1 + (c >> 1);
      ep[c] = (z >= 0 && z < g.nz && cy >= 0 && cx >= 0 && cy < g.ny && cx < g.nx)
                  ? T(E[cy + (long long)g.NY * cx + plane * z])
                  : T(0);
    }
  };
  auto buf = [](int z) { return (z + 3) % 3; };  // z >= -1

  T a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0};
  double pq = 0.0;
  // cell layers k = zc0-1 .. zc1-1 from node layers k (buffer buf(k)) and k+1
  const int kBeg = zc0 - 1;
  load_moduli(kBeg);
  load_stage(kBeg);
  store_stage(buf(kBeg));
  load_stage(kBeg + 1);
  store_stage(buf(kBeg + 1));
  __syncthreads();
  for (int k = kBeg; k < zc1; ++k) {
    const bool more = k + 2 <= zc1;
    if (more) load_stage(k + 2);  // in flight during this layer's FMAs
    const T* s0 = sU[buf(k)];
    const T* s1 = sU[buf(k + 1)];
    const T e00 = ep[0], e10 = ep[1], e01 = ep[2], e11 = ep[3];
    load_moduli(k + 1);  // in flight as well
    // cells (y-1+CY, x-1+CX) of layer k.  The Ke pointer is made opaque per
    // layer so its 576 entries are loaded where they are used instead of
    // being hoisted out of the loop (1152 SGPRs, spilled to VGPR lanes).
    KeP<T> Kq = (KeP<T>)Ke;
    asm volatile("" : "+s"(Kq));
    if (e00 != T(0)) cell_pair<T, 0, 0>(Kq, s0, s1, tx, ty, e00, a0, a1);
    if (e10 != T(0)) cell_pair<T, 1, 0>(Kq, s0, s1, tx, ty, e10, a0, a1);
    if (e01 != T(0)) cell_pair<T, 0, 1>(Kq, s0, s1, tx, ty, e01, a0, a1);
    if (e11 != T(0)) cell_pair<T, 1, 1>(Kq, s0, s1, tx, ty, e11, a0, a1);
    // the third buffer held node layer k-1, which no thread reads any more
    // (every thread passed the previous layer's barrier)
    if (more) store_stage(buf(k + 2));
    if (k >= zc0 && own) {  // node layer k complete
      const long long n = yy + (long long)g.NY * xx + plane * k;
      const uint8_t f = fix[n];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double uc = u[c * NN + n];
        const T ku = ((f >> c) & 1) ? T(uc) : a0[c];
        if (EPI == EPI_KU) {
          const TY v = TY(ku);
          y[c * NN + n] = v;
          pq += uc * double(v);
        } else if (EPI == EPI_RESID) {
          y[c * NN + n] = TY(T(rv[c * NN + n]) - ku);
        } else {
          const T rc = T(rv[c * NN + n]), dc = T(dv[c * NN + n]);
          const TY v = dc != T(0) ? TY(T(uc) + T(omega) * (rc - ku) / dc) : TY(0);
          y[c * NN + n] = v;
          pq += double(v) * double(rv[c * NN + n]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      a0[r] = a1[r];
      a1[r] = 0;
    }
    __syncthreads();  // layer k+2 staged; buffer buf(k) free for layer k+3
  }
  if (reduce) {
    // one partial per work-group, flattened block index
    __shared__ double shp[AK_T / 64];
    double v = wave_sum(pq);
    if ((tid & 63) == 0) shp[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < AK_T / 64; ++w) s += shp[w];
      partials[blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z)] = s;
    }
  }
}

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
// cycle).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cell_Ks_u(Geo g, const T* __restrict__ Ks, const T* __restrict__ u,
                                                      T* __restrict__ Yc, const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int plane = g.NY * g.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= plane) return;
  const int cx = idx / g.NY, cy = idx - cx * g.NY;
  const int cz = blockIdx.y;
  if (cy >= g.ny || cx >= g.nx || cz >= g.nz) return;
  const long long NN = g.NN;
  const long long cell = idx + (long long)plane * cz;
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
      const T v = Ks[(long long)(q * (q + 1) / 2 + p) * NN + cell];
      ye[p] += v * ue[q];
      if (p != q) ye[q] += v * ue[p];
    }
  }
#pragma unroll
  for (int i = 0; i < 24; ++i) Yc[(long long)i * NN + cell] = ye[i];
}

template <typename T, int EPI>
__global__ __launch_bounds__(kBlock) void k_node_gather_Ks(Geo g, const T* __restrict__ Yc, const T* __restrict__ u,
                                                           T* __restrict__ y, const T* __restrict__ rv,
                                                           const T* __restrict__ dv, double omega, const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int plane = g.NY * g.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= plane) return;
  const int x = idx / g.NY, yy = idx - x * g.NY;
  const int z = blockIdx.y;
  const long long NN = g.NN;
  const long long n = idx + (long long)plane * z;
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

// z-chunk of the tiled kernel: AK_ZC node layers, doubled until the grid fits
// the partials buffer (one partial per work-group)
static int applyk_zc(const Geo& g) {
  const long long bxy = (long long)((g.NY + AK_TY - 1) / AK_TY) * ((g.NX + AK_TX - 1) / AK_TX);
  int zc = AK_ZC;
  while (bxy * ((g.NZ + zc - 1) / zc) > kMaxPartials) zc *= 2;
  return zc;
}

dim3 applyk_grid(const Geo& g) {
  const int zc = applyk_zc(g);
  return dim3((g.NY + AK_TY - 1) / AK_TY, (g.NX + AK_TX - 1) / AK_TX, (g.NZ + zc - 1) / zc);
}

int applyk_tiled_blocks(const Geo& g) {
  const dim3 d = applyk_grid(g);
  return (int)(d.x * d.y * d.z);
}

void launch_apply_K_tiled(hipStream_t s, const Geo& g, const double* Ke, const double* E, const uint8_t* fix,
                          const double* u, double* y, double* partials, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_tiled<double, double, double, double, EPI_KU, double, double>), applyk_grid(g),
                     dim3(AK_T), 0, s, g, applyk_zc(g), Ke, E, fix, u, y, partials, fl, partials ? 1 : 0,
                     (const double*)nullptr, (const double*)nullptr, 0.0);
}

static dim3 node_grid(const Geo& g) { return dim3((g.NY * g.NX + kBlock - 1) / kBlock, g.NZ); }

long long coarse_ks_tmp_len(const Geo& g) { return 24LL * g.NN; }

template <typename T>
static void coarse_resid(hipStream_t s, const Geo& g, const T* Ks, const T* x, const T* r, T* d, T* tmp,
                         const int* fl) {
  hipLaunchKernelGGL(k_cell_Ks_u<T>, node_grid(g), dim3(kBlock), 0, s, g, Ks, x, tmp, fl);
  hipLaunchKernelGGL((k_node_gather_Ks<T, EPI_RESID>), node_grid(g), dim3(kBlock), 0, s, g, (const T*)tmp, x, d, r,
                     (const T*)nullptr, 0.0, fl);
}

template <typename T>
static void coarse_smooth(hipStream_t s, const Geo& g, const T* Ks, const T* x, const T* r, const T* diag,
                          double omega, T* out, T* tmp, const int* fl) {
  hipLaunchKernelGGL(k_cell_Ks_u<T>, node_grid(g), dim3(kBlock), 0, s, g, Ks, x, tmp, fl);
  hipLaunchKernelGGL((k_node_gather_Ks<T, EPI_SMOOTH>), node_grid(g), dim3(kBlock), 0, s, g, (const T*)tmp, x, out,
                     r, diag, omega, fl);
}

void launch_coarse_resid32(hipStream_t s, const Geo& g, const float* Ks, const float* x, const float* r, float* d,
                           float* tmp, const int* fl) {
  coarse_resid(s, g, Ks, x, r, d, tmp, fl);
}
void launch_coarse_smooth32(hipStream_t s, const Geo& g, const float* Ks, const float* x, const float* r,
                            const float* diag, double omega, float* out, float* tmp, const int* fl) {
  coarse_smooth(s, g, Ks, x, r, diag, omega, out, tmp, fl);
}
void launch_coarse_resid64(hipStream_t s, const Geo& g, const double* Ks, const double* x, const double* r,
                           double* d, double* tmp, const int* fl) {
  coarse_resid(s, g, Ks, x, r, d, tmp, fl);
}
void launch_coarse_smooth64(hipStream_t s, const Geo& g, const double* Ks, const double* x, const double* r,
                            const double* diag, double omega, double* out, double* tmp, const int* fl) {
  coarse_smooth(s, g, Ks, x, r, diag, omega, out, tmp, fl);
}

// FP64 fine-level smoother products of the 'Standard' cycle (761, 775) with
// E*Ke and the boundary-condition rows, through the tiled K*u.
void launch_fine_resid64(hipStream_t s, const Geo& g, const double* Ke, const double* E, const uint8_t* fix,
                         const double* x, const double* r, double* d, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_tiled<double, double, double, double, EPI_RESID, double, double>), applyk_grid(g),
                     dim3(AK_T), 0, s, g, applyk_zc(g), Ke, E, fix, x, d, (double*)nullptr, fl, 0, r,
                     (const double*)nullptr, 0.0);
}
void launch_fine_smooth64(hipStream_t s, const Geo& g, const double* Ke, const double* E, const uint8_t* fix,
                          const double* x, const double* r, const double* diag, double omega, double* out,
                          double* partials, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_tiled<double, double, double, double, EPI_SMOOTH, double, double>), applyk_grid(g),
                     dim3(AK_T), 0, s, g, applyk_zc(g), Ke, E, fix, x, out, partials, fl, partials ? 1 : 0, r, diag,
                     omega);
}

}  // namespace t3x
