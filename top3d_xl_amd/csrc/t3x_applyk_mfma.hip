// Fine-level K*u on the matrix cores (Solving_KbyU_MatrixFree, TOP3D_XL.m:641-726),
// used by the FP32 'Standard' V-cycle smoother.
//
// Per cell layer the product the reference forms as U_e * Ke (714) is done as
// Y = Ke * U on v_mfma_*_16x16x4: A = Ke (24 rows padded to 32, 6 k-blocks of
// 4 columns: 12 register-resident fragments per lane, loaded once), B = the
// 24 displacements of 16 cells along y gathered from the LDS-staged node
// layers, C = 24 rows x 16 cells.  Columns are scaled by E_e, written to LDS,
// and each node thread then adds the rows of its node from the 4 cells
// around its column in a fixed order (deterministic, no atomics).
//
// Tiling: a work-group owns 15x15 (y, x) node columns and the 16x16 cells
// around them, and marches 32 node layers in z (node layers k and k+1 and
// the cell moduli of layer k double-buffered in LDS, as in the vector
// kernel); wave w computes the cells of x-rows 4w..4w+3.  Boundary
// conditions: fixed DOFs are zeroed when staged (P_free u) and fixed rows
// return u (703-715, 1338-1437).
#include <hip/hip_runtime.h>

#include "t3x_kernels.h"
#include "t3x_reduce.h"

namespace t3x {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// v_mfma_f{64,32}_16x16x4 operand layouts (gfx950):
//   A[i][k] at lane i + 16k, B[k][j] at lane j + 16k;
//   D[i][j]: f64 lane j + 16 (i % 4), register i / 4;
//            f32 lane j + 16 (i / 4), register i % 4.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 V;
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef f4 V;
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int MK_NT = 15;            // node columns per tile edge
constexpr int MK_CT = MK_NT + 1;     // cells per tile edge (16: one MFMA column group)
constexpr int MK_HT = MK_NT + 2;     // staged nodes per tile edge (17)
constexpr int MK_ZC = 32;            // node layers per work-group
constexpr int MK_T = 256;

}  // namespace

enum { MEPI_KU = 0, MEPI_RESID = 1, MEPI_SMOOTH = 2 };

// T: arithmetic/LDS type; TE, TU, TY, TR, TD: types of E, u, y, r, diag in HBM.
template <typename T, typename TE, typename TU, typename TY, int EPI, typename TR, typename TD>
__global__ __launch_bounds__(MK_T) void k_apply_K_mfma(Geo g, const T* __restrict__ Ke, const TE* __restrict__ E,
                                                       const uint8_t* __restrict__ fix, const TU* __restrict__ u,
                                                       TY* __restrict__ y, double* __restrict__ partials,
                                                       const int* fl, int reduce, const TR* __restrict__ rv,
                                                       const TD* __restrict__ dv, double omega) {
  using M = Mfma<T>;
  if (fl && fl[FL_DONE]) return;
  __shared__ T sU[2][3][MK_HT][MK_HT];      // [buffer][component][x][y]
  __shared__ T sE[2][MK_CT][MK_CT];         // [buffer][cell x][cell y]
  __shared__ T sY[24][MK_CT * MK_CT];       // [row][cell y + 16 cell x]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int y0 = blockIdx.x * MK_NT, x0 = blockIdx.y * MK_NT;
  const int zc0 = blockIdx.z * MK_ZC;
  const int zc1 = min(g.NZ, zc0 + MK_ZC);
  const long long NN = g.NN;
  const long long plane = (long long)g.NY * g.NX;
  // Ke fragments: A[i][k] = Ke[16 mb + i][4 kb + k] (rows >= 24 are zero)
  T ka[2][6];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int kb = 0; kb < 6; ++kb) {
      const int i = 16 * mb + (lane & 15), k = 4 * kb + (lane >> 4);
      ka[mb][kb] = i < 24 ? Ke[i * 24 + k] : T(0);
    }
  // ---- staging, split into global loads (issued a layer ahead, into
  // registers) and LDS stores (after the matrix-core phase has released the
  // buffer), so the load latency overlaps the MFMAs.
  constexpr int NST = (MK_HT * MK_HT + MK_T - 1) / MK_T;  // staged nodes per thread (2)
  T pu[NST][3], pe;
  auto load_stage = [&](int z) {
#pragma unroll
    for (int s2 = 0; s2 < NST; ++s2) {
      const int t = tid + s2 * MK_T;
      const int hy = t % MK_HT, hx = t / MK_HT;
      const int gy = y0 - 1 + hy, gx = x0 - 1 + hx;
      pu[s2][0] = pu[s2][1] = pu[s2][2] = T(0);
      if (t < MK_HT * MK_HT && z >= 0 && z < g.NZ && gy >= 0 && gx >= 0 && gy < g.NY && gx < g.NX) {
        const long long n = gy + (long long)g.NY * gx + plane * z;
        const uint8_t f = fix[n];
        pu[s2][0] = (f & 1) ? T(0) : T(u[n]);
        pu[s2][1] = (f & 2) ? T(0) : T(u[NN + n]);
        pu[s2][2] = (f & 4) ? T(0) : T(u[2 * NN + n]);
      }
    }
    {
      const int ey = tid % MK_CT, ex = tid / MK_CT;  // MK_CT * MK_CT == MK_T
      const int gy = y0 - 1 + ey, gx = x0 - 1 + ex;
      pe = T(0);
      if (z >= 0 && z < g.nz && gy >= 0 && gx >= 0 && gy < g.ny && gx < g.nx)
        pe = T(E[gy + (long long)g.NY * gx + plane * z]);
    }
  };
  auto store_stage = [&](int b) {
#pragma unroll
    for (int s2 = 0; s2 < NST; ++s2) {
      const int t = tid + s2 * MK_T;
      if (t < MK_HT * MK_HT) {
        const int hy = t % MK_HT, hx = t / MK_HT;
        sU[b][0][hx][hy] = pu[s2][0];
        sU[b][1][hx][hy] = pu[s2][1];
        sU[b][2][hx][hy] = pu[s2][2];
      }
    }
    sE[b][tid / MK_CT][tid % MK_CT] = pe;
  };
  // node thread of the gather: (ty, tx) in 15 x 15
  const int ty = tid % 16, tx = tid / 16;
  const bool gnode = ty < MK_NT && tx < MK_NT;
  const int yy = y0 + ty, xx = x0 + tx;
  const bool own = gnode && yy < g.NY && xx < g.NX;
  T a0[3] = {0, 0, 0}, a1[3] = {0, 0, 0};
  double pq = 0.0;
  const int kBeg = zc0 - 1;
  load_stage(kBeg);
  store_stage(kBeg & 1);
  load_stage(kBeg + 1);
  store_stage((kBeg + 1) & 1);
  for (int k = kBeg; k < zc1; ++k) {
    const bool more = k + 2 <= zc1;
    if (more) load_stage(k + 2);  // in flight during this layer's MFMAs
    // epilogue operands of node layer k, also loaded ahead
    const bool epi = k >= zc0 && own;
    const long long n = yy + (long long)g.NY * xx + plane * k;
    uint8_t f = 0;
    double uo[3] = {0.0, 0.0, 0.0}, ro[3] = {0.0, 0.0, 0.0};
    T dd[3] = {T(0), T(0), T(0)};
    if (epi) {
      f = fix[n];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        uo[c] = u[c * NN + n];
        if (EPI != MEPI_KU) ro[c] = double(rv[c * NN + n]);
        if (EPI == MEPI_SMOOTH) dd[c] = T(dv[c * NN + n]);
      }
    }
    __syncthreads();  // node layers k, k+1 and cell layer k's moduli staged; sY free
    const int bk = k & 1, bk1 = (k + 1) & 1;
    // ---- matrix-core phase: cells (cy, cx), cx = 4 wave .. 4 wave + 3
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int cx = 4 * wave + q, cy = lane & 15;
      typename M::V c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < 6; ++kb) {
        const int d = 4 * kb + (lane >> 4);  // dof 0..23
        const int a = d / 3, comp = d - 3 * (d / 3);
        const int hy = cy + kLocDy[a], hx = cx + kLocDx[a];
        const T b = sU[kLocDz[a] ? bk1 : bk][comp][hx][hy];
        c0 = M::mma(ka[0][kb], b, c0);
        c1 = M::mma(ka[1][kb], b, c1);
      }
      const T e = sE[bk][cx][cy];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = M::row(lane, r);
        sY[i][cy + MK_CT * cx] = c0[r] * e;
        if (i < 8) sY[16 + i][cy + MK_CT * cx] = c1[r] * e;
      }
    }
    __syncthreads();  // sY complete; node layer k's buffer free
    if (more) store_stage(bk);
    // ---- gather: rows of this node (bottom: layer k, top: layer k+1) from its 4 cells
    if (gnode) {
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const int CY = cc & 1, CX = cc >> 1;
        const int col = (ty + CY) + MK_CT * (tx + CX);
        const int jb = loc_index(1 - CY, 1 - CX, 0), jt = loc_index(1 - CY, 1 - CX, 1);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          a0[r] += sY[3 * jb + r][col];
          a1[r] += sY[3 * jt + r][col];
        }
      }
    }
    if (epi) {  // node layer k complete
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T ku = ((f >> c) & 1) ? T(uo[c]) : a0[c];
        if (EPI == MEPI_KU) {
          const TY v = TY(ku);
          y[c * NN + n] = v;
          pq += uo[c] * double(v);
        } else if (EPI == MEPI_RESID) {
          y[c * NN + n] = TY(T(ro[c]) - ku);
        } else {
          const TY v = dd[c] != T(0) ? TY(T(uo[c]) + T(omega) * (T(ro[c]) - ku) / dd[c]) : TY(0);
          y[c * NN + n] = v;
          pq += double(v) * ro[c];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      a0[r] = a1[r];
      a1[r] = 0;
    }
  }
  if (reduce) {
    __shared__ double shp[MK_T / 64];
    const double v = wave_sum(pq);
    if (lane == 0) shp[wave] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < MK_T / 64; ++w) s += shp[w];
      partials[blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z)] = s;
    }
  }
}

dim3 applyk_mfma_grid(const Geo& g) {
  return dim3((g.NY + MK_NT - 1) / MK_NT, (g.NX + MK_NT - 1) / MK_NT, (g.NZ + MK_ZC - 1) / MK_ZC);
}

int applyk_mfma_blocks(const Geo& g) {
  const dim3 d = applyk_mfma_grid(g);
  return (int)(d.x * d.y * d.z);
}

void launch_fine_resid32_mfma(hipStream_t s, const Geo& g, const float* Ke, const float* E, const uint8_t* fix,
                              const float* x, const double* r, float* d, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_mfma<float, float, float, float, MEPI_RESID, double, float>), applyk_mfma_grid(g),
                     dim3(MK_T), 0, s, g, Ke, E, fix, x, d, (double*)nullptr, fl, 0, r, (const float*)nullptr, 0.0);
}

void launch_fine_smooth32_mfma(hipStream_t s, const Geo& g, const float* Ke, const float* E, const uint8_t* fix,
                               const float* x, const double* r, const float* diag, double omega, float* out,
                               double* partials, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_mfma<float, float, float, float, MEPI_SMOOTH, double, float>), applyk_mfma_grid(g),
                     dim3(MK_T), 0, s, g, Ke, E, fix, x, out, partials, fl, partials ? 1 : 0, r, diag, omega);
}

}  // namespace t3x
