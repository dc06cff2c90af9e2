// Fine-level K*u with the hex8 element operator in sum-factorised form
// (Solving_KbyU_MatrixFree, TOP3D_XL.m:641-726, with the element matrix of
// FEA_VoxelBasedElementStiffnessMatrix, 1254-1292).
//
// For the unit brick with trilinear shape functions, every block of Ke is a
// sum of tensor products of three 1-D 2x2 matrices, one per axis:
//   K1 = d d^T, M1 = h h^T + d d^T / 12, C1 = d h^T (and C1^T),
//   d = (-1, 1), h = (1/2, 1/2),
// with Ke(a i, b j) = lam T_ij + mu T_ji + mu delta_ij sum_k T_kk, where T_ij
// takes the derivative of the test function along i and of the trial
// function along j.  In the basis of sums and differences along each axis
// every 1-D factor maps one coefficient to at most one coefficient, so
//   Ke u_e = S^T W(lam, mu) S u_e,
// S: sums and differences of each component's 8 nodal values along x, y and
// z (24 adds per component), W: 21 coupled coefficients with 8 constants.
// 196 flops per cell against 1152 for the dense 24x24 product; the
// operator is the same up to rounding (the constants follow from E, nu and a
// unit cell, the same inputs as the dense Ke).
//
// Tiling: a work-group owns 15x15 (y, x) node columns and the 16x16 cells
// around them and marches a chunk of node layers in z (two node layers and
// the cell moduli double-buffered in LDS, the next layer's global loads in
// flight during the cell phase); one thread per cell forms E_e Ke u_e, then
// each node thread adds the rows of its node from the 4 cells around its
// column in a fixed order (deterministic, no atomics).  A thread's cells
// of consecutive layers share a face: the x-y transform of that face and
// the top-face output of the cell below stay in registers, so each layer
// costs one face analysis and one face synthesis (about 150 flops per cell
// instead of 196, a third fewer LDS reads).  Boundary conditions: fixed DOFs
// are zeroed where the cells read them (P_free u) and fixed rows return u
// (703-715, 1338-1437).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "t3x_kernels.h"
#include "t3x_reduce.h"

namespace t3x {

namespace {

constexpr int SF_NT = 15;          // node columns per tile edge
constexpr int SF_CT = SF_NT + 1;   // cells per tile edge
constexpr int SF_HT = SF_NT + 2;   // staged nodes per tile edge
constexpr int SF_ZC = 26;          // node layers per work-group (at most; see sf_zc): FP64 K*p at
                                   // 256x512x256 0.60 ms against 0.62 (16) and 0.63 (40), interleaved runs
constexpr int SF_ZC_SPARSE = 16;   // the same on domains with void tiles (shorter chunks: more tiles skipped)
constexpr int SF_ZC_MIN = 4;       // fewest node layers per work-group when splitting for parallelism
constexpr int SF_MIN_GROUPS = 2048;  // work-groups below which the z chunks are shortened
constexpr int SF_T = SF_CT * SF_CT;  // 256 threads: one per cell

// Position bits of local node a in the transform: x = dx, y = 1 - dy (the
// reference's y axis points up the rows), z = dz.
__device__ __forceinline__ constexpr int sf_pos(int a) {
  return kLocDx[a] + 2 * (1 - kLocDy[a]) + 4 * kLocDz[a];
}

// Value of lane + 1 within its row of 16 lanes (DPP row_shl:1; the last lane
// of a row reads 0).  Every lane of the wave must be active.
__device__ __forceinline__ float next_in_row(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xF, 0xF, false));
}
__device__ __forceinline__ double next_in_row(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x101, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x101, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// One face of a cell (4 nodes at positions p = x + 2y, the z = 0 half of
// sf_pos) in place: values -> (sum, difference) coefficients along x, then
// along y, same bit layout.  The z stage of the cell transform combines the
// bottom and top faces' coefficients: (top + bottom, top - bottom).
template <typename T>
__device__ __forceinline__ void face_analyse(T f[4]) {
#pragma unroll
  for (int b = 1; b < 4; b <<= 1)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (!(j & b)) {
        const T s = f[j | b] + f[j], d = f[j | b] - f[j];
        f[j] = s;
        f[j | b] = d;
      }
}

// Transpose of face_analyse's differences: coefficients -> nodal values.
template <typename T>
__device__ __forceinline__ void face_synthesise(T f[4]) {
#pragma unroll
  for (int b = 2; b >= 1; b >>= 1)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (!(j & b)) {
        const T lo = f[j] - f[j | b], hi = f[j] + f[j | b];
        f[j] = lo;
        f[j | b] = hi;
      }
}

// E * W in transform space for one cell: analysed coefficients X, Y, Z of
// the three components -> ox, oy, oz.  c[0..7]: the constants for E = 1
// (sf_coefficients), scaled here by e.
template <typename T>
__device__ __forceinline__ void sf_couple(const T X[8], const T Y[8], const T Z[8], T ox[8], T oy[8], T oz[8],
                                          const T* c, T e) {
  T k[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) k[i] = c[i] * e;
  const T t = X[1] + Y[2] + Z[4];
  ox[0] = oy[0] = oz[0] = T(0);
  ox[1] = k[0] * X[1] + k[1] * t;
  oy[2] = k[0] * Y[2] + k[1] * t;
  oz[4] = k[0] * Z[4] + k[1] * t;
  ox[2] = oy[1] = k[2] * (X[2] + Y[1]);
  ox[4] = oz[1] = k[2] * (X[4] + Z[1]);
  oy[4] = oz[2] = k[2] * (Y[4] + Z[2]);
  ox[3] = k[3] * X[3] + k[4] * Z[6];
  ox[5] = k[3] * X[5] + k[4] * Y[6];
  oy[3] = k[3] * Y[3] + k[4] * Z[5];
  oy[6] = k[3] * Y[6] + k[4] * X[5];
  oz[5] = k[3] * Z[5] + k[4] * Y[3];
  oz[6] = k[3] * Z[6] + k[4] * X[3];
  ox[6] = k[5] * X[6] + k[6] * (Y[5] + Z[3]);
  oy[5] = k[5] * Y[5] + k[6] * (X[6] + Z[3]);
  oz[3] = k[5] * Z[3] + k[6] * (X[6] + Y[5]);
  ox[7] = k[7] * X[7];
  oy[7] = k[7] * Y[7];
  oz[7] = k[7] * Z[7];
}

}  // namespace

SfCoef sf_coefficients(double nu) {
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  const double c2 = mu / 6 + (lam + mu) / 12, c3 = mu / 48 + (lam + mu) / 144;
  return SfCoef{{mu / 8, lam / 16, mu / 16, c2 / 4, lam / 48, mu / 24, mu / 48, c3}};
}

enum { SEPI_KU = 0, SEPI_RESID = 1, SEPI_SMOOTH = 2 };

// Epilogues per node (as the dense kernels):
//   SEPI_KU     y = Ku                         partials: u.y (the PCG's p.q)
//   SEPI_RESID  y = r - Ku                     (the 'Standard' cycle's defect, 761)
//   SEPI_SMOOTH y = u + omega (r - Ku) / d     partials: y.r (post-smoothing, 775)
// T: arithmetic/LDS type; TE, TU, TY, TR, TD: types of E, u, y, r, diag in HBM.
template <typename T, typename TE, typename TU, typename TY, int EPI, typename TR, typename TD>
// (the FP32 residual instance bounded to five waves per SIMD: 95 VGPRs
// without scratch; the smoothing instance would spill)
__global__ __launch_bounds__(SF_T) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && EPI == SEPI_RESID ? 5 : 1)))
void k_apply_K_sf(Geo g, int zc, SfCoef cf, const TE* __restrict__ E,
                                                     const uint8_t* __restrict__ fix, const TU* __restrict__ u,
                                                     TY* __restrict__ y, double* __restrict__ partials,
                                                     const int* fl, int reduce, const TR* __restrict__ rv,
                                                     const TD* __restrict__ dv, double omega) {
  if (fl && fl[FL_DONE]) return;
  constexpr int CS = SF_HT * SF_HT;       // one component of one staged node layer
  __shared__ T sU[2 * 3 * CS];            // [buffer][component][x][y]
  __shared__ T sE[2][SF_CT * SF_CT];      // [buffer][cell y + 16 cell x]
  __shared__ uint8_t sF[2][CS];           // fixation bits of the staged nodes
  // Gather of the cell rows into node rows: the y-neighbour pair is summed in
  // registers across lanes and only the x-neighbour partial sums (3 rows)
  // go through LDS.
  __shared__ T sR[3][SF_CT * SF_CT];  // [row][cell y + 16 cell x]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware tile order (xcd_tile_order): neighbouring tiles share an L2
  // On a domain with void tiles the grid is the list of tiles that hold a
  // solid cell (the others' owned nodes are void and stay 0).
  const dim3 tile = xcd_tile_order();
  int bx = tile.x, by = tile.y, bz = tile.z;
  const size_t tileId = bx + (size_t)gridDim.x * (by + (size_t)gridDim.y * bz);  // partials slot
  if (g.tiles) {
    const int lin = g.tiles[tileId];
    const int tx = (g.NY + SF_NT - 1) / SF_NT, txy = tx * ((g.NX + SF_NT - 1) / SF_NT);
    bz = lin / txy;
    by = (lin - bz * txy) / tx;
    bx = lin - bz * txy - by * tx;
  }
  const int y0 = bx * SF_NT, x0 = by * SF_NT;
  const int zc0 = bz * zc;
  const int zc1 = min(g.NZ, zc0 + zc);
  const long long NN = g.NN;
  const long long plane = (long long)g.NY * g.NX;
  T cst[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cst[i] = T(cf.k[i]);
  // staging: global loads a layer ahead into registers, LDS stores after the
  // cell phase has released the buffer.  u is staged unmasked beside its
  // fixation bits: the cells apply P_free when they read their nodes, and
  // the node's own u and bits for the epilogue come from the same stage.
  constexpr int NST = (CS + SF_T - 1) / SF_T;  // staged nodes per thread (2)
  T pu[NST][3], pe;
  uint8_t pf[NST];
  auto load_stage = [&](int z) {
#pragma unroll
    for (int s2 = 0; s2 < NST; ++s2) {
      const int t = tid + s2 * SF_T;
      const int hy = t % SF_HT, hx = t / SF_HT;
      const int gy = y0 - 1 + hy, gx = x0 - 1 + hx;
      pu[s2][0] = pu[s2][1] = pu[s2][2] = T(0);
      pf[s2] = 0;
      if (t < CS && z >= 0 && z < g.NZ && gy >= 0 && gx >= 0 && gy < g.NY && gx < g.NX) {
        const long long n = gy + (long long)g.NY * gx + plane * z;
        pf[s2] = fix[n];
        pu[s2][0] = T(u[n]);
        pu[s2][1] = T(u[NN + n]);
        pu[s2][2] = T(u[2 * NN + n]);
      }
    }
    {
      const int ey = tid % SF_CT, ex = tid / SF_CT;
      const int gy = y0 - 1 + ey, gx = x0 - 1 + ex;
      pe = T(0);
      if (z >= 0 && z < g.nz && gy >= 0 && gx >= 0 && gy < g.ny && gx < g.nx)
        pe = T(E[gy + (long long)g.NY * gx + plane * z]);
    }
  };
  auto store_stage = [&](int b) {
#pragma unroll
    for (int s2 = 0; s2 < NST; ++s2) {
      const int t = tid + s2 * SF_T;
      if (t < CS) {
        const int hy = t % SF_HT, hx = t / SF_HT;
        sU[(b * 3 + 0) * CS + hx * SF_HT + hy] = pu[s2][0];
        sU[(b * 3 + 1) * CS + hx * SF_HT + hy] = pu[s2][1];
        sU[(b * 3 + 2) * CS + hx * SF_HT + hy] = pu[s2][2];
        sF[b][hx * SF_HT + hy] = pf[s2];
      }
    }
    sE[b][tid] = pe;
  };
  // cell thread: (cy, cx) in 16 x 16; node thread of the gather: (ty, tx) in 15 x 15
  const int cy = tid % SF_CT, cx = tid / SF_CT;
  const int ty = cy, tx = cx;
  const bool gnode = ty < SF_NT && tx < SF_NT;
  const int yy = y0 + ty, xx = x0 + tx;
  const bool own = gnode && yy < g.NY && xx < g.NX;
  // the 4 nodes of cell (cy, cx) in one staged layer, P_free applied as
  // read, face-analysed
  auto read_face = [&](int b, T F[3][4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {  // local nodes 0..3: the z = 0 face
      const int p = sf_pos(a);
      const int h = (cx + kLocDx[a]) * SF_HT + (cy + kLocDy[a]);
      const int o = (b * 3) * CS + h;
      const uint8_t fa = sF[b][h];
      F[0][p] = (fa & 1) ? T(0) : sU[o];
      F[1][p] = (fa & 2) ? T(0) : sU[o + CS];
      F[2][p] = (fa & 4) ? T(0) : sU[o + 2 * CS];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) face_analyse(F[c]);
  };
  double pq = 0.0;
  const int kBeg = zc0 - 1;
  load_stage(kBeg);
  store_stage(kBeg & 1);
  load_stage(kBeg + 1);
  store_stage((kBeg + 1) & 1);
  __syncthreads();
  // The cell of layer k shares its bottom face with the top face of the cell
  // of layer k - 1: the face's analysed coefficients (Fb) and the previous
  // cell's top-face output coefficients (Ot) stay in registers, so a step
  // reads and analyses one node layer and synthesises one (the sum of the
  // two cells' coefficients on the shared face).
  T Fb[3][4], Ot[3][4];
  read_face(kBeg & 1, Fb);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) Ot[c][j] = T(0);
  for (int k = kBeg; k < zc1; ++k) {
    const bool more = k + 2 <= zc1;
    if (more) load_stage(k + 2);  // in flight during this layer's cell phase
    const bool epi = k >= zc0 && own;
    const long long n = yy + (long long)g.NY * xx + plane * k;
    uint8_t f = 0;
    double uo[3] = {0.0, 0.0, 0.0}, ro[3] = {0.0, 0.0, 0.0};
    T dd[3] = {T(0), T(0), T(0)};
    if (epi) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (EPI != SEPI_KU) ro[c] = double(rv[c * NN + n]);
        if (EPI == SEPI_SMOOTH) dd[c] = T(dv[c * NN + n]);
      }
    }
    __syncthreads();  // node layers k, k+1 and cell layer k's moduli staged; sR free
    const int bk = k & 1, bk1 = (k + 1) & 1;
    // ---- cell phase: cell (cy, cx) of layer k.  Lanes hold cells y-fastest in
    // rows of 16, so the cell at y + 1 is the next lane of the row: the node
    // thread (ty, tx) = (cy, cx) sums its rows from cells (cy, cx) and
    // (cy + 1, cx) in registers (q), and the pair at x + 1 comes through LDS
    // (sR) from the thread of column cx + 1, which formed the same sum for
    // the node at its x - 1.
    T q[3];
    {
      T Ft[3][4];
      read_face(bk1, Ft);
      if (epi) {  // this node's bits and unmasked u, before its buffer is restaged
        const int h = (tx + 1) * SF_HT + (ty + 1);
        f = sF[bk][h];
#pragma unroll
        for (int c = 0; c < 3; ++c) uo[c] = double(sU[(bk * 3 + c) * CS + h]);
      }
      T X[8], Yv[8], Z[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {  // z stage: positions j (sum) and j + 4 (difference)
        X[j] = Ft[0][j] + Fb[0][j];
        X[j + 4] = Ft[0][j] - Fb[0][j];
        Yv[j] = Ft[1][j] + Fb[1][j];
        Yv[j + 4] = Ft[1][j] - Fb[1][j];
        Z[j] = Ft[2][j] + Fb[2][j];
        Z[j + 4] = Ft[2][j] - Fb[2][j];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) Fb[c][j] = Ft[c][j];
      T ox[8], oy[8], oz[8];
      sf_couple(X, Yv, Z, ox, oy, oz, cst, sE[bk][tid]);
      // transposed z stage: the bottom face (node layer k) joins the previous
      // cell's top face, then the face is synthesised once
      T S[3][4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        S[0][j] = Ot[0][j] + (ox[j] - ox[j + 4]);
        S[1][j] = Ot[1][j] + (oy[j] - oy[j + 4]);
        S[2][j] = Ot[2][j] + (oz[j] - oz[j + 4]);
        Ot[0][j] = ox[j] + ox[j + 4];
        Ot[1][j] = oy[j] + oy[j + 4];
        Ot[2][j] = oz[j] + oz[j + 4];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) face_synthesise(S[c]);
      // face position of local (dy, dx): dx + 2 (1 - dy); (1, 1), (0, 1) are
      // this node column, (1, 0), (0, 0) the column at x - 1
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        q[c] = S[c][1] + next_in_row(S[c][3]);
        sR[c][tid] = S[c][0] + next_in_row(S[c][2]);
      }
    }
    __syncthreads();  // sR complete; node layer k's buffer free
    // FP64: stored on every layer (past the chunk's last layer the stage is
    // stale and lands in a buffer no later layer reads), so no load into the
    // stage is outstanding on any path when the next layer overwrites it and
    // the compiler does not wait there for this layer's global stores (FP64
    // K*p 0.61 -> 0.59 ms at 256x512x256; the FP32 instances measured slower
    // with it, 0.63 -> 0.66 ms, and keep the conditional store).
    if (sizeof(T) == 8 || more) store_stage(bk);
    if (epi) {  // node layer k complete
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T a = q[c] + sR[c][tid + SF_CT];
        const T ku = ((f >> c) & 1) ? T(uo[c]) : a;
        if (EPI == SEPI_KU) {
          const TY v = TY(ku);
          y[c * NN + n] = v;
          pq += uo[c] * double(v);
        } else if (EPI == SEPI_RESID) {
          y[c * NN + n] = TY(T(ro[c]) - ku);
        } else {
          const TY v = dd[c] != T(0) ? TY(T(uo[c]) + T(omega) * (T(ro[c]) - ku) / dd[c]) : TY(0);
          y[c * NN + n] = v;
          pq += double(v) * ro[c];
        }
      }
    }
  }
  if (reduce) {
    __shared__ double shp[SF_T / 64];
    const double v = wave_sum(pq);
    if (lane == 0) shp[wave] = v;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < SF_T / 64; ++w) s += shp[w];
      partials[tileId] = s;
    }
  }
}

// Unit compliance ce_e = u_e^T Ke u_e with the unmodified Ke (579-598), in
// the sum-factorised form: the synthesis is the transpose of the analysis,
// so u_e^T (S^T W S) u_e = (S u_e)^T W (S u_e) with W the coupling for E = 1
// (about 150 flops per cell against 1200 for the dense 24x24 product; the
// same value up to rounding).  One thread per cell, n the cell's lower
// corner node.
__global__ __launch_bounds__(256) void k_compliance_sf(Geo g, SfCoef cf, const uint8_t* __restrict__ emask,
                                                      const double* __restrict__ u, double* __restrict__ ce) {
  const long long NN = g.NN;
  const long long n = (long long)blockIdx.x * 256 + threadIdx.x;
  if (n >= NN) return;
  if (!emask[n]) {
    ce[n] = 0.0;
    return;
  }
  const long long plane = (long long)g.NY * g.NX;
  double Fb[3][4], Ft[3][4];  // the z = 0 and z = 1 faces, positions x + 2 (1 - y)
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long m = n + kLocDy[a] + (long long)g.NY * kLocDx[a] + plane * kLocDz[a];
    const int p = sf_pos(a) & 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (kLocDz[a]) Ft[c][p] = u[c * NN + m];
      else Fb[c][p] = u[c * NN + m];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    face_analyse(Fb[c]);
    face_analyse(Ft[c]);
  }
  double X[8], Y[8], Z[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    X[j] = Ft[0][j] + Fb[0][j];
    X[j + 4] = Ft[0][j] - Fb[0][j];
    Y[j] = Ft[1][j] + Fb[1][j];
    Y[j + 4] = Ft[1][j] - Fb[1][j];
    Z[j] = Ft[2][j] + Fb[2][j];
    Z[j + 4] = Ft[2][j] - Fb[2][j];
  }
  double cst[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cst[i] = cf.k[i];
  double ox[8], oy[8], oz[8];
  sf_couple(X, Y, Z, ox, oy, oz, cst, 1.0);
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += X[j] * ox[j] + Y[j] * oy[j] + Z[j] * oz[j];
  ce[n] = s;
}

void launch_compliance_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const uint8_t* emask, const double* u,
                          double* ce) {
  const long long b = (g.NN + 255) / 256;
  hipLaunchKernelGGL(k_compliance_sf, dim3((unsigned)(b < 1 ? 1 : b)), dim3(256), 0, s, g, cf, emask, u, ce);
}

static long long sf_tiles_xy(int NY, int NX) {
  return (long long)((NY + SF_NT - 1) / SF_NT) * ((NX + SF_NT - 1) / SF_NT);
}

// z chunk: the NZ node layers split into equal chunks of about SF_ZC
// layers, more (down to SF_ZC_MIN layers) while the grid has fewer than
// SF_MIN_GROUPS work-groups, fewer while it would write more than
// kMaxPartials partials; a node plane with more tiles than that is refused
// at set-up (sf_plane_fits)
static int sf_zc(const Geo& g, bool sparse) {
  const long long bxy = sf_tiles_xy(g.NY, g.NX);
  const int zcMax = sparse ? SF_ZC_SPARSE : SF_ZC;
  long long nch = (g.NZ + zcMax - 1) / zcMax;
  const long long want = (SF_MIN_GROUPS + bxy - 1) / bxy;
  if (nch < want) nch = std::min<long long>(want, (g.NZ + SF_ZC_MIN - 1) / SF_ZC_MIN);
  if (nch > kMaxPartials / bxy) nch = std::max<long long>(1, kMaxPartials / bxy);
  if (nch < 1) nch = 1;
  return (int)((g.NZ + nch - 1) / nch);
}

bool sf_plane_fits(int NY, int NX) { return sf_tiles_xy(NY, NX) <= kMaxPartials; }

static dim3 sf_grid_full(const Geo& g, bool sparse) {
  const int zc = sf_zc(g, sparse);
  return dim3((g.NY + SF_NT - 1) / SF_NT, (g.NX + SF_NT - 1) / SF_NT, (g.NZ + zc - 1) / zc);
}

// the launch shape of a level: the full tile grid, or one work-group per
// listed tile (sparse chunking) when the level carries a tile list
static dim3 sf_grid_of(const Geo& g) { return g.tiles ? dim3(std::max(g.nTiles, 1)) : sf_grid_full(g, false); }
static int sf_zc_of(const Geo& g) { return sf_zc(g, g.tiles != nullptr); }

std::vector<int> sf_active_tiles(const Geo& g, const std::vector<uint8_t>& emask) {
  const dim3 d = sf_grid_full(g, true);
  const int zc = sf_zc(g, true);
  std::vector<uint8_t> f((size_t)d.x * d.y * d.z, 0);
  // tile (bx, by, bz) stages cells y in [SF_NT bx - 1, SF_NT bx + SF_NT - 1],
  // x likewise, z in [zc bz - 1, min(NZ, zc (bz + 1)) - 1]
  auto range = [](int c, int step, int n, int& lo, int& hi) {
    lo = std::min(n - 1, c / step);      // smallest b with step b + step - 1 >= c
    hi = std::min(n - 1, (c + 1) / step);  // largest b with step b - 1 <= c
  };
  for (int z = 0; z < g.nz; ++z)
    for (int x = 0; x < g.nx; ++x)
      for (int y = 0; y < g.ny; ++y) {
        if (!emask[y + (size_t)g.NY * (x + (size_t)g.NX * z)]) continue;
        int y0, y1, x0, x1, z0, z1;
        range(y, SF_NT, (int)d.x, y0, y1);
        range(x, SF_NT, (int)d.y, x0, x1);
        range(z, zc, (int)d.z, z0, z1);
        for (int bz = z0; bz <= z1; ++bz)
          for (int by = x0; by <= x1; ++by)
            for (int bx = y0; bx <= y1; ++bx) f[bx + (size_t)d.x * (by + (size_t)d.y * bz)] = 1;
      }
  std::vector<int> list;
  for (size_t t = 0; t < f.size(); ++t)
    if (f[t]) list.push_back((int)t);
  if (list.size() == f.size()) return {};
  return list;
}

int applyk_sf_blocks(const Geo& g) {
  const dim3 d = sf_grid_of(g);
  return (int)(d.x * d.y * d.z);
}

void launch_apply_K_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                       const double* u, double* y, double* partials, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_sf<double, double, double, double, SEPI_KU, double, double>), sf_grid_of(g), dim3(SF_T),
                     0, s, g, sf_zc_of(g), cf, E, fix, u, y, partials, fl, partials ? 1 : 0, (const double*)nullptr,
                     (const double*)nullptr, 0.0);
}

void launch_fine_resid64_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                            const double* x, const double* r, double* d, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_sf<double, double, double, double, SEPI_RESID, double, double>), sf_grid_of(g),
                     dim3(SF_T), 0, s, g, sf_zc_of(g), cf, E, fix, x, d, (double*)nullptr, fl, 0, r,
                     (const double*)nullptr, 0.0);
}

void launch_fine_smooth64_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                             const double* x, const double* r, const double* diag, double omega, double* out,
                             double* partials, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_sf<double, double, double, double, SEPI_SMOOTH, double, double>), sf_grid_of(g),
                     dim3(SF_T), 0, s, g, sf_zc_of(g), cf, E, fix, x, out, partials, fl, partials ? 1 : 0, r, diag,
                     omega);
}

void launch_fine_resid32_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const float* E, const uint8_t* fix,
                            const float* x, const double* r, float* d, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_sf<float, float, float, float, SEPI_RESID, double, float>), sf_grid_of(g), dim3(SF_T),
                     0, s, g, sf_zc_of(g), cf, E, fix, x, d, (double*)nullptr, fl, 0, r, (const float*)nullptr, 0.0);
}

void launch_fine_smooth32_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const float* E, const uint8_t* fix,
                             const float* x, const double* r, const float* diag, double omega, float* out,
                             double* partials, const int* fl) {
  hipLaunchKernelGGL((k_apply_K_sf<float, float, float, float, SEPI_SMOOTH, double, float>), sf_grid_of(g),
                     dim3(SF_T), 0, s, g, sf_zc_of(g), cf, E, fix, x, out, partials, fl, partials ? 1 : 0, r, diag,
                     omega);
}

}  // namespace t3x
