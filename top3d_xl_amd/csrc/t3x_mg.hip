// Multigrid transfers of the 'Adapted' V-cycle (TOP3D_XL.m:747-814) for gfx950.
//
// R = P^T with P the trilinear hat interpolation over the solid region; the
// reference's scatter into intermediate nodes, division by transferMatCoeffi
// and solid-node selection (783-814) reduce to masked tensor-product
// stencils with 1-D weights (S-|d|)/S, S the span (4 for the first
// coarsening, 2 after).  Vectors are zero outside the solid region, so only
// the coarse mask is applied.
//
// Restriction is separable: a y pass (fine -> NYc x NXf x NZf) and an x-z
// pass (-> coarse), each thread owning one output, a wave reading
// consecutive y.  The y pass reads every fine value once from HBM; the
// intermediate is 1/S of the fine vector.  Prolongation + Jacobi runs one
// thread per (y, x) column over a short run of z planes, so the coarse-column
// weights are computed once per thread and the grid still has tens of
// thousands of blocks.
//
// Templates over the vector types give the FP64 path (double everywhere) and
// the mixed path (FP64 fine residual, FP32 coarse levels and smoother
// output) from the same code.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "t3x_kernels.h"
#include "t3x_reduce.h"

namespace t3x {

namespace {

template <int S>
__device__ __forceinline__ double hat_w(int d) {
  return (double)(S - (d < 0 ? -d : d)) * (1.0 / S);
}

// y pass: tmp[c][z][x][Jy] = sum_dy w(dy) rf[c](S*Jy+dy, x, z)
template <int S, typename TI, typename TM>
__global__ __launch_bounds__(kBlock) void k_restrict_y(Geo gf, int NYc, const TI* __restrict__ rf,
                                                       TM* __restrict__ tmp, const int* fl, Act actF) {
  if (fl && fl[FL_DONE]) return;
  const int plane = NYc * gf.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= plane) return;
  const int x = idx / NYc, Jy = idx - x * NYc;
  const int z = blockIdx.y;
  const long long base = (long long)gf.NY * ((long long)x + (long long)gf.NX * z);
  const long long TN = (long long)plane * gf.NZ;
  const long long o = idx + (long long)plane * z;
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  // the S * 2 - 1 fine nodes lie in at most two chunks; void ones are 0
  const bool skip = actF.chunk && !actF.chunk[(base + max(S * Jy - (S - 1), 0)) >> 4] &&
                    !actF.chunk[(base + min(S * Jy + S - 1, gf.NY - 1)) >> 4];
  if (!skip) {
#pragma unroll
    for (int dy = -(S - 1); dy <= S - 1; ++dy) {
      const int y = S * Jy + dy;
      if (y < 0 || y >= gf.NY) continue;
      const double w = hat_w<S>(dy);
      const long long f = base + y;
      t0 += w * (double)rf[f];
      t1 += w * (double)rf[gf.NN + f];
      t2 += w * (double)rf[2 * gf.NN + f];
    }
  }
  tmp[o] = (TM)t0;
  tmp[TN + o] = (TM)t1;
  tmp[2 * TN + o] = (TM)t2;
}

// The y pass through LDS: block b takes the rows [R b, R b + R) (row
// x + NX z: NY fine values and NYc intermediate ones, both contiguous), loads
// the fine rows into LDS with coalesced loads (values of void 16-node chunks
// are 0 and not read), then computes one output per thread from LDS with the
// weights and summation order of k_restrict_y, so the result is the same.
// The direct kernel's loads are stride S across lanes, each fine value read
// by 2 S - 1 threads.
template <int S, typename TI, typename TM>
__global__ __launch_bounds__(kBlock) void k_restrict_y_lds(Geo gf, int NYc, int R, const TI* __restrict__ rf,
                                                           TM* __restrict__ tmp, const int* fl, Act actF) {
  if (fl && fl[FL_DONE]) return;
  extern __shared__ __align__(16) unsigned char smem[];
  TI* sv = (TI*)smem;  // [component][R NY]
  const long long rows = (long long)gf.NX * gf.NZ;
  const long long r0 = (long long)blockIdx.x * R;
  const int nr = (int)min((long long)R, rows - r0);
  const int len = nr * gf.NY, cs = R * gf.NY;
  const long long f0 = r0 * gf.NY;
  constexpr int UL = 4;  // fine values per component and thread in flight
  for (int i0 = threadIdx.x; i0 < len; i0 += UL * kBlock) {
    TI v[3][UL];
#pragma unroll
    for (int u = 0; u < UL; ++u) {
      const int i = i0 + u * kBlock;
      const long long f = f0 + i;
      const bool on = i < len && !(actF.chunk && !actF.chunk[f >> 4]);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c][u] = on ? rf[c * gf.NN + f] : TI(0);
    }
#pragma unroll
    for (int u = 0; u < UL; ++u) {
      const int i = i0 + u * kBlock;
      if (i < len) {
#pragma unroll
        for (int c = 0; c < 3; ++c) sv[c * cs + i] = v[c][u];
      }
    }
  }
  __syncthreads();
  const long long TN = (long long)NYc * rows;
  const int nout = nr * NYc;
  for (int o = threadIdx.x; o < nout; o += kBlock) {
    const int rr = o / NYc, Jy = o - rr * NYc;
    const TI* row = sv + rr * gf.NY;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int dy = -(S - 1); dy <= S - 1; ++dy) {
      const int y = S * Jy + dy;
      if (y < 0 || y >= gf.NY) continue;
      const double w = hat_w<S>(dy);
      t0 += w * (double)row[y];
      t1 += w * (double)row[cs + y];
      t2 += w * (double)row[2 * cs + y];
    }
    const long long oo = r0 * NYc + o;
    tmp[oo] = (TM)t0;
    tmp[TN + oo] = (TM)t1;
    tmp[2 * TN + oo] = (TM)t2;
  }
}

// x-z pass: rc[c](J) = mask(J) * sum_dz w(dz) sum_dx w(dx) tmp[c](Jy, S*Jx+dx, S*Jz+dz)
template <int S, typename TM, typename TO>
__global__ __launch_bounds__(kBlock) void k_restrict_xz(Geo gf, Geo gc, const TM* __restrict__ tmp,
                                                        const uint8_t* __restrict__ nmaskC, TO* __restrict__ rc,
                                                        const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int planeC = gc.NY * gc.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= planeC) return;
  const int Jx = idx / gc.NY, Jy = idx - Jx * gc.NY;
  const int Jz = blockIdx.y;
  const long long n = idx + (long long)planeC * Jz;
  const int plane = gc.NY * gf.NX;  // intermediate plane: NYc x NXf
  const long long TN = (long long)plane * gf.NZ;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (nmaskC[n]) {
#pragma unroll
    for (int dz = -(S - 1); dz <= S - 1; ++dz) {
      const int z = S * Jz + dz;
      if (z < 0 || z >= gf.NZ) continue;
      const double wz = hat_w<S>(dz);
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int dx = -(S - 1); dx <= S - 1; ++dx) {
        const int x = S * Jx + dx;
        if (x < 0 || x >= gf.NX) continue;
        const double wx = hat_w<S>(dx);
        const long long m = Jy + (long long)gc.NY * x + (long long)plane * z;
        s0 += wx * (double)tmp[m];
        s1 += wx * (double)tmp[TN + m];
        s2 += wx * (double)tmp[2 * TN + m];
      }
      a0 += wz * s0;
      a1 += wz * s1;
      a2 += wz * s2;
    }
  }
  rc[n] = (TO)a0;
  rc[gc.NN + n] = (TO)a1;
  rc[2 * gc.NN + n] = (TO)a2;
}

// out = omega*(r/d) + P xc + (omega*r)/d on solid nodes, 0 elsewhere (758,
// 770, 773); optional fused z.r partials (629).  Block (bx, bz) covers 256
// (y, x) columns and the node planes [bz*zc, bz*zc + zc); a thread walks its
// column plane by plane, reloading the 8 coarse values (clamped neighbours
// carry weight 0) only when a solid node needs another coarse z interval;
// void nodes are neither loaded nor stored.
template <int S, typename TX, typename TR, typename TD, typename TO>
__global__ __launch_bounds__(kBlock) void k_prolong_jacobi(Geo gf, Geo gc, double omega, int zc,
                                                           const TX* __restrict__ xc, const TR* __restrict__ rf,
                                                           const TD* __restrict__ diag,
                                                           const uint8_t* __restrict__ nmaskF, TO* __restrict__ out,
                                                           double* __restrict__ partials, const int* fl) {
  if (fl && fl[FL_DONE]) return;
  const int plane = gf.NY * gf.NX;
  const long long NN = gf.NN, NC = gc.NN;
  const long long planeC = (long long)gc.NY * gc.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  double zr = 0.0;
  if (idx < plane) {
    const int x = idx / gf.NY, y = idx - x * gf.NY;
    const int Jy = y / S, Jx = x / S;
    const int ty = y - S * Jy, tx = x - S * Jx;
    const double wy[2] = {(double)(S - ty) * (1.0 / S), (double)ty * (1.0 / S)};
    const double wx[2] = {(double)(S - tx) * (1.0 / S), (double)tx * (1.0 / S)};
    const int Jy1 = min(Jy + 1, gc.NY - 1), Jx1 = min(Jx + 1, gc.NX - 1);
    const long long cyx[4] = {Jy + (long long)gc.NY * Jx, Jy1 + (long long)gc.NY * Jx, Jy + (long long)gc.NY * Jx1,
                              Jy1 + (long long)gc.NY * Jx1};
    const int z0 = blockIdx.y * zc, z1 = min(gf.NZ, z0 + zc);
    // the 8 coarse values of this column's cell, reloaded only when a solid
    // node needs a coarse z interval other than the loaded one
    double cv[8][3];
    int cJz = -1;
    for (int z = z0; z < z1; ++z) {
      const long long n = idx + (long long)plane * z;
      const bool solid = nmaskF[n] != 0;
      const int Jz = z / S, tz = z - S * Jz;
      if (solid && Jz != cJz) {
        cJz = Jz;
        const int Jz1 = min(Jz + 1, gc.NZ - 1);
        const long long zo[2] = {planeC * Jz, planeC * Jz1};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const long long J = cyx[(c & 1) + 2 * ((c >> 1) & 1)] + zo[c >> 2];
          cv[c][0] = (double)xc[J];
          cv[c][1] = (double)xc[NC + J];
          cv[c][2] = (double)xc[2 * NC + J];
        }
      }
      // void nodes hold 0 in every level vector (all of their writers leave
      // them 0): no load, no store, no interpolation
      if (solid) {
        const double wz[2] = {(double)(S - tz) * (1.0 / S), (double)tz * (1.0 / S)};
        double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int a = c & 1, b = (c >> 1) & 1, d = c >> 2;
          const double w = wy[a] * wx[b] * wz[d];
          p[0] += w * cv[c][0];
          p[1] += w * cv[c][1];
          p[2] += w * cv[c][2];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double r = (double)rf[k * NN + n];
          const double dd = solid ? (double)diag[k * NN + n] : 1.0;
          const TO v = solid ? (TO)((omega * (r / dd) + p[k]) + (omega * r) / dd) : TO(0);
          zr += (double)v * r;
          out[k * NN + n] = v;
        }
      }
    }
  }
  if (partials) {
    double v[1] = {zr};
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x + (size_t)gridDim.x * blockIdx.y] = v[0];
  }
}

// out += P xc on solid nodes (the 'Standard' cycle's coarse correction, 770),
// the grid of k_prolong_jacobi.  A thread takes its planes kPZ at a time: the
// masks first, then every load the solid ones need (the coarse values of the
// group's coarse z layers, the fine values) issued together, then the
// arithmetic — two dependent memory round trips per kPZ planes (fine
// 256x512x256 FP32: 199 against 290 us plane by plane; the Jacobi form above,
// with twice the fine operands, keeps its plane-by-plane loop: grouped, its
// FP64 instance needs 212 VGPRs and ran 531 against 461 us).
constexpr int kPZ = 4;
template <int S, typename TX, typename TO>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(S == 4 && sizeof(TO) == 4 ? 4 : 1)))
void k_prolong_add(Geo gf, Geo gc, int zc, const TX* __restrict__ xc,
                                                        const uint8_t* __restrict__ nmaskF, TO* __restrict__ out,
                                                        const int* fl) {
  static_assert(kPZ % S == 0, "a plane group spans whole coarse intervals");
  constexpr int NL = kPZ / S + 1;  // coarse z layers of a plane group
  if (fl && fl[FL_DONE]) return;
  const int plane = gf.NY * gf.NX;
  const long long NN = gf.NN, NC = gc.NN;
  const long long planeC = (long long)gc.NY * gc.NX;
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx < plane) {
    const int x = idx / gf.NY, y = idx - x * gf.NY;
    const int Jy = y / S, Jx = x / S;
    const int ty = y - S * Jy, tx = x - S * Jx;
    const double wy[2] = {(double)(S - ty) * (1.0 / S), (double)ty * (1.0 / S)};
    const double wx[2] = {(double)(S - tx) * (1.0 / S), (double)tx * (1.0 / S)};
    const int Jy1 = min(Jy + 1, gc.NY - 1), Jx1 = min(Jx + 1, gc.NX - 1);
    const long long cyx[4] = {Jy + (long long)gc.NY * Jx, Jy1 + (long long)gc.NY * Jx, Jy + (long long)gc.NY * Jx1,
                              Jy1 + (long long)gc.NY * Jx1};
    const int z0 = blockIdx.y * zc, z1 = min(gf.NZ, z0 + zc);
    for (int zb = z0; zb < z1; zb += kPZ) {
      bool solid[kPZ];
      bool any = false;
#pragma unroll
      for (int i = 0; i < kPZ; ++i) {
        solid[i] = zb + i < z1 && nmaskF[idx + (long long)plane * (zb + i)] != 0;
        any |= solid[i];
      }
      // void nodes hold 0 in every level vector (all of their writers leave
      // them 0): no load, no store, no interpolation
      if (!any) continue;
      TX cv[NL][4][3];  // coarse layers Jz0 .. Jz0 + NL - 1 (clamped; a clamped one has weight 0)
      const int Jz0 = zb / S;
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        const long long zo = planeC * min(Jz0 + l, gc.NZ - 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const long long J = cyx[c] + zo;
          cv[l][c][0] = xc[J];
          cv[l][c][1] = xc[NC + J];
          cv[l][c][2] = xc[2 * NC + J];
        }
      }
      TO ov[kPZ][3];
#pragma unroll
      for (int i = 0; i < kPZ; ++i) {
        if (!solid[i]) continue;
        const long long n = idx + (long long)plane * (zb + i);
#pragma unroll
        for (int k = 0; k < 3; ++k) ov[i][k] = out[k * NN + n];
      }
#pragma unroll
      for (int i = 0; i < kPZ; ++i) {
        if (!solid[i]) continue;
        const long long n = idx + (long long)plane * (zb + i);
        const int l0 = i / S, tz = i - S * l0;  // zb is a multiple of S: plane zb + i lies in layer Jz0 + i / S
        const double wz[2] = {(double)(S - tz) * (1.0 / S), (double)tz * (1.0 / S)};
        double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int a = c & 1, b = (c >> 1) & 1, d = c >> 2;
          const double w = wy[a] * wx[b] * wz[d];
          const int cc = a + 2 * b;
#pragma unroll
          for (int k = 0; k < 3; ++k) p[k] += w * (double)(d ? cv[l0 + 1][cc][k] : cv[l0][cc][k]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k * NN + n] = (TO)((double)ov[i][k] + p[k]);
      }
    }
  }
}

constexpr size_t kRestrictLds = 48 * 1024;  // LDS bytes of the y pass's rows at most

template <typename TI, typename TM, typename TO>
void restrict_impl(hipStream_t s, const Geo& gf, const Geo& gc, int span, const TI* rf, TM* tmp,
                   const uint8_t* nmaskC, TO* rc, const int* fl, Act actF) {
  const int plane = gc.NY * gf.NX;
  const dim3 g1((plane + kBlock - 1) / kBlock, gf.NZ);
  const dim3 g2((gc.NY * gc.NX + kBlock - 1) / kBlock, gc.NZ);
  // the LDS y pass: about one intermediate value per thread (R rows a block)
  // while the rows fit kRestrictLds bytes
  const int R = std::max(1, (kBlock + gc.NY / 2) / gc.NY);
  const size_t lds = (size_t)R * gf.NY * 3 * sizeof(TI);
  const long long rows = (long long)gf.NX * gf.NZ;
  const bool viaLds = lds <= kRestrictLds && (rows + R - 1) / R <= 0x7fffffffLL;
  const dim3 gl((unsigned)((rows + R - 1) / R));
  if (span == 4) {
    if (viaLds)
      hipLaunchKernelGGL((k_restrict_y_lds<4, TI, TM>), gl, dim3(kBlock), lds, s, gf, gc.NY, R, rf, tmp, fl, actF);
    else
      hipLaunchKernelGGL((k_restrict_y<4, TI, TM>), g1, dim3(kBlock), 0, s, gf, gc.NY, rf, tmp, fl, actF);
    hipLaunchKernelGGL((k_restrict_xz<4, TM, TO>), g2, dim3(kBlock), 0, s, gf, gc, tmp, nmaskC, rc, fl);
  } else {
    if (viaLds)
      hipLaunchKernelGGL((k_restrict_y_lds<2, TI, TM>), gl, dim3(kBlock), lds, s, gf, gc.NY, R, rf, tmp, fl, actF);
    else
      hipLaunchKernelGGL((k_restrict_y<2, TI, TM>), g1, dim3(kBlock), 0, s, gf, gc.NY, rf, tmp, fl, actF);
    hipLaunchKernelGGL((k_restrict_xz<2, TM, TO>), g2, dim3(kBlock), 0, s, gf, gc, tmp, nmaskC, rc, fl);
  }
}

// grid (y-x tiles, z chunks) and the planes per chunk (a multiple of kPZ);
// at most kMaxPartials blocks (a node plane beyond that is refused at set-up:
// prolong_plane_fits)
dim3 prolong_grid(const Geo& gf, int* zc) {
  const int bx = (gf.NY * gf.NX + kBlock - 1) / kBlock;
  int c = kPZ;
  while ((long long)bx * ((gf.NZ + c - 1) / c) > kMaxPartials && c < gf.NZ) c *= 2;
  *zc = c;
  return dim3(bx, (gf.NZ + c - 1) / c);
}

template <typename TX, typename TR, typename TD, typename TO>
void prolong_impl(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const TX* xc, const TR* rf,
                  const TD* diag, const uint8_t* nmaskF, TO* out, const RedWs* ws, double* sc, const int* fl) {
  int mode = 0;
  if (ws) mode = (fl && sc) ? 2 : 1;
  int zc = 0;
  const dim3 g = prolong_grid(gf, &zc);
  double* part = mode ? ws->partials : nullptr;
  if (span == 4)
    hipLaunchKernelGGL((k_prolong_jacobi<4, TX, TR, TD, TO>), g, dim3(kBlock), 0, s, gf, gc, omega, zc, xc, rf, diag,
                       nmaskF, out, part, fl);
  else
    hipLaunchKernelGGL((k_prolong_jacobi<2, TX, TR, TD, TO>), g, dim3(kBlock), 0, s, gf, gc, omega, zc, xc, rf, diag,
                       nmaskF, out, part, fl);
  if (mode) launch_fin1(s, mode == 1 ? FIN_ZR_INIT : FIN_ZR_LOOP, (int)(g.x * g.y), *ws, sc, (int*)fl, nullptr);
}

// x = omega * r / d on solid nodes (the 'Standard' cycle's pre-smoothing, 757)
template <typename TR, typename TD, typename TO>
__global__ __launch_bounds__(kBlock) void k_jacobi0(const TR* __restrict__ r, const TD* __restrict__ d,
                                                    TO* __restrict__ x, double omega, long long n, const int* fl,
                                                    Act act) {
  if (fl && fl[FL_DONE]) return;
  const long long m = act_items(act, n);
  const int nc = act_ncomp(act, n);
  for (long long g = (long long)blockIdx.x * kBlock + threadIdx.x; g < m; g += (long long)gridDim.x * kBlock) {
    const long long b = act_node(act, g);
    if (act.list && b >= act.NN) continue;
    for (int c = 0; c < nc; ++c) {
      const long long i = b + c * act.NN;
      const double dd = (double)d[i];
      x[i] = dd != 0.0 ? (TO)(omega * ((double)r[i] / dd)) : TO(0);
    }
  }
}

}  // namespace

long long restrict_tmp_len(const Geo& gf, const Geo& gc) { return 3LL * gc.NY * gf.NX * gf.NZ; }

void launch_restrict(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* rf, double* tmp,
                     const uint8_t* nmaskC, double* rc, const int* fl, Act actF) {
  restrict_impl<double, double, double>(s, gf, gc, span, rf, tmp, nmaskC, rc, fl, actF);
}
void launch_restrict(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* rf, float* tmp,
                     const uint8_t* nmaskC, float* rc, const int* fl, Act actF) {
  restrict_impl<double, float, float>(s, gf, gc, span, rf, tmp, nmaskC, rc, fl, actF);
}
void launch_restrict(hipStream_t s, const Geo& gf, const Geo& gc, int span, const float* rf, float* tmp,
                     const uint8_t* nmaskC, float* rc, const int* fl, Act actF) {
  restrict_impl<float, float, float>(s, gf, gc, span, rf, tmp, nmaskC, rc, fl, actF);
}

void launch_prolong_jacobi(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const double* xc,
                           const double* rf, const double* diag, const uint8_t* nmaskF, double* out,
                           const RedWs* ws, double* sc, const int* fl) {
  prolong_impl(s, gf, gc, span, omega, xc, rf, diag, nmaskF, out, ws, sc, fl);
}
void launch_prolong_jacobi(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const float* xc,
                           const float* rf, const float* diag, const uint8_t* nmaskF, float* out,
                           const RedWs* ws, double* sc, const int* fl) {
  prolong_impl(s, gf, gc, span, omega, xc, rf, diag, nmaskF, out, ws, sc, fl);
}
void launch_prolong_jacobi(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const float* xc,
                           const double* rf, const float* diag, const uint8_t* nmaskF, float* out,
                           const RedWs* ws, double* sc, const int* fl) {
  prolong_impl(s, gf, gc, span, omega, xc, rf, diag, nmaskF, out, ws, sc, fl);
}

template <typename T>
static void prolong_add_impl(hipStream_t s, const Geo& gf, const Geo& gc, int span, const T* xc, T* x,
                             const uint8_t* nmaskF, const int* fl) {
  int zc = 0;
  const dim3 g = prolong_grid(gf, &zc);
  if (span == 4)
    hipLaunchKernelGGL((k_prolong_add<4, T, T>), g, dim3(kBlock), 0, s, gf, gc, zc, xc, nmaskF, x, fl);
  else
    hipLaunchKernelGGL((k_prolong_add<2, T, T>), g, dim3(kBlock), 0, s, gf, gc, zc, xc, nmaskF, x, fl);
}
void launch_prolong_add(hipStream_t s, const Geo& gf, const Geo& gc, int span, const float* xc, float* x,
                        const uint8_t* nmaskF, const int* fl) {
  prolong_add_impl(s, gf, gc, span, xc, x, nmaskF, fl);
}
void launch_prolong_add(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* xc, double* x,
                        const uint8_t* nmaskF, const int* fl) {
  prolong_add_impl(s, gf, gc, span, xc, x, nmaskF, fl);
}

void launch_jacobi0(hipStream_t s, const double* r, const float* d, float* x, double omega, long long n,
                    const int* fl, Act act) {
  hipLaunchKernelGGL((k_jacobi0<double, float, float>), dim3(stream_blocks(act_items(act, n))), dim3(kBlock), 0, s, r, d, x, omega,
                     n, fl, act);
}
void launch_jacobi0(hipStream_t s, const double* r, const double* d, double* x, double omega, long long n,
                    const int* fl, Act act) {
  hipLaunchKernelGGL((k_jacobi0<double, double, double>), dim3(stream_blocks(act_items(act, n))), dim3(kBlock), 0, s, r, d, x,
                     omega, n, fl, act);
}
void launch_jacobi0(hipStream_t s, const float* r, const float* d, float* x, double omega, long long n,
                    const int* fl, Act act) {
  hipLaunchKernelGGL((k_jacobi0<float, float, float>), dim3(stream_blocks(act_items(act, n))), dim3(kBlock), 0, s, r, d, x, omega,
                     n, fl, act);
}

bool prolong_plane_fits(int NY, int NX) {
  return ((long long)NY * NX + kBlock - 1) / kBlock <= kMaxPartials;
}

}  // namespace t3x
