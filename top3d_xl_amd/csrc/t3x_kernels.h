// Kernel launchers of libt3x (all asynchronous on the given stream).
#pragma once

#include "t3x_internal.h"

namespace t3x {

// Scalar slots in device memory (double sc[] / int fl[]) for the solvers.
enum ScalarSlot {
  SC_NORMB = 0, SC_ZR, SC_PQ, SC_RR, SC_ZRN, SC_RES, SC_LAMBDA, SC_BETA,
  SC_TOL, SC_COUNT_PCG  // one PCG block = SC_COUNT_PCG doubles (<= 16)
};
enum FlagSlot { FL_DONE = 0, FL_ITS, FL_ITER, FL_COUNT };  // <= 8

// Geometry passed by value to kernels.  tiles / nTiles: the finest level's
// sum-factorised K*u work-group tiles that hold a solid cell (linear ids in
// the full tile grid, ascending), or null when every tile does; the K*u then
// launches one work-group per listed tile.
struct Geo {
  int ny, nx, nz;  // cells
  int NY, NX, NZ;  // nodes
  long long NN;
  const int* tiles;
  int nTiles;
};
inline Geo geo(const Level& L) {
  return Geo{L.ny, L.nx, L.nz, L.NY, L.NX, L.NZ, L.NN, L.sfTiles, L.nSfTiles};
}

// Node chunks of a finest-level vector that hold a solid node: chunk[n >> 4]
// for node n (16 nodes: one 128-byte line of a component in FP64), a nodal
// 3-vector's element i standing for node i mod NN.  The vector kernels skip
// void chunks (their entries are 0 and stay 0); chunk null: every chunk.
// list: the same chunks as ascending ids (nList of them), so that the vector
// kernels launch work only for them.
struct Act {
  const uint8_t* chunk = nullptr;
  long long NN = 0;
  const int* list = nullptr;
  int nList = 0;
};
__device__ __forceinline__ bool act_skip(const Act& a, long long i) {
  if (!a.chunk) return false;
  if (i >= a.NN) i -= a.NN;
  if (i >= a.NN) i -= a.NN;
  return !a.chunk[i >> 4];
}
// A vector kernel over n elements (n = NN or 3 NN when a.list is set) runs
// act_items work items; item g is node act_node(g) (at or past NN: none) and
// covers its act_ncomp components, element node + c NN.  Without a list,
// item g is element g.
__host__ __device__ __forceinline__ long long act_items(const Act& a, long long n) {
  return a.list ? 16LL * a.nList : n;
}
__host__ __device__ __forceinline__ int act_ncomp(const Act& a, long long n) { return a.list ? (int)(n / a.NN) : 1; }
__device__ __forceinline__ long long act_node(const Act& a, long long g) {
  return a.list ? 16LL * a.list[g >> 4] + (g & 15) : g;
}

// Reduction workspace: partials (kMaxRedBlocks * 8 doubles) + one counter.
struct RedWs {
  double* partials;
  unsigned int* counter;
};

// One-block finalisers: fold nb partials in a fixed order, then apply the
// solver's scalar step (CG step lengths, convergence flag).
enum FinOp { FIN_PQ = 0, FIN_ZR_INIT, FIN_ZR_LOOP, FIN_RR, FIN_SUM, FIN_SQRT };
void launch_fin1(hipStream_t s, int op, int nb, const RedWs& ws, double* sc, int* fl, double* out);

// ---------------------------------------------------------------- FEM
// sum-factorised K*u (t3x_applyk_sf.hip): the isotropic unit-brick Ke as
// 8 constants; partials: one per work-group (nullptr: none)
struct SfCoef {
  double k[8];
};
SfCoef sf_coefficients(double nu);
int applyk_sf_blocks(const Geo& g);
// Whether a level whose node grid has NY x NX nodes per plane fits the launch
// shapes of the kernels that write one reduction partial per work-group
// (at most kMaxPartials even with a single z chunk).
bool sf_plane_fits(int NY, int NX);
// The K*u tiles whose staged cells include a solid one (linear ids in the
// full tile grid of the sparse chunking, ascending); empty when all do.
// g.tiles is ignored.
std::vector<int> sf_active_tiles(const Geo& g, const std::vector<uint8_t>& emask);
bool filter_plane_fits(int NY, int NX);
bool prolong_plane_fits(int NY, int NX);
void launch_apply_K_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                       const double* u, double* y, double* partials, const int* fl);
void launch_fine_resid64_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                            const double* x, const double* r, double* d, const int* fl);
void launch_fine_smooth64_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                             const double* x, const double* r, const double* diag, double omega, double* out,
                             double* partials, const int* fl);
void launch_fine_resid32_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const float* E, const uint8_t* fix,
                            const float* x, const double* r, float* d, const int* fl);
void launch_fine_smooth32_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const float* E, const uint8_t* fix,
                             const float* x, const double* r, const float* diag, double omega, float* out,
                             double* partials, const int* fl);
// fine-level K*u (y = Ku, with the PCG's p.q reduction when pcgStep)
void launch_apply_K(hipStream_t s, const Geo& g, const SfCoef& cf, const double* E, const uint8_t* fix,
                    const double* u, double* y, const RedWs& ws, double* sc, const int* fl, bool pcgStep);
// unit compliance per cell, u_e^T Ke u_e (579-598), sum-factorised
void launch_compliance_sf(hipStream_t s, const Geo& g, const SfCoef& cf, const uint8_t* emask, const double* u,
                          double* ce);
void launch_diag_fine(hipStream_t s, const Geo& g, const double* Ke, const double* E, const uint8_t* fix,
                      const uint8_t* nmask, double* diag);
void launch_diag_coarse(hipStream_t s, const Geo& g, const double* Ks, const uint8_t* emask,
                        const uint8_t* nmask, double* diag);
void launch_galerkin_l2(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* T, const double* E,
                        const uint8_t* emaskC, const uint8_t* bcFlagC, double* Ks);
void launch_galerkin_l2_bc(hipStream_t s, const Geo& gf, const Geo& gc, int span, int nBc, const int* bcCells,
                           const double* Ke, const double* E, const uint8_t* emaskF, const uint8_t* fix,
                           double* Ks);
void launch_galerkin_coarse(hipStream_t s, const Geo& gf, const Geo& gc, const double* KsF, const uint8_t* emaskF,
                            const uint8_t* emaskC, double* KsC);

// ---------------------------------------------------------------- MG (t3x_mg.hip)
// Separable restriction through an intermediate of restrict_tmp_len(gf, gc)
// elements; prolongation + weighted Jacobi with optional fused z.r.
long long restrict_tmp_len(const Geo& gf, const Geo& gc);
// actF: the fine level's chunk flags (finest level only; the y pass reads no
// fine value from a void chunk).
void launch_restrict(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* rf, double* tmp,
                     const uint8_t* nmaskC, double* rc, const int* fl, Act actF = {});
void launch_restrict(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* rf, float* tmp,
                     const uint8_t* nmaskC, float* rc, const int* fl, Act actF = {});
void launch_restrict(hipStream_t s, const Geo& gf, const Geo& gc, int span, const float* rf, float* tmp,
                     const uint8_t* nmaskC, float* rc, const int* fl, Act actF = {});
void launch_prolong_jacobi(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const double* xc,
                           const double* rf, const double* diag, const uint8_t* nmaskF, double* out,
                           const RedWs* ws, double* sc, const int* fl);
void launch_prolong_jacobi(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const float* xc,
                           const float* rf, const float* diag, const uint8_t* nmaskF, float* out,
                           const RedWs* ws, double* sc, const int* fl);
void launch_prolong_jacobi(hipStream_t s, const Geo& gf, const Geo& gc, int span, double omega, const float* xc,
                           const double* rf, const float* diag, const uint8_t* nmaskF, float* out,
                           const RedWs* ws, double* sc, const int* fl);
// 'Standard' V-cycle pieces: x += P xc; x = omega r / d
void launch_prolong_add(hipStream_t s, const Geo& gf, const Geo& gc, int span, const float* xc, float* x,
                        const uint8_t* nmaskF, const int* fl);
void launch_prolong_add(hipStream_t s, const Geo& gf, const Geo& gc, int span, const double* xc, double* x,
                        const uint8_t* nmaskF, const int* fl);
void launch_jacobi0(hipStream_t s, const double* r, const double* d, double* x, double omega, long long n,
                    const int* fl, Act act = {});
void launch_jacobi0(hipStream_t s, const double* r, const float* d, float* x, double omega, long long n,
                    const int* fl, Act act = {});
void launch_jacobi0(hipStream_t s, const float* r, const float* d, float* x, double omega, long long n,
                    const int* fl, Act act = {});
// K*u smoothers of coarse levels with the stored Galerkin matrices (FP32 copy
// or the FP64 originals; t3x_applyk.hip).  resid: d = r - Kx; smooth:
// out = x + omega (r - Kx) / diag.  tmp: coarse_ks_tmp_len(g) elements of
// scratch (per-cell products).
long long coarse_ks_tmp_len(const Geo& g);
// emask, nmask: the level's cell and node masks (void cells read no matrix
// entry, void nodes no product).
void launch_coarse_resid32(hipStream_t s, const Geo& g, const float* Ks, const float* x, const float* r, float* d,
                           float* tmp, const int* fl, const uint8_t* emask, const uint8_t* nmask);
void launch_coarse_smooth32(hipStream_t s, const Geo& g, const float* Ks, const float* x, const float* r,
                            const float* diag, double omega, float* out, float* tmp, const int* fl,
                            const uint8_t* emask, const uint8_t* nmask);
// FP16 matrices with per-cell power-of-two scales (the products formed in FP32)
void launch_coarse_resid32(hipStream_t s, const Geo& g, const __half* Ks, const float* scale, const float* x,
                           const float* r, float* d, float* tmp, const int* fl, const uint8_t* emask,
                           const uint8_t* nmask);
void launch_coarse_smooth32(hipStream_t s, const Geo& g, const __half* Ks, const float* scale, const float* x,
                            const float* r, const float* diag, double omega, float* out, float* tmp, const int* fl,
                            const uint8_t* emask, const uint8_t* nmask);
// Ks (kTri x NN, FP64) -> FP16 entries Ks / scale, scale the power of two at
// or above the cell's largest |entry| (1 for a zero cell)
void launch_scale_ks_f16(hipStream_t s, const Geo& g, const double* Ks, __half* Ksh, float* scale);
void launch_coarse_resid64(hipStream_t s, const Geo& g, const double* Ks, const double* x, const double* r,
                           double* d, double* tmp, const int* fl, const uint8_t* emask, const uint8_t* nmask);
void launch_coarse_smooth64(hipStream_t s, const Geo& g, const double* Ks, const double* x, const double* r,
                            const double* diag, double omega, double* out, double* tmp, const int* fl,
                            const uint8_t* emask, const uint8_t* nmask);
void launch_coarse_gather(hipStream_t s, int n, const long long* map, const double* r, double* v, const int* fl);
void launch_coarse_gather(hipStream_t s, int n, const long long* map, const float* r, double* v, const int* fl);
void launch_coarse_scatter(hipStream_t s, int n, const long long* map, const double* v, double* x, long long NN3,
                           const int* fl);
void launch_coarse_scatter(hipStream_t s, int n, const long long* map, const double* v, float* x, long long NN3,
                           const int* fl);
void launch_band_assemble(hipStream_t s, const Geo& g, const double* Ks, const uint8_t* emask, const int* nodePos,
                          int bw, double* band, long long nBand);
int launch_band_factor(hipStream_t s, int n, int bw, double* band, int* status);
void launch_band_post(hipStream_t s, int n, int bw, int nb, const double* band, double* bandT, double* dinv);
void launch_band_solve(hipStream_t s, int n, int bw, int nb, const double* band, const double* bandT,
                       const double* dinv, double* v, const int* fl);

// ---------------------------------------------------------------- vectors
void launch_fill(hipStream_t s, double* p, double v, long long n);
// dst[0..n) = vals[0..n), n <= 8, as kernel arguments: host scalars reach the
// device in stream order without a host-to-device copy from host memory.
void launch_set_vals(hipStream_t s, double* dst, int n, const double* vals);
void launch_set_int(hipStream_t s, int* dst, int v);
void launch_copy(hipStream_t s, double* dst, const double* src, long long n);
void launch_sub(hipStream_t s, double* r, const double* b, const double* q, long long n);  // r = b - q
void launch_norm2(hipStream_t s, const double* v, long long n, const RedWs& ws, double* out);
void launch_dot(hipStream_t s, const double* a, const double* b, long long n, const RedWs& ws, double* out);
// r -= lambda q with r.r; with xj, also xj = omega r / dj (first Jacobi step of
// the FP32 'Standard' V-cycle).  y and p are updated later by launch_pcg_pupdate.
void launch_pcg_update(hipStream_t s, double* y, double* r, const double* p, const double* q, long long n,
                       const RedWs& ws, double* sc, int* fl, const float* dj = nullptr, float* xj = nullptr,
                       double omega = 0.0, Act act = {});
void launch_pcg_pupdate(hipStream_t s, double* y, double* p, const double* z, long long n, const double* sc,
                        const int* fl, Act act = {});
void launch_pcg_finish(hipStream_t s, double* y, const double* p, long long n, const double* sc, const int* fl,
                       Act act = {});
// mixed precision: FP32 z = M r, FP64 y, r, p, q
void launch_pcg_pupdate(hipStream_t s, double* y, double* p, const float* z, long long n, const double* sc,
                        const int* fl, Act act = {});
void launch_stream_copy(hipStream_t s, double* dst, const double* src, long long nDoubles);  // nDoubles even
void launch_convert(hipStream_t s, float* dst, const double* src, long long n);
void launch_convert(hipStream_t s, double* dst, const float* src, long long n);
void launch_pcg_init(hipStream_t s, double* sc, int* fl, double tol);
void launch_scale_diag(hipStream_t s, double* z, const double* dinv, const double* r, long long n,
                       const RedWs* ws, double* sc, const int* fl);
void launch_gather_compact(hipStream_t s, const int* map, int n, int ncomp, long long NN, const double* dense,
                           double* compact);
void launch_scatter_compact(hipStream_t s, const int* map, int n, int ncomp, long long NN, const double* compact,
                            double* dense);

// ---------------------------------------------------------------- filter
void launch_filter_e2n(hipStream_t s, const Geo& g, const uint8_t* emask, const double* src, double* b);
void launch_filter_apply(hipStream_t s, const Geo& g, const double* kern, const uint8_t* emask, const double* u,
                         double* y, const RedWs* ws, double* sc, const int* fl);
void launch_filter_n2e(hipStream_t s, const Geo& g, const uint8_t* emask, const double* u, double* tar);
void launch_filter_diag(hipStream_t s, const Geo& g, const double* kern, const uint8_t* emask,
                        const uint8_t* nmask, double* dinv);

// ---------------------------------------------------------------- optimiser
void launch_simp(hipStream_t s, long long n, const uint8_t* emask, const double* xPhys, double Emin, double E0,
                 double* E);
void launch_masked_sum(hipStream_t s, long long n, const uint8_t* emask, const double* a, const double* b,
                       int mode, const RedWs& ws, double* out);
void launch_masked_maxdiff(hipStream_t s, long long n, const uint8_t* emask, const double* a, const double* b,
                           const RedWs& ws, double* out);
void launch_oc_pass(hipStream_t s, long long n, const uint8_t* emask, const uint8_t* passive, const double* x,
                    const double* dc, const RedWs& ws, double* oc, int depth);
void launch_oc_final(hipStream_t s, long long n, const uint8_t* emask, const uint8_t* passive, double* x,
                     const double* dc, const RedWs& ws, double* oc);
void launch_set_where(hipStream_t s, long long n, const uint8_t* mask, double* a, double v);
void launch_elementwise(hipStream_t s, int op, long long n, const uint8_t* emask, double* a, const double* b,
                        const double* c, double p0, double p1, double p2);
enum { OP_DC = 0, OP_MUL, OP_DIVMAX, OP_HEAVI, OP_DHEAVI, OP_PNORMGRAD, OP_COPY, OP_SET, OP_DIVC,
       OP_MMA_BOUNDS, OP_MULC, OP_AXPYC };
enum MmaSlot {
  MM_LAM = 0, MM_MU, MM_EPSI, MM_ERR, MM_LOOP, MM_DONE, MM_B, MM_Y, MM_Z, MM_GRAD, MM_HESS, MM_S1, MM_S2,
  MM_TOL, MM_C, MM_A, MM_NEWTON, MM_GX, MM_COUNT
};
void launch_mma_gensub(hipStream_t s, long long n, const uint8_t* emask, int iter, const double* x,
                       const double* xmin, const double* xmax, double* xold1, double* xold2, double* low,
                       double* upp, const double* dfdx, const double* dgdx, double* alpha, double* beta,
                       double* p0, double* q0, double* pij, double* qij, const RedWs& ws, double* mm, double nAct);
void launch_mma_step(hipStream_t s, long long n, const uint8_t* emask, double* x, const double* low,
                     const double* upp, const double* alpha, const double* beta, const double* p0, const double* q0,
                     const double* pij, const double* qij, const RedWs& ws, double* mm);

}  // namespace t3x
