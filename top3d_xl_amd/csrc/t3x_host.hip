// Host side of libt3x: model set-up, the per-iteration drivers and the C ABI.
//
// Reference lines are TOP3D_XL.m.  There is no CPU compute path: every
// numerical step runs in the HIP kernels of t3x_kernels.hip; a missing GPU is
// an error (T3X_ERR_NO_DEVICE).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/t3x.h"
#include "t3x_coarse.h"
#include "t3x_pool.h"
#include "t3x_kernels.h"
#include "t3x_reduce.h"

using namespace t3x;

namespace {

double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Zeroed device allocation.  The fill is ordered on `stream` and waited for:
// a plain hipMemset runs on the null stream, which is not ordered with the
// non-blocking problem streams, and may still be in flight (and then land
// after the first kernel that writes the buffer) when it returns.
template <class T>
T* dalloc(size_t n, hipStream_t stream) {
  void* p = nullptr;
  T3X_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  T3X_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream));
  T3X_HIP(hipStreamSynchronize(stream));
  return (T*)p;
}

// ------------------------------------------------------------ element stiffness
// Ke of the unit trilinear brick for E = 1 from 2x2x2 Gauss quadrature of
// B^T D B (exact for a brick), nodes in the reference's local order at
// physical (x, y, z) = (dx, -dy, dz).  Equals the closed form of
// FEA_VoxelBasedElementStiffnessMatrix (1254-1292) to rounding.
std::vector<double> gauss_ke(double nu, double cell) {
  const double lam = nu / ((1 + nu) * (1 - 2 * nu)), mu = 1.0 / (2 * (1 + nu));
  double D[6][6] = {};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D[i][j] = lam;
  for (int i = 0; i < 3; ++i) D[i][i] += 2 * mu;
  for (int i = 3; i < 6; ++i) D[i][i] = mu;
  double nat[8][3];
  for (int a = 0; a < 8; ++a) {
    nat[a][0] = kLocDx[a] ? 1.0 : -1.0;   // x
    nat[a][1] = kLocDy[a] ? -1.0 : 1.0;   // y = -row
    nat[a][2] = kLocDz[a] ? 1.0 : -1.0;   // z
  }
  std::vector<double> K(576, 0.0);
  const double g = 1.0 / std::sqrt(3.0);
  for (int q = 0; q < 8; ++q) {
    const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g, zeta = (q & 4) ? g : -g;
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
      const double s = nat[a][0], t = nat[a][1], p = nat[a][2];
      // d/dx = 2 d/dxi on the unit cube
      const double bx = 2 * 0.125 * s * (1 + t * eta) * (1 + p * zeta);
      const double by = 2 * 0.125 * t * (1 + s * xi) * (1 + p * zeta);
      const double bz = 2 * 0.125 * p * (1 + s * xi) * (1 + t * eta);
      B[0][3 * a] = bx; B[1][3 * a + 1] = by; B[2][3 * a + 2] = bz;
      B[3][3 * a + 1] = bz; B[3][3 * a + 2] = by;
      B[4][3 * a] = bz; B[4][3 * a + 2] = bx;
      B[5][3 * a] = by; B[5][3 * a + 1] = bx;
    }
    const double w = 0.125;  // det J
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k)
          for (int l = 0; l < 6; ++l) s += B[k][i] * D[k][l] * B[l][j];
        K[j * 24 + i] += s * w;
      }
  }
  for (int i = 0; i < 24; ++i)  // exact symmetry
    for (int j = 0; j < i; ++j) {
      const double v = 0.5 * (K[j * 24 + i] + K[i * 24 + j]);
      K[j * 24 + i] = K[i * 24 + j] = v * cell;
    }
  for (int i = 0; i < 24; ++i) K[i * 25] *= cell;
  return K;
}

// Interpolation weight of coarse local node J at child-local node a of child
// (ly,lx,lz) of a span-s cell.
double hat_weight(int span, int ly, int lx, int lz, int a, int J) {
  const int ty = ly + kLocDy[a], tx = lx + kLocDx[a], tz = lz + kLocDz[a];
  const double wy = kLocDy[J] ? (double)ty / span : (double)(span - ty) / span;
  const double wx = kLocDx[J] ? (double)tx / span : (double)(span - tx) / span;
  const double wz = kLocDz[J] ? (double)tz / span : (double)(span - tz) / span;
  return wy * wx * wz;
}

// T_ss = Q_ss^T Ke Q_ss (upper triangle) for the span^3 children (950-954).
std::vector<double> l2_templates(const std::vector<double>& Ke, int span) {
  const int s3 = span * span * span;
  std::vector<double> T((size_t)s3 * kTri);
  for (int ss = 0; ss < s3; ++ss) {
    const int ly = ss % span, lx = (ss / span) % span, lz = ss / (span * span);
    double Q[24][24] = {};
    for (int a = 0; a < 8; ++a)
      for (int J = 0; J < 8; ++J) {
        const double w = hat_weight(span, ly, lx, lz, a, J);
        for (int c = 0; c < 3; ++c) Q[3 * a + c][3 * J + c] = w;
      }
    double KQ[24][24];
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0.0;
        for (int k = 0; k < 24; ++k) s += Ke[k * 24 + i] * Q[k][j];
        KQ[i][j] = s;
      }
    for (int q = 0; q < 24; ++q)
      for (int p = 0; p <= q; ++p) {
        double s = 0.0;
        for (int k = 0; k < 24; ++k) s += Q[k][p] * KQ[k][q];
        T[(size_t)ss * kTri + tri_index(p, q)] = s;
      }
  }
  return T;
}

}  // namespace

// ============================================================== the problem
struct OptState {
  bool active = false;
  int mode = 0, nLoop = 0, loop = 0, loopbeta = 0;
  double V0 = 0.0, rMin = 0.0, rHat = 0.0, cSolid = 0.0;
  double change = 1.0, sharp = 1.0, beta = 1.0;  // 85-86, 233, 284-285
  std::vector<t3x_iter> hist;
  std::vector<int> lssIts;      // one entry per elasticity solve (lssIts(end+1), 121)
  std::vector<double> resAll;   // final relative residual of each solve
};

// HIP-event timing of kernel groups on the problem's stream.
enum TimedGroup { TG_APPLYK = 0, TG_VCYCLE, TG_COARSE, TG_ASSEMBLE, TG_COUNT };
struct GroupTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> pool;
  double ms = 0.0;
  long long count = 0;
};

// parts per million of the device's work-group slots held by this process's
// handles for the one-launch coarsest recurrences
static std::atomic<long long> gChainShare{0};

struct t3x_problem {
  // Caller-provided resources (the serving engine's pooled slot): device
  // allocations are carved in order out of [arena, arena + arenaSize) instead
  // of hipMalloc'ed, the stream and the pinned read-back buffer are lent, and
  // destroying the problem releases none of them.
  char* arena = nullptr;
  size_t arenaSize = 0, arenaUsed = 0;
  bool ownStream = true, ownHostBuf = true;
  // coarsest solve: one launch per recurrence (k_solve_chain) where allowed
  // (its groups must all be resident together: see gChainShare)
  bool oneLaunchCoarse = true;
  long long chainShare = 0;  // its part (per million) of the device slots held for one-launch solves
  OptState opt;
  GroupTimer timers[TG_COUNT];
  t3x_params P{};
  int nely = 0, nelx = 0, nelz = 0;
  int L = 0;
  std::vector<Level> lv;
  hipStream_t stream = nullptr;
  std::string err;
  int warnings = 0;
  // finest-level device state
  double* dKe = nullptr;
  SfCoef sf{};  // Ke of the unit brick as the constants of the sum-factorised K*u
  double* dT2 = nullptr;
  uint8_t* fix = nullptr;
  uint8_t* act = nullptr;  // finest node chunks (16 nodes) with a solid node, or null when all are
  int* actList = nullptr;  // ids of those chunks (ascending), nActList of them
  int nActList = 0;
  Act fineAct() const { return Act{act, lv[0].NN, actList, nActList}; }
  uint8_t* passive = nullptr;  // passive cells (passiveElements_), or null
  uint8_t* active = nullptr;   // solid and not passive (MMA design variables, 274-276)
  int nPassive = 0;
  double* bvec = nullptr;
  double *F = nullptr, *U = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *tmp3 = nullptr;
  // load cases (loadingCond_, 481-493): F_ columns and their solutions; F/U
  // above are case 0; objective weights objWeightingList_ (493, 558)
  int nCases = 1;
  std::vector<double*> Fc, Uc;
  std::vector<double> wObj;
  double* ceTmp = nullptr;  // one case's element compliance (nCases > 1)
  double *E = nullptr, *x = nullptr, *xPhys = nullptr, *ce = nullptr, *dc = nullptr;
  // mixed precision (precision == 1): FP32 V-cycle output; the 'Standard'
  // cycle adds FP32 E and the fine-level smoother vectors
  float* z32 = nullptr;
  float *E32 = nullptr, *x32 = nullptr, *d32 = nullptr;
  // FP64 'Standard' cycle: fine-level smoother state and defect
  double *x64 = nullptr, *d64 = nullptr;
  void* rtmp = nullptr;  // restriction intermediate / coarse smoother products (largest use)
  double *ca = nullptr, *cb = nullptr, *cc = nullptr, *cd = nullptr;  // cell scratch
  // filters (node scalars)
  double *fb = nullptr, *fx = nullptr, *fr = nullptr, *fz = nullptr, *fp = nullptr, *fq = nullptr;
  Filter filt[2];
  double* dFiltKern = nullptr;  // 2*64
  // L2 cells with a BC child
  int nBc2 = 0;
  int* dBc2 = nullptr;
  uint8_t* dBcFlag2 = nullptr;
  // coarsest band
  int cN = 0, cBW = 0;
  CoarseSolver cs;
  long long* dCMap = nullptr;
  int* dNodePos = nullptr;
  // the factorisation of the coarsest matrix: band (factored in place),
  // scratch, and the solve operators (Dinv, DT, G, H)
  struct CoarseFactor {
    double *band = nullptr, *tinv = nullptr, *pan = nullptr, *Dinv = nullptr, *DT = nullptr, *G = nullptr,
           *H = nullptr;
    int* status = nullptr;
  };
  CoarseFactor cf;
  const CoarseFactor& fac() const { return cf; }
  double *cva = nullptr, *cvb = nullptr;
  ChainState* chainSt = nullptr;
  bool assembled = false;
  // compact <-> dense maps (finest)
  int nNodes = 0, nCells = 0;
  int *dNodeMap = nullptr, *dCellMap = nullptr;
  std::vector<int> hNodeMap, hCellMap;
  // host copies for the passive-layer generator: nodes left in the fixings,
  // nodes loaded in any case, the passive cells (dense node/cell indexing)
  std::vector<uint8_t> hFixNode, hLoaded, hPassive;
  // reductions / scalars
  RedWs ws{nullptr, nullptr};
  double* sc = nullptr;   // [0..15] elastic PCG, [16..31] filter PCG, [32..] misc
  int* fl = nullptr;      // [0..7] elastic, [8..15] filter
  double* oc = nullptr;   // 8
  double* mm = nullptr;   // MM_COUNT
  double* hbuf = nullptr; // pinned read-back
  // MMA state (dense cells)
  double *mx = nullptr, *mxold1 = nullptr, *mxold2 = nullptr, *mlow = nullptr, *mupp = nullptr, *malpha = nullptr,
         *mbeta = nullptr, *mp0 = nullptr, *mq0 = nullptr, *mpij = nullptr, *mqij = nullptr, *mxmin = nullptr,
         *mxmax = nullptr;
  std::vector<void*> allocs;
  // one PCG iteration captured as a HIP graph, per solution vector y (the
  // kernels' pointers are fixed for a problem; only y differs between solves)
  std::vector<std::pair<const double*, hipGraphExec_t>> pcgGraphs;

  long long bytes = 0;  // device bytes allocated by this problem
  template <class T>
  T* alloc(size_t n) {
    const size_t nb = std::max<size_t>(n, 1) * sizeof(T);
    if (arena) {
      const size_t off = (arenaUsed + 255) & ~(size_t)255;
      if (off + nb > arenaSize) throw HipError("device memory slot exhausted (footprint estimate exceeded)");
      T* p = (T*)(arena + off);
      T3X_HIP(hipMemsetAsync(p, 0, nb, stream));
      T3X_HIP(hipStreamSynchronize(stream));
      arenaUsed = off + nb;
      bytes += (long long)nb;
      return p;
    }
    T* p = dalloc<T>(n, stream);
    allocs.push_back(p);
    bytes += (long long)nb;
    return p;
  }
  hipEvent_t ev() {
    hipEvent_t e;
    T3X_HIP(hipEventCreate(&e));
    return e;
  }
  // begin/end of a timed launch group (no-op unless params.timing)
  // timing 1: every group; 2: only the fine K*p (each event pair is a
  // barrier packet on the stream, ~1% of a design iteration for all groups)
  bool timed(int g) const { return P.timing == 1 || (P.timing == 2 && g == TG_APPLYK); }
  void tbegin(int g, hipEvent_t* e) {
    if (!timed(g)) return;
    GroupTimer& T = timers[g];
    if (T.pool.size() < 2) { T.pool.push_back(ev()); T.pool.push_back(ev()); }
    e[0] = T.pool.back(); T.pool.pop_back();
    e[1] = T.pool.back(); T.pool.pop_back();
    T3X_HIP(hipEventRecord(e[0], stream));
  }
  void tend(int g, hipEvent_t* e) {
    if (!timed(g)) return;
    GroupTimer& T = timers[g];
    T3X_HIP(hipEventRecord(e[1], stream));
    T.pending.push_back({e[0], e[1]});
    if (T.pending.size() > 4096) tflush(g);
  }
  void tflush(int g) {
    GroupTimer& T = timers[g];
    for (auto& pr : T.pending) {
      T3X_HIP(hipEventSynchronize(pr.second));
      float ms = 0.f;
      T3X_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      T.ms += ms;
      T.count += 1;
      T.pool.push_back(pr.first);
      T.pool.push_back(pr.second);
    }
    T.pending.clear();
  }
  ~t3x_problem() {
    if (chainShare) gChainShare.fetch_sub(chainShare);
    for (auto& g : pcgGraphs) (void)hipGraphExecDestroy(g.second);
    for (auto& T : timers) {
      for (auto& pr : T.pending) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
      for (auto e : T.pool) (void)hipEventDestroy(e);
    }
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocs) (void)hipFree(p);
    if (hbuf && ownHostBuf) (void)hipHostFree(hbuf);
    if (stream && ownStream) (void)hipStreamDestroy(stream);
  }
  const Level& f() const { return lv[0]; }
  double* xTilde() { return ca; }   // LOCAL-mode state between iterations
  double* volMax() { return cb; }
  long long NN() const { return lv[0].NN; }
  bool mixed() const { return P.precision >= 1; }  // FP32 V-cycle
  bool f16() const { return P.precision == 2; }    // ... with FP16 coarse Galerkin copies
  const uint8_t* design_mask() const { return active ? active : lv[0].emask; }
  bool standard() const { return P.vcycle == 1; }  // 'Standard' V-cycle (FP64 or, mixed, FP32)
  void sync() { T3X_HIP(hipStreamSynchronize(stream)); }
  // Host-to-device upload ordered on the problem stream (a plain hipMemcpy from
  // pageable memory may return before its DMA lands and is not ordered with a
  // non-blocking stream), waited for so the host buffer may be released.
  void upload(void* dst, const void* src, size_t bytes) {
    T3X_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    sync();
  }
};

namespace {

// --------------------------------------------------------- hierarchy set-up
// Level count and padding (1651-1671), non-dyadic merge (1085-1096):
// per level (ny, nx, nz, span) of the padded grid.
struct LevelDims {
  int ny, nx, nz, span;
};
// coarsestControl <= 0 (not the reference's: an option of this library):
// the reference's count for 50000, then one level more while the coarsest
// box's band Cholesky, n bw^2 with n = 3 nodes and bw = 3 (1 + c0 + c0 c1)
// over its sorted node extents c0 <= c1, exceeds kAutoCoarseCost and the
// dimensions divide (the fine grid keeps its padding).  Cuboid
// 256x512x256: 2.4e10, unchanged; 192x384x192: 3.5e11 -> 3.8e9.
constexpr double kAutoCoarseCost = 1e11;
std::vector<LevelDims> level_dims(int nely, int nelx, int nelz, long long nv, int coarsestControl) {
  const bool autoDepth = coarsestControl <= 0;
  if (autoDepth) coarsestControl = 50000;
  int cnt = 0;
  while (nv >= coarsestControl) {
    ++cnt;
    nv = (long long)std::llround((double)nv / 8.0);
  }
  cnt = std::max(3, cnt);
  if (autoDepth) {
    auto band_cost = [&](int c) {  // coarsest box of cnt = c: cells fine / 2^c per axis
      const int pw = 1 << c;
      long long e[3] = {(nely + pw - 1) / pw + 1LL, (nelx + pw - 1) / pw + 1LL, (nelz + pw - 1) / pw + 1LL};
      std::sort(e, e + 3);
      const double n = 3.0 * e[0] * e[1] * e[2], bw = 3.0 * (1 + e[0] + e[0] * e[1]);
      return n * bw * bw;
    };
    auto divides = [&](int c) {
      const int pw = 1 << c;
      return nely % pw == 0 && nelx % pw == 0 && nelz % pw == 0 && std::min({nely, nelx, nelz}) / pw >= 2;
    };
    while (band_cost(cnt) > kAutoCoarseCost && divides(cnt) && divides(cnt + 1)) ++cnt;
  }
  const int pw = 1 << cnt;
  auto padded = [&](int n) { return (n + pw - 1) / pw * pw; };
  // (cnt + 1) levels, then one merged by the span-4 first coarsening
  std::vector<LevelDims> d(cnt);
  d[0] = {padded(nely), padded(nelx), padded(nelz), 0};
  for (int l = 1; l < cnt; ++l) {
    const int s = (l == 1) ? 4 : 2;
    d[l] = {d[l - 1].ny / s, d[l - 1].nx / s, d[l - 1].nz / s, s};
  }
  return d;
}

// The fine node plane (NY x NX nodes; coarser planes are smaller) must fit the
// launch shapes that write one reduction partial per work-group.
bool plane_fits(int NY, int NX) {
  return sf_plane_fits(NY, NX) && filter_plane_fits(NY, NX) && prolong_plane_fits(NY, NX);
}
const char* kPlaneTooLarge =
    "model too large: the padded (nely+1) x (nelx+1) node plane exceeds the launch-shape limit (about 1.4e7 nodes)";

void build_levels(t3x_problem& P, const uint8_t* vox) {
  long long nv = 0;
  const long long nvox = (long long)P.nely * P.nelx * P.nelz;
  for (long long i = 0; i < nvox; ++i) nv += vox[i] ? 1 : 0;
  const std::vector<LevelDims> dims = level_dims(P.nely, P.nelx, P.nelz, nv, P.P.coarsestControl);
  P.L = (int)dims.size();
  P.lv.assign(P.L, Level());
  for (int l = 0; l < P.L; ++l) {
    Level& V = P.lv[l];
    V.ny = dims[l].ny; V.nx = dims[l].nx; V.nz = dims[l].nz; V.span = dims[l].span;
    V.NY = V.ny + 1; V.NX = V.nx + 1; V.NZ = V.nz + 1;
    V.NN = (long long)V.NY * V.NX * V.NZ;
    V.hEmask.assign(V.NN, 0);
    V.hNmask.assign(V.NN, 0);
  }
  Level& F = P.lv[0];
  // finest cells
  for (int iz = 0; iz < P.nelz; ++iz)
    for (int ix = 0; ix < P.nelx; ++ix)
      for (int iy = 0; iy < P.nely; ++iy)
        if (vox[iy + (long long)P.nely * (ix + (long long)P.nelx * iz)])
          F.hEmask[iy + (long long)F.NY * (ix + (long long)F.NX * iz)] = 1;
  // coarse cells: non-empty if any child is (1152)
  for (int l = 1; l < P.L; ++l) {
    Level& C = P.lv[l];
    const Level& Fn = P.lv[l - 1];
    const int s = C.span;
    for (int iz = 0; iz < Fn.nz; ++iz)
      for (int ix = 0; ix < Fn.nx; ++ix)
        for (int iy = 0; iy < Fn.ny; ++iy)
          if (Fn.hEmask[iy + (long long)Fn.NY * (ix + (long long)Fn.NX * iz)])
            C.hEmask[iy / s + (long long)C.NY * (ix / s + (long long)C.NX * (iz / s))] = 1;
  }
  for (int l = 0; l < P.L; ++l) {
    Level& V = P.lv[l];
    V.nCells = 0;
    for (int iz = 0; iz < V.nz; ++iz)
      for (int ix = 0; ix < V.nx; ++ix)
        for (int iy = 0; iy < V.ny; ++iy) {
          const long long c = iy + (long long)V.NY * (ix + (long long)V.NX * iz);
          if (!V.hEmask[c]) continue;
          ++V.nCells;
          for (int j = 0; j < 8; ++j)
            V.hNmask[(iy + kLocDy[j]) + (long long)V.NY * ((ix + kLocDx[j]) + (long long)V.NX * (iz + kLocDz[j]))] = 1;
        }
    V.nNodes = 0;
    for (long long n = 0; n < V.NN; ++n) V.nNodes += V.hNmask[n];
  }
}

// Upper bound of the device bytes setup() and a whole run allocate: the same
// per-level arrays, the coarsest band with the bandwidth of a full box in
// its best axis order (a sub-domain's ordering is never wider), and 4 KiB of
// allocation slack per array.
long long footprint_estimate(int nely, int nelx, int nelz, long long nv, const t3x_params& prm, int mode,
                             int nCases = 1) {
  const std::vector<LevelDims> d = level_dims(nely, nelx, nelz, nv, prm.coarsestControl);
  const int L = (int)d.size();
  const bool mixed = prm.precision >= 1;
  auto NNof = [&](int l) { return (long long)(d[l].ny + 1) * (d[l].nx + 1) * (d[l].nz + 1); };
  long long b = 0, nalloc = 0;
  auto add = [&](long long bytes, int count = 1) { b += bytes * count; nalloc += count; };
  long long tmp = 0;
  for (int l = 0; l < L; ++l) {
    const long long NN = NNof(l);
    add(NN, 2);                                      // emask, nmask
    if (l < L - 1) add(24 * NN);                     // diag
    if (l >= 1) add(24 * NN, 2), add(8LL * kTri * NN);  // r, x, Ks
    if (!mixed && prm.vcycle == 1 && l >= 1 && l < L - 1) add(24 * NN);  // t
    if (mixed) {
      if (l < L - 1) add(12 * NN);                   // diagf
      if (l >= 1) add(12 * NN, 2);                   // rf, xf
      if (prm.vcycle == 1 && l >= 1 && l < L - 1)  // tf, Ksf (or Ksh and its scales)
        add(12 * NN), add((prm.precision == 2 ? 2LL : 4LL) * kTri * NN + (prm.precision == 2 ? 4 * NN : 0));
    }
    if (l + 1 < L) tmp = std::max(tmp, 3LL * (d[l + 1].ny + 1) * (d[l].nx + 1) * (d[l].nz + 1));
    if (prm.vcycle == 1 && l >= 1 && l < L - 1) tmp = std::max(tmp, (mixed ? 12 : 24) * NN);  // coarse smoother products
  }
  add(8 * tmp);                                      // restriction intermediate
  const long long NN = NNof(0);
  add(NN);                                           // fix
  add(24 * NN, 8);                                   // F, U, r, z, p, q, tmp3, bvec
  if (nCases > 1) add(24 * NN, 2 * (nCases - 1)), add(8 * NN);  // further F_ columns and solutions, ceTmp
  add(8 * NN, 9);                                    // cell arrays
  add(8 * NN, 6);                                    // filter vectors
  add(8 * NN, 2);                                    // filter diagonals
  if (mixed) add(12 * NN);                           // z32
  if (prm.vcycle == 1 && mixed) add(4 * NN), add(12 * NN, 2);                // E32, x32, d32
  if (prm.vcycle == 1 && !mixed) add(24 * NN, 2);                            // x64, d64
  if (mode == 1) add(8 * NN, 13);                    // MMA state
  add(576 * 8);                                      // Ke
  add(64LL * kTri * 8);                              // level-2 templates
  add(4 * NNof(1) + NNof(1));                        // BC cell list + flags
  add(8 * NN, 1);                                    // compact node + cell maps (2 x int32)
  // coarsest band
  int c[3] = {d[L - 1].ny + 1, d[L - 1].nx + 1, d[L - 1].nz + 1};
  std::sort(c, c + 3);
  const long long nC = 3LL * c[0] * c[1] * c[2];
  const long long bw = 3LL * (1 + c[0] + (long long)c[0] * c[1]) + 2;
  const long long SB = std::max<long long>(64, (bw + 1 + 63) / 64 * 64);
  const long long np = (nC + SB - 1) / SB * SB;
  add(8 * nC);                                       // coarsest map
  add(4 * NNof(L - 1));                              // node positions
  add(8 * np * (bw + 1));                            // band
  add(8 * np * 64);                                  // tile inverses
  add(8 * 2 * ((bw + 63) / 64 + 1) * 64 * 64);       // factorization panels
  add(8 * (np / SB) * SB * SB, 4);                   // Dinv, DT, G, H
  add(8 * 3 * np, 2);                                // solve vectors (with the tagged words)
  add((long long)sizeof(ChainState));                 // one-launch recurrences' words
  add(8LL * std::max(kMaxRedBlocks * 8, kMaxPartials) + 4096 * 4);  // reductions, scalars, flags
  add(4LL * kMaxPartials);                           // K*u tile list (irregular domains)
  add(NN / 16 + 1);                                  // node-chunk flags (irregular domains)
  add(4 * (NN / 16 + 1));                            // and their id list
  return b + 4096 * nalloc;
}

int setup(t3x_problem& P, const uint8_t* vox, int nFix, const int* fixNodes, const uint8_t* fixState, int nCases,
          const int* nPerCase, const int* loadNodes, const double* loadVec) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    P.err = "no HIP device visible: libt3x has no CPU path";
    return T3X_ERR_NO_DEVICE;
  }
  build_levels(P, vox);
  Level& F = P.lv[0];
  const long long NN = F.NN;
  if (3 * NN >= (1LL << 31)) {  // the fine-level kernels index DOFs with 32-bit offsets
    P.err = "model too large: 3 x (nely+1)(nelx+1)(nelz+1) must stay below 2^31";
    return T3X_ERR_INPUT;
  }
  if (!plane_fits(F.NY, F.NX)) {
    P.err = kPlaneTooLarge;
    return T3X_ERR_INPUT;
  }
  // ---- boundary conditions (536-557 or explicit), 1294-1336
  std::vector<long long> fixD;
  std::vector<std::vector<long long>> loadD;
  std::vector<uint8_t> fs;
  std::vector<std::vector<double>> lv;
  if (nFix < 0) {
    loadD.resize(1);
    lv.resize(1);
    for (int jz = 0; jz <= P.nelz; ++jz)
      for (int jy = 0; jy <= P.nely; ++jy) fixD.push_back(jy + (long long)F.NY * (0 + (long long)F.NX * jz));
    fs.assign(fixD.size() * 3, 1);
    const int zTop = (int)std::llround(P.nelz / 6.0 + 1.0);
    for (int jz = 0; jz < zTop; ++jz)
      for (int jy = 0; jy <= P.nely; ++jy)
        loadD[0].push_back(jy + (long long)F.NY * (P.nelx + (long long)F.NX * jz));
    const double qv = 1.0 / (double)loadD[0].size();
    for (size_t i = 0; i < loadD[0].size(); ++i) {
      lv[0].push_back(0.0 * qv); lv[0].push_back(0.0 * qv); lv[0].push_back(-1.0 * qv);
    }
    P.wObj.assign(1, 1.0);  // 558
  } else {
    const long long uy = P.nely + 1, ux = P.nelx + 1;
    auto adapt = [&](int id) {
      const long long jy = id % uy, jx = (id / uy) % ux, jz = id / uy / ux;
      return jy + (long long)F.NY * (jx + (long long)F.NX * jz);
    };
    for (int i = 0; i < nFix; ++i) {
      fixD.push_back(adapt(fixNodes[i]));
      for (int c = 0; c < 3; ++c) fs.push_back(fixState[3 * i + c]);
    }
    loadD.resize(nCases);
    lv.resize(nCases);
    long long off = 0;
    for (int ii = 0; ii < nCases; ++ii)
      for (int i = 0; i < nPerCase[ii]; ++i, ++off) {
        loadD[ii].push_back(adapt(loadNodes[off]));
        for (int c = 0; c < 3; ++c) lv[ii].push_back(loadVec[3 * off + c]);
      }
    P.wObj.assign(nCases, 1.0 / nCases);  // 493
  }
  P.nCases = (int)loadD.size();
  std::vector<uint8_t> hfix(NN, 0);
  std::vector<std::vector<double>> hF(P.nCases);
  std::vector<uint8_t> loaded(NN, 0);  // loaded in any case: released from the fixings (1301-1305)
  int nF = 0;
  for (int ii = 0; ii < P.nCases; ++ii) {
    // a node listed twice keeps its first entry (unique(), 499-518)
    std::vector<uint8_t> seen(NN, 0);
    hF[ii].assign(3 * NN, 0.0);
    int nL = 0;
    for (size_t i = 0; i < loadD[ii].size(); ++i) {
      const long long d = loadD[ii][i];
      if (d < 0 || d >= NN || !F.hNmask[d] || seen[d]) continue;
      seen[d] = 1;
      loaded[d] = 1;
      ++nL;
      for (int c = 0; c < 3; ++c) hF[ii][c * NN + d] = lv[ii][3 * i + c];
    }
    if (nL == 0) { P.err = "No Loads!"; return T3X_ERR_NO_LOADS; }  // 1309
  }
  std::vector<uint8_t> fixSeen(NN, 0);
  P.hFixNode.assign(NN, 0);
  for (size_t i = 0; i < fixD.size(); ++i) {
    const long long d = fixD[i];
    if (d < 0 || d >= NN || !F.hNmask[d] || fixSeen[d]) continue;
    fixSeen[d] = 1;
    if (loaded[d]) continue;  // loaded nodes are released (1303)
    P.hFixNode[d] = 1;
    uint8_t m = 0;
    for (int c = 0; c < 3; ++c) m |= (fs[3 * i + c] ? 1 : 0) << c;
    hfix[d] = m;
  }
  for (long long d = 0; d < NN; ++d) nF += hfix[d] ? 1 : 0;
  if (nF == 0) { P.err = "No Fixations!"; return T3X_ERR_NO_FIX; }
  P.hLoaded = loaded;

  if (!P.stream) T3X_HIP(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
  // ---- device arrays per level
  for (int l = 0; l < P.L; ++l) {
    Level& V = P.lv[l];
    V.emask = P.alloc<uint8_t>(V.NN);
    V.nmask = P.alloc<uint8_t>(V.NN);
    P.upload(V.emask, V.hEmask.data(), V.NN);
    P.upload(V.nmask, V.hNmask.data(), V.NN);
    if (l < P.L - 1) V.diag = P.alloc<double>(3 * V.NN);
    if (l >= 1) {
      V.r = P.alloc<double>(3 * V.NN);
      V.x = P.alloc<double>(3 * V.NN);
      V.Ks = P.alloc<double>((size_t)kTri * V.NN);
    }
    if (P.standard() && !P.mixed() && l >= 1 && l < P.L - 1) V.t = P.alloc<double>(3 * V.NN);
    if (l == 0) {  // K*u work-groups only for the tiles that hold a solid cell
      const std::vector<int> tl = sf_active_tiles(geo(V), V.hEmask);
      if (!tl.empty()) {
        V.sfTiles = P.alloc<int>(tl.size());
        V.nSfTiles = (int)tl.size();
        P.upload(V.sfTiles, tl.data(), tl.size() * sizeof(int));
      }
      // vector kernels skip 16-node chunks without a solid node
      std::vector<uint8_t> ch((size_t)((V.NN + 15) / 16), 0);
      bool all = true;
      for (long long n = 0; n < V.NN; ++n)
        if (V.hNmask[n]) ch[n >> 4] = 1;
      for (uint8_t c : ch) all = all && c;
      if (!all) {
        P.act = P.alloc<uint8_t>(ch.size());
        P.upload(P.act, ch.data(), ch.size());
        std::vector<int> ids;
        for (size_t c = 0; c < ch.size(); ++c)
          if (ch[c]) ids.push_back((int)c);
        P.nActList = (int)ids.size();
        P.actList = P.alloc<int>(ids.size());
        if (!ids.empty()) P.upload(P.actList, ids.data(), ids.size() * sizeof(int));
      }
    }
    if (P.mixed()) {
      if (l < P.L - 1) V.diagf = P.alloc<float>(3 * V.NN);
      if (l >= 1) {
        V.rf = P.alloc<float>(3 * V.NN);
        V.xf = P.alloc<float>(3 * V.NN);
      }
      if (P.standard() && l >= 1 && l < P.L - 1) {
        V.tf = P.alloc<float>(3 * V.NN);
        if (P.f16()) {
          V.Ksh = P.alloc<__half>((size_t)kTri * V.NN);
          V.Kscale = P.alloc<float>(V.NN);
        } else {
          V.Ksf = P.alloc<float>((size_t)kTri * V.NN);
        }
      }
    }
  }
  {
    long long tl = 0;
    for (int l = 0; l + 1 < P.L; ++l) tl = std::max(tl, restrict_tmp_len(geo(P.lv[l]), geo(P.lv[l + 1])));
    if (P.standard())  // also the per-cell products of the coarse smoothers (floats when mixed)
      for (int l = 1; l + 1 < P.L; ++l)
        tl = std::max(tl, P.mixed() ? (coarse_ks_tmp_len(geo(P.lv[l])) + 1) / 2 : coarse_ks_tmp_len(geo(P.lv[l])));
    P.rtmp = P.alloc<double>(tl);
  }
  P.fix = P.alloc<uint8_t>(NN);
  P.upload(P.fix, hfix.data(), NN);
  for (int ii = 0; ii < P.nCases; ++ii) {
    P.Fc.push_back(P.alloc<double>(3 * NN));
    P.upload(P.Fc[ii], hF[ii].data(), 3 * NN * sizeof(double));
    P.Uc.push_back(P.alloc<double>(3 * NN));
  }
  P.F = P.Fc[0];
  P.U = P.Uc[0];
  if (P.nCases > 1) P.ceTmp = P.alloc<double>(NN);
  for (double** v : {&P.r, &P.z, &P.p, &P.q, &P.tmp3}) *v = P.alloc<double>(3 * NN);
  for (double** v : {&P.E, &P.x, &P.xPhys, &P.ce, &P.dc, &P.ca, &P.cb, &P.cc, &P.cd}) *v = P.alloc<double>(NN);
  for (double** v : {&P.fb, &P.fx, &P.fr, &P.fz, &P.fp, &P.fq}) *v = P.alloc<double>(NN);
  // ---- Ke and L2 templates
  const std::vector<double> Ke = gauss_ke(P.P.nu, 1.0);
  P.dKe = P.alloc<double>(576);
  P.upload(P.dKe, Ke.data(), 576 * sizeof(double));
  P.sf = sf_coefficients(P.P.nu);
  if (P.mixed()) P.z32 = P.alloc<float>(3 * NN);
  if (P.standard() && !P.mixed()) {
    P.x64 = P.alloc<double>(3 * NN);
    P.d64 = P.alloc<double>(3 * NN);
  }
  if (P.standard() && P.mixed()) {
    P.E32 = P.alloc<float>(NN);
    P.x32 = P.alloc<float>(3 * NN);
    P.d32 = P.alloc<float>(3 * NN);
  }
  const std::vector<double> T2 = l2_templates(Ke, P.lv[1].span);
  P.dT2 = P.alloc<double>(T2.size());
  P.upload(P.dT2, T2.data(), T2.size() * sizeof(double));
  // ---- L2 cells containing a fine cell that touches a fixed node (1397-1422)
  {
    const Level& C = P.lv[1];
    std::vector<uint8_t> flag(C.NN, 0);
    for (int iz = 0; iz < F.nz; ++iz)
      for (int ix = 0; ix < F.nx; ++ix)
        for (int iy = 0; iy < F.ny; ++iy) {
          if (!F.hEmask[iy + (long long)F.NY * (ix + (long long)F.NX * iz)]) continue;
          bool t = false;
          for (int j = 0; j < 8; ++j)
            t |= hfix[(iy + kLocDy[j]) + (long long)F.NY * ((ix + kLocDx[j]) + (long long)F.NX * (iz + kLocDz[j]))] != 0;
          if (t) flag[iy / C.span + (long long)C.NY * (ix / C.span + (long long)C.NX * (iz / C.span))] = 1;
        }
    std::vector<int> list;
    for (long long c = 0; c < C.NN; ++c)
      if (flag[c]) list.push_back((int)c);
    P.nBc2 = (int)list.size();
    P.dBc2 = P.alloc<int>(list.size());
    if (!list.empty()) P.upload(P.dBc2, list.data(), list.size() * sizeof(int));
    P.dBcFlag2 = P.alloc<uint8_t>(C.NN);
    P.upload(P.dBcFlag2, flag.data(), C.NN);
  }
  // ---- coarsest band ordering: the axis permutation with the smallest bandwidth
  {
    const Level& C = P.lv[P.L - 1];
    const int dims[3] = {C.NY, C.NX, C.NZ};
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<int> bestPos;
    int bestBW = -1;
    for (auto& pm : perms) {
      std::vector<std::pair<long long, long long>> keys;
      for (long long n = 0; n < C.NN; ++n) {
        if (!C.hNmask[n]) continue;
        const int co[3] = {(int)(n % C.NY), (int)((n / C.NY) % C.NX), (int)(n / C.NY / C.NX)};
        keys.push_back({co[pm[0]] + (long long)dims[pm[0]] * (co[pm[1]] + (long long)dims[pm[1]] * co[pm[2]]), n});
      }
      std::sort(keys.begin(), keys.end());
      std::vector<int> pos(C.NN, -1);
      for (size_t k = 0; k < keys.size(); ++k) pos[keys[k].second] = (int)k;
      int bw = 0;
      for (int iz = 0; iz < C.nz; ++iz)
        for (int ix = 0; ix < C.nx; ++ix)
          for (int iy = 0; iy < C.ny; ++iy) {
            if (!C.hEmask[iy + (long long)C.NY * (ix + (long long)C.NX * iz)]) continue;
            int lo = 1 << 30, hi = -1;
            for (int j = 0; j < 8; ++j) {
              const int pp = pos[(iy + kLocDy[j]) + (long long)C.NY * ((ix + kLocDx[j]) + (long long)C.NX * (iz + kLocDz[j]))];
              lo = std::min(lo, pp);
              hi = std::max(hi, pp);
            }
            bw = std::max(bw, hi - lo);
          }
      if (bestBW < 0 || bw < bestBW) { bestBW = bw; bestPos = pos; }
    }
    P.cN = 3 * C.nNodes;
    P.cBW = 3 * bestBW + 2;
    std::vector<long long> cmap(P.cN);
    for (long long n = 0; n < C.NN; ++n)
      if (bestPos[n] >= 0)
        for (int c = 0; c < 3; ++c) cmap[3 * bestPos[n] + c] = c * C.NN + n;
    P.dCMap = P.alloc<long long>(P.cN);
    P.upload(P.dCMap, cmap.data(), P.cN * sizeof(long long));
    P.dNodePos = P.alloc<int>(C.NN);
    P.upload(P.dNodePos, bestPos.data(), C.NN * sizeof(int));
    P.cs.plan(P.cN, P.cBW);
    auto& f = P.cf;
    f.band = P.alloc<double>(P.cs.band_len());
    f.tinv = P.alloc<double>(P.cs.tinv_len());
    f.pan = P.alloc<double>(P.cs.panel_len());
    f.Dinv = P.alloc<double>(P.cs.blk_len());
    f.DT = P.alloc<double>(P.cs.blk_len());
    f.G = P.alloc<double>(P.cs.blk_len());
    f.H = P.alloc<double>(P.cs.blk_len());
    f.status = P.alloc<int>(1);
    // np doubles, then (one launch per recurrence) 2 np tagged words each,
    // zeroed: tags start at 1, so no word left in recycled memory can match
    P.cva = P.alloc<double>(3 * (size_t)P.cs.np);
    P.cvb = P.alloc<double>(3 * (size_t)P.cs.np);
    P.chainSt = P.alloc<ChainState>(1);
    T3X_HIP(hipMemsetAsync(P.cva, 0, 3 * (size_t)P.cs.np * sizeof(double), P.stream));
    T3X_HIP(hipMemsetAsync(P.cvb, 0, 3 * (size_t)P.cs.np * sizeof(double), P.stream));
    T3X_HIP(hipMemsetAsync(P.chainSt, 0, sizeof(ChainState), P.stream));
    P.sync();
    // the handles of the process that use the one-launch form together
    // need at most all of the device's work-group slots, so recurrences
    // running at the same time can always all be resident
    if (P.oneLaunchCoarse) {
      const long long need = P.cs.chain_share();
      long long held = gChainShare.load();
      while (held + need <= 1000000 && !gChainShare.compare_exchange_weak(held, held + need)) {
      }
      if (held + need <= 1000000) {
        P.cs.chain = P.chainSt;
        P.chainShare = need;
      }
    }
  }
  // ---- compact maps (eleMapBack / nodMapBack order)
  for (long long n = 0; n < NN; ++n) {
    if (F.hNmask[n]) P.hNodeMap.push_back((int)n);
    if (F.hEmask[n]) P.hCellMap.push_back((int)n);
  }
  P.nNodes = (int)P.hNodeMap.size();
  P.nCells = (int)P.hCellMap.size();
  P.dNodeMap = P.alloc<int>(P.nNodes);
  P.dCellMap = P.alloc<int>(P.nCells);
  P.upload(P.dNodeMap, P.hNodeMap.data(), P.nNodes * sizeof(int));
  P.upload(P.dCellMap, P.hCellMap.data(), P.nCells * sizeof(int));
  // ---- reductions and scalars
  P.ws.partials = P.alloc<double>((size_t)std::max(kMaxRedBlocks * 8, kMaxPartials));
  P.ws.counter = P.alloc<unsigned int>(4);
  P.sc = P.alloc<double>(64);
  P.fl = P.alloc<int>(16);
  P.oc = P.alloc<double>(8);
  P.mm = P.alloc<double>(MM_COUNT);
  P.dFiltKern = P.alloc<double>(128);
  if (!P.hbuf) T3X_HIP(hipHostMalloc((void**)&P.hbuf, 256 * sizeof(double), hipHostMallocDefault));
  return T3X_OK;
}

// ------------------------------------------------------------- assembly
// Solving_AssembleFEAstencil (919-1066).
int assemble(t3x_problem& P) {
  hipStream_t s = P.stream;
  const Geo g0 = geo(P.lv[0]);
  hipEvent_t ea[2];
  P.tbegin(TG_ASSEMBLE, ea);
  for (int l = 1; l < P.L; ++l) {
    Level& C = P.lv[l];
    const Geo gc = geo(C), gf = geo(P.lv[l - 1]);
    if (l == 1) {
      launch_galerkin_l2(s, gf, gc, C.span, P.dT2, P.E, C.emask, P.dBcFlag2, C.Ks);
      launch_galerkin_l2_bc(s, gf, gc, C.span, P.nBc2, P.dBc2, P.dKe, P.E, P.lv[0].emask, P.fix, C.Ks);
    } else {
      launch_galerkin_coarse(s, gf, gc, P.lv[l - 1].Ks, P.lv[l - 1].emask, C.emask, C.Ks);
    }
    if (l < P.L - 1) launch_diag_coarse(s, gc, C.Ks, C.emask, C.nmask, C.diag);
  }
  launch_diag_fine(s, g0, P.dKe, P.E, P.fix, P.lv[0].nmask, P.lv[0].diag);
  if (P.mixed())
    for (int l = 0; l < P.L - 1; ++l) launch_convert(s, P.lv[l].diagf, P.lv[l].diag, 3 * P.lv[l].NN);
  if (P.standard() && P.mixed()) {
    launch_convert(s, P.E32, P.E, P.NN());
    for (int l = 1; l < P.L - 1; ++l) {
      Level& V = P.lv[l];
      if (V.Ksh) launch_scale_ks_f16(s, geo(V), V.Ks, V.Ksh, V.Kscale);
      else launch_convert(s, V.Ksf, V.Ks, (long long)kTri * V.NN);
    }
  }
  const Level& C = P.lv[P.L - 1];
  auto& f = P.cf;
  launch_band_assemble(s, geo(C), C.Ks, C.emask, P.dNodePos, P.cBW, f.band, (long long)P.cs.band_len());
  P.cs.pad_identity(s, f.band);
  launch_set_int(s, f.status, 0);
  P.cs.factor(s, f.band, f.tinv, f.pan, f.Dinv, f.DT, f.G, f.H, f.status);
  P.tend(TG_ASSEMBLE, ea);
  T3X_HIP(hipMemcpyAsync(P.hbuf, f.status, sizeof(int), hipMemcpyDeviceToHost, s));
  P.sync();
  if (*(const int*)P.hbuf) { P.err = "coarsest matrix is not positive definite"; return T3X_ERR_NOT_PD; }
  P.assembled = true;
  return T3X_OK;
}

// ------------------------------------------------------------- V-cycle
// Solving_Vcycle, 'Adapted' (747-781); fuse: z.r reduction into the last
// prolongation (612 / 629).
void vcycle(t3x_problem& P, const double* r, double* z, double* sc, const int* fl, bool fuse) {
  hipStream_t s = P.stream;
  const int L = P.L;
  double* tmp = (double*)P.rtmp;
  hipEvent_t ev[2], ec[2];
  P.tbegin(TG_VCYCLE, ev);
  launch_restrict(s, geo(P.lv[0]), geo(P.lv[1]), P.lv[1].span, r, tmp, P.lv[1].nmask, P.lv[1].r, fl, P.fineAct());
  for (int l = 1; l < L - 1; ++l)
    launch_restrict(s, geo(P.lv[l]), geo(P.lv[l + 1]), P.lv[l + 1].span, (const double*)P.lv[l].r, tmp,
                    P.lv[l + 1].nmask, P.lv[l + 1].r, fl);
  Level& C = P.lv[L - 1];
  P.tbegin(TG_COARSE, ec);
  P.cs.solve(s, P.dCMap, C.r, C.x, P.fac().Dinv, P.fac().DT, P.fac().G, P.fac().H, P.cva, P.cvb, fl);
  P.tend(TG_COARSE, ec);
  for (int l = L - 2; l >= 1; --l)
    launch_prolong_jacobi(s, geo(P.lv[l]), geo(P.lv[l + 1]), P.lv[l + 1].span, P.P.omega,
                          (const double*)P.lv[l + 1].x, (const double*)P.lv[l].r, (const double*)P.lv[l].diag,
                          P.lv[l].nmask, P.lv[l].x, nullptr, nullptr, fl);
  launch_prolong_jacobi(s, geo(P.lv[0]), geo(P.lv[1]), P.lv[1].span, P.P.omega, (const double*)P.lv[1].x, r,
                        (const double*)P.lv[0].diag, P.lv[0].nmask, z, fuse ? &P.ws : nullptr, sc, fl);
  P.tend(TG_VCYCLE, ev);
}

// The same operator with FP32 level vectors and smoother output; the fine
// residual stays FP64 and the coarsest solve is the FP64 band Cholesky.
void vcycle_mixed(t3x_problem& P, const double* r, float* z, double* sc, const int* fl, bool fuse) {
  hipStream_t s = P.stream;
  const int L = P.L;
  float* tmp = (float*)P.rtmp;
  hipEvent_t ev[2], ec[2];
  P.tbegin(TG_VCYCLE, ev);
  launch_restrict(s, geo(P.lv[0]), geo(P.lv[1]), P.lv[1].span, r, tmp, P.lv[1].nmask, P.lv[1].rf, fl, P.fineAct());
  for (int l = 1; l < L - 1; ++l)
    launch_restrict(s, geo(P.lv[l]), geo(P.lv[l + 1]), P.lv[l + 1].span, (const float*)P.lv[l].rf, tmp,
                    P.lv[l + 1].nmask, P.lv[l + 1].rf, fl);
  Level& C = P.lv[L - 1];
  P.tbegin(TG_COARSE, ec);
  P.cs.solve(s, P.dCMap, C.rf, C.xf, P.fac().Dinv, P.fac().DT, P.fac().G, P.fac().H, P.cva, P.cvb, fl);
  P.tend(TG_COARSE, ec);
  for (int l = L - 2; l >= 1; --l)
    launch_prolong_jacobi(s, geo(P.lv[l]), geo(P.lv[l + 1]), P.lv[l + 1].span, P.P.omega,
                          (const float*)P.lv[l + 1].xf, (const float*)P.lv[l].rf, (const float*)P.lv[l].diagf,
                          P.lv[l].nmask, P.lv[l].xf, nullptr, nullptr, fl);
  launch_prolong_jacobi(s, geo(P.lv[0]), geo(P.lv[1]), P.lv[1].span, P.P.omega, (const float*)P.lv[1].xf, r,
                        (const float*)P.lv[0].diagf, P.lv[0].nmask, z, fuse ? &P.ws : nullptr, sc, fl);
  P.tend(TG_VCYCLE, ev);
}

// 'Standard' V-cycle (759-777) in FP32: x = w r/d; d = r - K x; restrict d
// down to the coarsest FP64 solve; then x += P x_coarse; x += w (r - K x)/d
// on the way up.  Fine-level K*u with E*Ke and the BC rows, coarse levels
// with the FP32 copies of the Galerkin matrices; z.r fused into the last
// smoothing step.
// x0Ready: the fine x = w r/d was already formed (by the PCG update kernel).
void vcycle_standard(t3x_problem& P, const double* r, float* z, double* sc, const int* fl, bool fuse,
                     bool x0Ready = false) {
  hipStream_t s = P.stream;
  const int L = P.L;
  const double om = P.P.omega;
  float* tmp = (float*)P.rtmp;
  const Geo g0 = geo(P.lv[0]);
  hipEvent_t ev[2], ec[2];
  P.tbegin(TG_VCYCLE, ev);
  if (!x0Ready) launch_jacobi0(s, r, P.lv[0].diagf, P.x32, om, 3 * P.lv[0].NN, fl, P.fineAct());  // 757
  launch_fine_resid32_sf(s, g0, P.sf, P.E32, P.fix, P.x32, r, P.d32, fl);                  // 761
  launch_restrict(s, g0, geo(P.lv[1]), P.lv[1].span, (const float*)P.d32, tmp, P.lv[1].nmask, P.lv[1].rf, fl,
                  P.fineAct());
  for (int l = 1; l < L - 1; ++l) {
    Level& V = P.lv[l];
    launch_jacobi0(s, (const float*)V.rf, V.diagf, V.xf, om, 3 * V.NN, fl);
    if (V.Ksh) launch_coarse_resid32(s, geo(V), V.Ksh, V.Kscale, V.xf, V.rf, V.tf, tmp, fl, V.emask, V.nmask);
    else launch_coarse_resid32(s, geo(V), V.Ksf, V.xf, V.rf, V.tf, tmp, fl, V.emask, V.nmask);
    launch_restrict(s, geo(V), geo(P.lv[l + 1]), P.lv[l + 1].span, (const float*)V.tf, tmp, P.lv[l + 1].nmask,
                    P.lv[l + 1].rf, fl);
  }
  Level& C = P.lv[L - 1];
  P.tbegin(TG_COARSE, ec);
  P.cs.solve(s, P.dCMap, C.rf, C.xf, P.fac().Dinv, P.fac().DT, P.fac().G, P.fac().H, P.cva, P.cvb, fl);  // 767
  P.tend(TG_COARSE, ec);
  for (int l = L - 2; l >= 1; --l) {
    Level& V = P.lv[l];
    launch_prolong_add(s, geo(V), geo(P.lv[l + 1]), P.lv[l + 1].span, P.lv[l + 1].xf, V.xf, V.nmask, fl);  // 770
    if (V.Ksh)
      launch_coarse_smooth32(s, geo(V), V.Ksh, V.Kscale, V.xf, V.rf, V.diagf, om, V.tf, tmp, fl, V.emask, V.nmask);
    else
      launch_coarse_smooth32(s, geo(V), V.Ksf, V.xf, V.rf, V.diagf, om, V.tf, tmp, fl, V.emask, V.nmask);  // 775
    std::swap(V.xf, V.tf);
  }
  launch_prolong_add(s, g0, geo(P.lv[1]), P.lv[1].span, P.lv[1].xf, P.x32, P.lv[0].nmask, fl);
  launch_fine_smooth32_sf(s, g0, P.sf, P.E32, P.fix, P.x32, r, P.lv[0].diagf, om, z,
                          fuse ? P.ws.partials : nullptr, fl);
  if (fuse) launch_fin1(s, fl ? FIN_ZR_LOOP : FIN_ZR_INIT, applyk_sf_blocks(g0), P.ws, sc, (int*)fl, nullptr);
  P.tend(TG_VCYCLE, ev);
}

// 'Standard' V-cycle (759-777) in FP64: the same steps as the FP32 form above
// with the FP64 Galerkin matrices, diagonals and level vectors.
void vcycle_standard64(t3x_problem& P, const double* r, double* z, double* sc, const int* fl, bool fuse) {
  hipStream_t s = P.stream;
  const int L = P.L;
  const double om = P.P.omega;
  double* tmp = (double*)P.rtmp;
  const Geo g0 = geo(P.lv[0]);
  hipEvent_t ev[2], ec[2];
  P.tbegin(TG_VCYCLE, ev);
  launch_jacobi0(s, r, (const double*)P.lv[0].diag, P.x64, om, 3 * P.lv[0].NN, fl, P.fineAct());  // 757
  launch_fine_resid64_sf(s, g0, P.sf, P.E, P.fix, P.x64, r, P.d64, fl);                      // 761
  launch_restrict(s, g0, geo(P.lv[1]), P.lv[1].span, (const double*)P.d64, tmp, P.lv[1].nmask, P.lv[1].r, fl,
                  P.fineAct());
  for (int l = 1; l < L - 1; ++l) {
    Level& V = P.lv[l];
    launch_jacobi0(s, (const double*)V.r, (const double*)V.diag, V.x, om, 3 * V.NN, fl);
    launch_coarse_resid64(s, geo(V), V.Ks, V.x, V.r, V.t, tmp, fl, V.emask, V.nmask);
    launch_restrict(s, geo(V), geo(P.lv[l + 1]), P.lv[l + 1].span, (const double*)V.t, tmp, P.lv[l + 1].nmask,
                    P.lv[l + 1].r, fl);
  }
  Level& C = P.lv[L - 1];
  P.tbegin(TG_COARSE, ec);
  P.cs.solve(s, P.dCMap, C.r, C.x, P.fac().Dinv, P.fac().DT, P.fac().G, P.fac().H, P.cva, P.cvb, fl);  // 767
  P.tend(TG_COARSE, ec);
  for (int l = L - 2; l >= 1; --l) {
    Level& V = P.lv[l];
    launch_prolong_add(s, geo(V), geo(P.lv[l + 1]), P.lv[l + 1].span, (const double*)P.lv[l + 1].x, V.x, V.nmask,
                       fl);                                                                   // 770
    launch_coarse_smooth64(s, geo(V), V.Ks, V.x, V.r, V.diag, om, V.t, tmp, fl, V.emask, V.nmask);  // 775
    std::swap(V.x, V.t);
  }
  launch_prolong_add(s, g0, geo(P.lv[1]), P.lv[1].span, (const double*)P.lv[1].x, P.x64, P.lv[0].nmask, fl);
  launch_fine_smooth64_sf(s, g0, P.sf, P.E, P.fix, P.x64, r, P.lv[0].diag, om, z, fuse ? P.ws.partials : nullptr, fl);
  if (fuse) launch_fin1(s, fl ? FIN_ZR_LOOP : FIN_ZR_INIT, applyk_sf_blocks(g0), P.ws, sc, (int*)fl, nullptr);
  P.tend(TG_VCYCLE, ev);
}

// ------------------------------------------------------------- PCG
// The coarsest solve's failure word (read into hbuf[255] after a solve): a
// wait of the one-launch recurrences expired, the solve is garbage.
static void check_chain(t3x_problem& P) {
  int failed = 0;
  std::memcpy(&failed, P.hbuf + 255, sizeof(int));
  if (!P.cs.chain || failed == 0) return;
  throw HipError("coarsest solve: a wait between work-groups expired (one-launch recurrences)");
}

// Solving_PCG (600-639) with device-side scalars and convergence flag.
int pcg_elastic(t3x_problem& P, const double* b, double* y, double tol, int maxIT, double* resOut, bool* warned) {
  hipStream_t s = P.stream;
  const long long n3 = 3 * P.NN();
  double* sc = P.sc;
  int* fl = P.fl;
  const Geo g0 = geo(P.lv[0]);
  const bool mx = P.mixed();
  launch_pcg_init(s, sc, fl, tol);
  launch_norm2(s, b, n3, P.ws, &sc[SC_NORMB]);                                              // 607
  launch_apply_K(s, g0, P.sf, P.E, P.fix, y, P.q, P.ws, sc, nullptr, false);                // 609
  launch_sub(s, P.r, b, P.q, n3);
  if (mx) {
    if (P.standard()) vcycle_standard(P, P.r, P.z32, sc, nullptr, true);                   // 610, 612
    else vcycle_mixed(P, P.r, P.z32, sc, nullptr, true);
    launch_convert(s, P.p, P.z32, n3);                                                      // 611
  } else {
    if (P.standard()) vcycle_standard64(P, P.r, P.z, sc, nullptr, true);
    else vcycle(P, P.r, P.z, sc, nullptr, true);
    launch_copy(s, P.p, P.z, n3);
  }
  const int ce = std::max(1, P.P.checkEvery);
  int its = maxIT;
  bool done = false;
  // one iteration: K*p (614-615), r and the residual norm (617-622), z = M r
  // (628-629), y and p (616, 630)
  auto iteration = [&]() {
    hipEvent_t ek[2];
    P.tbegin(TG_APPLYK, ek);
    launch_apply_K(s, g0, P.sf, P.E, P.fix, P.p, P.q, P.ws, sc, fl, true);                  // 614-615
    P.tend(TG_APPLYK, ek);
    if (mx && P.standard())  // 617-622, with the first fine Jacobi step of 628-629 fused
      launch_pcg_update(s, y, P.r, P.p, P.q, n3, P.ws, sc, fl, P.lv[0].diagf, P.x32, P.P.omega, P.fineAct());
    else
      launch_pcg_update(s, y, P.r, P.p, P.q, n3, P.ws, sc, fl, nullptr, nullptr, 0.0, P.fineAct());  // 617-622
    if (mx) {
      if (P.standard()) vcycle_standard(P, P.r, P.z32, sc, fl, true, true);                // 628-629
      else vcycle_mixed(P, P.r, P.z32, sc, fl, true);
      launch_pcg_pupdate(s, y, P.p, P.z32, n3, sc, fl, P.fineAct());                        // 616, 630
    } else {
      if (P.standard()) vcycle_standard64(P, P.r, P.z, sc, fl, true);                         // 628-629
      else vcycle(P, P.r, P.z, sc, fl, true);
      launch_pcg_pupdate(s, y, P.p, P.z, n3, sc, fl, P.fineAct());
    }
  };
  // With params.graphs (and no kernel-group timing) the iteration is captured
  // once per solution vector and replayed: every kernel reads its scalars and
  // the convergence flag from device memory, so the same graph serves every
  // iteration.  Measured on MI355X it neither gains nor loses (0.27 ms per
  // PCG iteration at 32x64x32, 1.44 ms at 128x256x128, either way): the
  // kernels' own dispatch latency, not host launch cost, bounds small
  // problems, so it is off by default.
  hipGraphExec_t graph = nullptr;
  if (P.P.graphs && !P.P.timing) {
    for (auto& g : P.pcgGraphs)
      if (g.first == y) graph = g.second;
    if (!graph) {
      hipGraph_t gr = nullptr;
      T3X_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        iteration();
      } catch (...) {  // leave the stream out of capture mode before reporting
        (void)hipStreamEndCapture(s, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      T3X_HIP(hipStreamEndCapture(s, &gr));
      const hipError_t e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      T3X_HIP(e);
      P.pcgGraphs.push_back({y, graph});
    }
  }
  for (int it = 1; it <= maxIT; ++it) {
    if (graph) T3X_HIP(hipGraphLaunch(graph, s));
    else iteration();
    if (it % ce == 0 || it == maxIT) {
      T3X_HIP(hipMemcpyAsync(P.hbuf, fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, s));
      P.sync();
      const int* hf = (const int*)P.hbuf;
      if (hf[FL_DONE]) { done = true; its = hf[FL_ITS]; break; }
    }
  }
  launch_pcg_finish(s, y, P.p, n3, sc, fl, P.fineAct());
  T3X_HIP(hipMemcpyAsync(P.hbuf, sc, SC_COUNT_PCG * sizeof(double), hipMemcpyDeviceToHost, s));
  if (P.cs.chain) T3X_HIP(hipMemcpyAsync(P.hbuf + 255, &P.cs.chain->fail, sizeof(int), hipMemcpyDeviceToHost, s));
  P.sync();
  check_chain(P);
  if (resOut) *resOut = P.hbuf[SC_RES];
  if (warned) *warned = !done;  // 634-637
  return its;
}

// ------------------------------------------------------------- PDE filter
void filter_setup(t3x_problem& P, int slot, double rMin) {
  // TopOpti_SetupPDEfilter_matrixFree (816-868): iKEF = r^2 KEF0 + KEF1 with
  // KEF0 = sum over 8 Gauss points of natural-coordinate gradient products and
  // KEF1 the consistent mass with detJ = 1/8.
  const double g = 1.0 / std::sqrt(3.0);
  double K0[8][8] = {}, K1[8][8] = {};
  for (int q = 0; q < 8; ++q) {
    const double sq = (q & 1) ? g : -g, tq = (q & 2) ? g : -g, pq = (q & 4) ? g : -g;
    double N[8], ds[8], dt[8], dp[8];
    for (int a = 0; a < 8; ++a) {
      const double sa = kLocDx[a] ? 1.0 : -1.0, ta = kLocDy[a] ? -1.0 : 1.0, pa = kLocDz[a] ? 1.0 : -1.0;
      N[a] = 0.125 * (1 + sa * sq) * (1 + ta * tq) * (1 + pa * pq);
      ds[a] = 0.125 * sa * (1 + ta * tq) * (1 + pa * pq);
      dt[a] = 0.125 * ta * (1 + sa * sq) * (1 + pa * pq);
      dp[a] = 0.125 * pa * (1 + sa * sq) * (1 + ta * tq);
    }
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        K0[a][b] += ds[a] * ds[b] + dt[a] * dt[b] + dp[a] * dp[b];
        K1[a][b] += N[a] * (1.0 / 8.0) * N[b];
      }
  }
  const double ir = (rMin * 1.0) / 2.0 / std::sqrt(3.0);  // 850
  Filter& Fl = P.filt[slot];
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) Fl.kernel[b * 8 + a] = ir * ir * K0[a][b] + K1[a][b];
  P.upload(P.dFiltKern + slot * 64, Fl.kernel, 64 * sizeof(double));
  if (!Fl.diagInv) Fl.diagInv = P.alloc<double>(P.NN());
  launch_filter_diag(P.stream, geo(P.lv[0]), P.dFiltKern + slot * 64, P.lv[0].emask, P.lv[0].nmask, Fl.diagInv);
  Fl.ready = true;
  Fl.r = rMin;
}

// TopOpti_ConductPDEFiltering_matrixFree (870-902); src/dst are dense cell arrays.
int filter_run(t3x_problem& P, int slot, const double* src, double* dst) {
  hipStream_t s = P.stream;
  const Geo g0 = geo(P.lv[0]);
  const long long NN = P.NN();
  double* sc = P.sc + 16;
  int* fl = P.fl + 8;
  const double* kern = P.dFiltKern + slot * 64;
  const double* dinv = P.filt[slot].diagInv;
  launch_filter_e2n(s, g0, P.lv[0].emask, src, P.fb);  // 880-886
  launch_pcg_init(s, sc, fl, P.P.tol / 10.0);          // 891
  launch_norm2(s, P.fb, NN, P.ws, &sc[SC_NORMB]);
  launch_fill(s, P.fx, 0.0, NN);
  launch_filter_apply(s, g0, kern, P.lv[0].emask, P.fx, P.fq, nullptr, nullptr, nullptr);
  launch_sub(s, P.fr, P.fb, P.fq, NN);
  launch_scale_diag(s, P.fz, dinv, P.fr, NN, &P.ws, sc, nullptr);
  launch_copy(s, P.fp, P.fz, NN);
  const int maxIT = P.P.maxIT;
  const int ce = std::max(1, P.P.checkEvery);
  int its = maxIT;
  for (int it = 1; it <= maxIT; ++it) {
    launch_filter_apply(s, g0, kern, P.lv[0].emask, P.fp, P.fq, &P.ws, sc, fl);
    launch_pcg_update(s, P.fx, P.fr, P.fp, P.fq, NN, P.ws, sc, fl, nullptr, nullptr, 0.0, P.fineAct());
    launch_scale_diag(s, P.fz, dinv, P.fr, NN, &P.ws, sc, fl);
    launch_pcg_pupdate(s, P.fx, P.fp, P.fz, NN, sc, fl, P.fineAct());
    if (it % ce == 0 || it == maxIT) {
      T3X_HIP(hipMemcpyAsync(P.hbuf, fl, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
      P.sync();
      const int* hf = (const int*)P.hbuf;
      if (hf[FL_DONE]) { its = hf[FL_ITS]; break; }
    }
  }
  launch_pcg_finish(s, P.fx, P.fp, NN, sc, fl, P.fineAct());
  launch_filter_n2e(s, g0, P.lv[0].emask, P.fx, dst);  // 894-901
  return its;
}

double read_scalar(t3x_problem& P, const double* d) {
  T3X_HIP(hipMemcpyAsync(P.hbuf, d, sizeof(double), hipMemcpyDeviceToHost, P.stream));
  P.sync();
  return P.hbuf[0];
}

double masked_sum(t3x_problem& P, const double* a, const double* b, int mode) {
  launch_masked_sum(P.stream, P.NN(), P.lv[0].emask, a, b, mode, P.ws, P.sc + 32);
  return read_scalar(P, P.sc + 32);
}

// ------------------------------------------------------------- OC (145-155)
double oc_update(t3x_problem& P, double* x, const double* xPhys, const double* dc, double V0, int* nBis) {
  hipStream_t s = P.stream;
  const double fval = masked_sum(P, xPhys, nullptr, 0) / (double)P.nCells - V0;  // 145
  const double init[8] = {0.0, 1e9, fval, 0.0, 0.0, 0.0, 0.0, 0.0};            // 146
  launch_set_vals(s, P.oc, 8, init);
  for (int pass = 0; pass < 64; ++pass) {
    launch_oc_pass(s, P.NN(), P.lv[0].emask, P.passive, x, dc, P.ws, P.oc, 3);
    if (pass % 8 == 7) {
      T3X_HIP(hipMemcpyAsync(P.hbuf, P.oc, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
      P.sync();
      if (P.hbuf[3] != 0.0) break;
    }
  }
  launch_oc_final(s, P.NN(), P.lv[0].emask, P.passive, x, dc, P.ws, P.oc);
  T3X_HIP(hipMemcpyAsync(P.hbuf, P.oc, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
  P.sync();
  if (nBis) *nBis = (int)P.hbuf[5];
  return P.hbuf[6];
}

// ------------------------------------------------------------- MMA (1804-2049)
void mma_alloc(t3x_problem& P) {
  if (P.mx) return;
  const long long NN = P.NN();
  for (double** v : {&P.mx, &P.mxold1, &P.mxold2, &P.mlow, &P.mupp, &P.malpha, &P.mbeta, &P.mp0, &P.mq0, &P.mpij,
                     &P.mqij, &P.mxmin, &P.mxmax})
    *v = P.alloc<double>(NN);
}

// One MMA update on the dense arrays: P.mx holds xval on entry, the new x on exit.
int mma_update(t3x_problem& P, int iter, const double* dfdx, double gx, const double* dgdx) {
  hipStream_t s = P.stream;
  const long long NN = P.NN();
  launch_set_vals(s, P.mm + MM_GX, 1, &gx);
  const uint8_t* am = P.design_mask();
  launch_mma_gensub(s, NN, am, iter, P.mx, P.mxmin, P.mxmax, P.mxold1, P.mxold2, P.mlow, P.mupp, dfdx, dgdx,
                    P.malpha, P.mbeta, P.mp0, P.mq0, P.mpij, P.mqij, P.ws, P.mm, (double)(P.nCells - P.nPassive));
  for (int k = 0; k < 4000; ++k) {
    launch_mma_step(s, NN, am, P.mx, P.mlow, P.mupp, P.malpha, P.mbeta, P.mp0, P.mq0, P.mpij, P.mqij, P.ws, P.mm);
    if (k % 4 == 3) {
      T3X_HIP(hipMemcpyAsync(P.hbuf, P.mm, MM_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
      P.sync();
      if (P.hbuf[MM_DONE] != 0.0) break;
    }
  }
  T3X_HIP(hipMemcpyAsync(P.hbuf, P.mm, MM_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
  P.sync();
  return (int)P.hbuf[MM_NEWTON];
}

// ------------------------------------------------------------- drivers
// TopOpti_ComputeUnitCompliance (579-598) over the load cases: ce = sum_j
// ce_j w_j (597), accumulated in case order.  A single case has weight 1
// (493 with one case, 558), so its ce is ce_0.
void compliance_cases(t3x_problem& P) {
  hipStream_t s = P.stream;
  const uint8_t* em = P.lv[0].emask;
  const Geo g0 = geo(P.lv[0]);
  if (P.nCases == 1) {
    launch_compliance_sf(s, g0, P.sf, em, P.Uc[0], P.ce);
    return;
  }
  for (int j = 0; j < P.nCases; ++j) {
    launch_compliance_sf(s, g0, P.sf, em, P.Uc[j], P.ceTmp);
    launch_elementwise(s, j == 0 ? OP_MULC : OP_AXPYC, P.NN(), em, P.ce, P.ceTmp, nullptr, P.wObj[j], 0, 0);
  }
}

// ------------------------------------------------------------- drivers
// TOP3D_XL_TO (40-223) / TOP3D_XL_PIO (225-440) split into begin / one
// design iteration / finish so callers (bench, serving engine) can drive
// the loop one iteration at a time.
int opt_begin(t3x_problem& P, int mode, double V0, int nLoop, double rMin, double rHat) {
  hipStream_t s = P.stream;
  const long long NN = P.NN();
  const uint8_t* em = P.lv[0].emask;
  OptState& S = P.opt;
  S = OptState();
  S.mode = mode; S.V0 = V0; S.nLoop = nLoop; S.rMin = rMin; S.rHat = rHat;
  S.hist.clear();
  filter_setup(P, 0, rMin);                 // 76 / 268
  if (mode == 1) filter_setup(P, 1, rHat);  // 269
  launch_elementwise(s, OP_SET, NN, em, P.x, nullptr, nullptr, V0, 0, 0);      // 82
  launch_elementwise(s, OP_SET, NN, em, P.xPhys, nullptr, nullptr, V0, 0, 0);  // 83
  launch_elementwise(s, OP_SET, NN, em, P.xTilde(), nullptr, nullptr, V0, 0, 0);  // 279
  launch_elementwise(s, OP_SET, NN, em, P.volMax(), nullptr, nullptr, V0, 0, 0);  // 278
  for (double* u : P.Uc) launch_fill(s, u, 0.0, 3 * NN);                          // 95
  if (mode == 1) mma_alloc(P);
  // solid reference solve (94-106 / 294-306)
  launch_elementwise(s, OP_SET, NN, em, P.E, nullptr, nullptr, P.P.E0, 0, 0);
  int rc = assemble(P);
  if (rc) return rc;
  for (int j = 0; j < P.nCases; ++j) {  // 101 / 301: one solve per F_ column
    bool warned = false;
    pcg_elastic(P, P.Fc[j], P.Uc[j], P.P.tol, P.P.maxIT, nullptr, &warned);
    if (warned) P.warnings++;
  }
  compliance_cases(P);
  S.cSolid = masked_sum(P, P.E, P.ce, 1);  // 104
  S.active = true;
  return T3X_OK;
}

// One design iteration (112-182 / 312-400).  Returns 1 when the loop
// condition (111 / 311) stops the optimisation, 0 otherwise.
int opt_step(t3x_problem& P, t3x_iter* it) {
  OptState& S = P.opt;
  if (!S.active) { P.err = "t3x_opt_step without t3x_opt_begin"; return -T3X_ERR_INPUT; }
  if (!(S.loop < S.nLoop && S.change > 0.0001 && S.sharp > 0.01)) return 1;  // 111 / 311
  hipStream_t s = P.stream;
  const long long NN = P.NN();
  const uint8_t* em = P.lv[0].emask;
  const double E0 = P.P.E0, Emin = P.P.Emin, V0 = S.V0;
  const double nE = (double)P.nCells;
  const double cSolid = S.cSolid;
  const double eta = 0.5;
  double* tmp = P.cc;
  const double t0 = wall();
  ++S.loop;
  ++S.loopbeta;
  launch_simp(s, NN, em, P.xPhys, Emin, E0, P.E);  // 116
  P.sync();
  const double ta = wall();
  int rc = assemble(P);  // 118
  if (rc) return -rc;
  const double tb = wall();
  int its = 0;
  double resFinal = 0.0;
  for (int j = 0; j < P.nCases; ++j) {  // 121 / 321: one warm-started solve per F_ column
    bool warned = false;
    double res = 0.0;
    const int itj = pcg_elastic(P, P.Fc[j], P.Uc[j], P.P.tol, P.P.maxIT, &res, &warned);
    if (warned) P.warnings++;
    S.lssIts.push_back(itj);
    S.resAll.push_back(res);
    its += itj;
    resFinal = std::max(resFinal, res);
  }
  const double tcg = wall();
  compliance_cases(P);                                                       // 124
  launch_elementwise(s, OP_DIVC, NN, em, tmp, P.ce, nullptr, cSolid, 0, 0);  // 125 ceNorm
  const double cObj = masked_sum(P, P.E, tmp, 1);                            // 126
  const double cDesign = cObj * cSolid;                                      // 127
  const double V = masked_sum(P, P.xPhys, nullptr, 0) / nE;                  // 128
  launch_elementwise(s, OP_DC, NN, em, P.dc, P.xPhys, P.ce, P.P.penal * (E0 - Emin), cSolid, 0);  // 129
  double tOpt = wall() - tcg, tFilt = 0.0, tLVF = 0.0, fval = 0.0;
  int fIts = 0;
  if (S.mode == 0) {
    const double t1 = wall();
    launch_elementwise(s, OP_MUL, NN, em, tmp, P.x, P.dc, 0, 0, 0);  // 137
    fIts += filter_run(P, 0, tmp, P.dc);
    launch_elementwise(s, OP_DIVMAX, NN, em, P.dc, P.x, nullptr, 0, 0, 0);
    P.sync();
    const double t2 = wall();
    tFilt += t2 - t1;
    fval = masked_sum(P, P.xPhys, nullptr, 0) / nE - V0;  // 145
    S.change = oc_update(P, P.x, P.xPhys, P.dc, V0, nullptr);  // 146-155
    launch_copy(s, P.xPhys, P.x, NN);  // 160
    P.sync();
    tOpt += wall() - t2;
  } else {
    const int pp = 16;  // pPIO (235)
    double* xhat = P.cd;
    double* xTilde = P.xTilde();
    double* volMax = P.volMax();
    const double t1 = wall();
    fIts += filter_run(P, 1, P.xPhys, xhat);  // 334
    const double sp = masked_sum(P, xhat, volMax, 3) / nE;
    const double fac = std::pow(sp, 1.0 / pp - 1.0);
    double* dfdx_pde = P.mxmin;  // scratch until the MMA bounds are set
    launch_elementwise(s, OP_PNORMGRAD, NN, em, dfdx_pde, xhat, volMax, fac, nE, 0);  // 335
    P.sync();
    const double t2 = wall();
    tLVF += t2 - t1;
    double* dx = P.mxmax;
    launch_elementwise(s, OP_DHEAVI, NN, em, dx, xTilde, nullptr, S.beta, eta, 0);  // 340
    launch_elementwise(s, OP_MUL, NN, em, tmp, P.dc, dx, 0, 0, 0);
    fIts += filter_run(P, 0, tmp, P.dc);  // 341
    P.sync();
    const double t3 = wall();
    tFilt += t3 - t2;
    double* dfdx = P.tmp3;
    fIts += filter_run(P, 1, dfdx_pde, dfdx);  // 346
    P.sync();
    const double t4 = wall();
    tLVF += t4 - t3;
    launch_elementwise(s, OP_MUL, NN, em, tmp, dfdx, dx, 0, 0, 0);
    fIts += filter_run(P, 0, tmp, dfdx);  // 350
    P.sync();
    const double t5 = wall();
    tFilt += t5 - t4;
    // MMA (354-369) over the active cells (274-276)
    if (S.loop == 1) {  // 355
      launch_copy(s, P.mxold1, P.x, NN);
      launch_copy(s, P.mxold2, P.x, NN);
      launch_fill(s, P.mlow, 0.0, NN);
      launch_fill(s, P.mupp, 0.0, NN);
    }
    fval = std::pow(sp, 1.0 / pp) - 1.0;                                                 // 360
    launch_elementwise(s, OP_MMA_BOUNDS, NN, em, P.mxmin, P.x, nullptr, 0.1, 0.0, 0);  // 364
    launch_elementwise(s, OP_MMA_BOUNDS, NN, em, P.mxmax, P.x, nullptr, 0.1, 1.0, 0);  // 365
    launch_copy(s, P.mx, P.x, NN);
    mma_update(P, S.loop, P.dc, fval, dfdx);  // 367
    launch_copy(s, tmp, P.x, NN);             // xval
    launch_copy(s, P.x, P.mx, NN);            // 369
    if (P.passive) launch_set_where(s, NN, P.passive, P.x, 1.0);  // 366: x(passive) = 1
    launch_masked_maxdiff(s, NN, P.design_mask(), P.x, tmp, P.ws, P.sc + 33);  // 368
    S.change = read_scalar(P, P.sc + 33);
    const double t6 = wall();
    tOpt += t6 - t5;
    fIts += filter_run(P, 0, P.x, xTilde);                                          // 374
    if (S.loopbeta >= 40) { S.beta = std::min(2 * S.beta, 128.0); S.loopbeta = 0; }  // 375-379
    launch_elementwise(s, OP_HEAVI, NN, em, P.xPhys, xTilde, nullptr, S.beta, eta, 0);  // 380
    if (P.passive) launch_set_where(s, NN, P.passive, P.xPhys, 1.0);                  // 381
    P.sync();
    tFilt += wall() - t6;
  }
  S.sharp = 4.0 * masked_sum(P, P.xPhys, nullptr, 2) / nE;  // 166 / 382
  const double tTot = wall() - t0;
  t3x_iter r{};
  r.loop = S.loop; r.c = cDesign; r.vol = V; r.cons = fval; r.sharp = S.sharp; r.change = S.change;
  r.pcgIts = its; r.filterIts = fIts; r.res = resFinal;
  r.t[0] = tb - ta; r.t[1] = tcg - tb; r.t[2] = tOpt; r.t[3] = tFilt; r.t[4] = tLVF; r.t[5] = tTot;
  S.hist.push_back(r);
  if (it) *it = r;
  return (S.loop < S.nLoop && S.change > 0.0001 && S.sharp > 0.01) ? 0 : 1;
}

int opt_result(t3x_problem& P, t3x_result* O) {
  hipStream_t s = P.stream;
  const long long NN = P.NN();
  const OptState& S = P.opt;
  const int n = (int)S.hist.size();
  for (int k = 0; k < n; ++k) {
    const t3x_iter& r = S.hist[k];
    if (O->cHist) O->cHist[k] = r.c;
    if (O->volHist) O->volHist[k] = r.vol;
    if (O->consHist) O->consHist[k] = r.cons;
    if (O->sharpHist) O->sharpHist[k] = r.sharp;
    if (O->tHist) std::memcpy(O->tHist + (size_t)k * 6, r.t, 6 * sizeof(double));
  }
  for (size_t k = 0; k < S.lssIts.size(); ++k) {  // one entry per solve, case fastest
    if (O->lssIts) O->lssIts[k] = S.lssIts[k];
    if (O->resHist) O->resHist[k] = S.resAll[k];
  }
  if (O->xPhys) {
    launch_gather_compact(s, P.dCellMap, P.nCells, 1, NN, P.xPhys, P.tmp3);
    T3X_HIP(hipMemcpyAsync(O->xPhys, P.tmp3, P.nCells * sizeof(double), hipMemcpyDeviceToHost, s));
    P.sync();
  }
  O->cSolid = S.cSolid;
  O->loops = n;
  O->warnings = P.warnings;
  return T3X_OK;
}

int run(t3x_problem& P, int mode, double V0, int nLoop, double rMin, double rHat, t3x_result* O) {
  int rc = opt_begin(P, mode, V0, nLoop, rMin, rHat);
  if (rc) return rc;
  for (;;) {
    const int k = P.opt.loop;
    rc = opt_step(P, nullptr);
    if (rc < 0) return -rc;
    if (O->xSnap && k < O->nSnap && P.opt.loop > k) {
      launch_gather_compact(P.stream, P.dCellMap, P.nCells, 1, P.NN(), P.xPhys, P.tmp3);
      T3X_HIP(hipMemcpyAsync(O->xSnap + (size_t)k * P.nCells, P.tmp3, P.nCells * sizeof(double),
                             hipMemcpyDeviceToHost, P.stream));
      P.sync();
    }
    if (rc == 1) break;
  }
  return opt_result(P, O);
}

// Upload a compact nodal vector (3 per node) into a zeroed dense SoA array.
void upload_nodal(t3x_problem& P, const double* h, double* d) {
  hipStream_t s = P.stream;
  launch_fill(s, d, 0.0, 3 * P.NN());
  double* stage = P.tmp3;  // compact staging (3*nNodes <= 3*NN)
  T3X_HIP(hipMemcpyAsync(stage, h, 3 * (size_t)P.nNodes * sizeof(double), hipMemcpyHostToDevice, s));
  launch_scatter_compact(s, P.dNodeMap, P.nNodes, 3, P.NN(), stage, d);
}
void download_nodal(t3x_problem& P, const double* d, double* h) {
  hipStream_t s = P.stream;
  launch_gather_compact(s, P.dNodeMap, P.nNodes, 3, P.NN(), d, P.tmp3);
  T3X_HIP(hipMemcpyAsync(h, P.tmp3, 3 * (size_t)P.nNodes * sizeof(double), hipMemcpyDeviceToHost, s));
  P.sync();
}
void upload_cells(t3x_problem& P, const double* h, double* d) {
  hipStream_t s = P.stream;
  launch_fill(s, d, 0.0, P.NN());
  T3X_HIP(hipMemcpyAsync(P.tmp3, h, (size_t)P.nCells * sizeof(double), hipMemcpyHostToDevice, s));
  launch_scatter_compact(s, P.dCellMap, P.nCells, 1, P.NN(), P.tmp3, d);
}
void download_cells(t3x_problem& P, const double* d, double* h) {
  hipStream_t s = P.stream;
  launch_gather_compact(s, P.dCellMap, P.nCells, 1, P.NN(), d, P.tmp3);
  T3X_HIP(hipMemcpyAsync(h, P.tmp3, (size_t)P.nCells * sizeof(double), hipMemcpyDeviceToHost, s));
  P.sync();
}

template <class F>
int guarded(t3x_problem* h, F f) {
  if (!h) return T3X_ERR_INPUT;
  try {
    return f();
  } catch (const HipError& e) {
    h->err = e.msg;
    return T3X_ERR_HIP;
  } catch (const std::exception& e) {
    h->err = e.what();
    return T3X_ERR_HIP;
  }
}

}  // namespace

static int create_impl(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                       const uint8_t* fixState, int nCases, const int* nPerCase, const int* loadNodes,
                       const double* loadVec, const t3x_params* params, const t3x::PoolSlot* slot,
                       t3x_problem** out) {
  if (!out) return T3X_ERR_INPUT;
  *out = nullptr;
  std::unique_ptr<t3x_problem> P(new t3x_problem());
  if (params) P->P = *params; else t3x_default_params(&P->P);
  if (slot) {  // pooled resources lent by the serving engine
    P->arena = (char*)slot->base;
    P->arenaSize = slot->size;
    P->stream = slot->stream;
    P->ownStream = false;
    P->hbuf = slot->hostBuf;
    P->ownHostBuf = false;
    P->oneLaunchCoarse = slot->oneLaunchCoarse;
  }
  P->nely = nely; P->nelx = nelx; P->nelz = nelz;
  if (!voxels || nelx < 3 || nely < 3 || nelz < 3) {  // 532-534
    P->err = "Inappropriate Input Model!";
    *out = P.release();
    return T3X_ERR_INPUT;
  }
  if (P->P.precision < 0 || P->P.precision > 2 || (P->P.vcycle != 0 && P->P.vcycle != 1)) {
    P->err = "params: precision must be 0 (FP64), 1 (mixed) or 2 (mixed, FP16 coarse matrices); vcycle 0 "
             "('Adapted') or 1 ('Standard')";
    *out = P.release();
    return T3X_ERR_INPUT;
  }
  if (nFix >= 0) {  // explicit boundary conditions: the arrays must be there
    long long nl = 0;
    bool ok = (nFix == 0 || (fixNodes && fixState)) && nCases >= 0 && (nCases == 0 || nPerCase);
    for (int ii = 0; ok && ii < nCases; ++ii) {
      ok = nPerCase[ii] >= 0;
      nl += nPerCase[ii];
    }
    if (!ok || (nl > 0 && (!loadNodes || !loadVec))) {
      P->err = "t3x_create: null or negative boundary-condition arguments";
      *out = P.release();
      return T3X_ERR_INPUT;
    }
  }
  int rc;
  try {
    if (nFix >= 0 && nCases < 1) {
      P->err = "No Loads!";
      rc = T3X_ERR_NO_LOADS;
    } else {
      rc = setup(*P, voxels, nFix, fixNodes, fixState, nCases, nPerCase, loadNodes, loadVec);
    }
  } catch (const HipError& e) {
    P->err = e.msg;
    rc = T3X_ERR_HIP;
  }
  *out = P.release();
  return rc;
}

// ================================================================ C ABI
extern "C" {

void t3x_default_params(t3x_params* p) {
  p->E0 = 1.0;
  p->nu = 0.3;
  p->Emin = 1.0e-6;
  p->penal = 3.0;
  p->tol = 1.0e-3;
  p->maxIT = 800;
  p->omega = 0.6;
  p->coarsestControl = 50000;
  p->precision = 0;
  p->checkEvery = 4;
  p->timing = 0;
  p->vcycle = 0;  // 'Adapted' (34)
  p->graphs = 0;
}

int t3x_create_multi(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                     const uint8_t* fixState, int nCases, const int* nPerCase, const int* loadNodes,
                     const double* loadVec, const t3x_params* params, t3x_problem** out) {
  return create_impl(voxels, nely, nelx, nelz, nFix, fixNodes, fixState, nCases, nPerCase, loadNodes, loadVec, params,
                     nullptr, out);
}

int t3x_create(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
               const uint8_t* fixState, int nLoad, const int* loadNodes, const double* loadVec,
               const t3x_params* params, t3x_problem** out) {
  return t3x_create_multi(voxels, nely, nelx, nelz, nFix, fixNodes, fixState, 1, &nLoad, loadNodes, loadVec, params,
                          out);
}

int t3x_num_load_cases(const t3x_problem* h) { return h ? h->nCases : 0; }

void t3x_destroy(t3x_problem* h) { delete h; }

const char* t3x_last_error(const t3x_problem* h) { return h ? h->err.c_str() : "null handle"; }

int t3x_info(const t3x_problem* h, int* info) {
  if (!h || !info) return T3X_ERR_INPUT;
  info[0] = h->L;
  for (int l = 0; l < h->L; ++l) {
    const Level& V = h->lv[l];
    int* o = info + 1 + 8 * l;
    o[0] = V.nx; o[1] = V.ny; o[2] = V.nz; o[3] = V.span; o[4] = V.nCells; o[5] = V.nNodes; o[6] = 0; o[7] = 0;
  }
  if (h->L > 0) info[1 + 8 * (h->L - 1) + 6] = h->cs.one_launch() ? 1 : 0;
  return T3X_OK;
}

int t3x_run(t3x_problem* h, int mode, double V0, int nLoop, double rMin, double rHat, t3x_result* out) {
  return guarded(h, [&] {
    if (mode != 0 && mode != 1) { h->err = "Wrong Input!"; return (int)T3X_ERR_INPUT; }
    t3x_result dummy{};
    return run(*h, mode, V0, nLoop, rMin, rHat, out ? out : &dummy);
  });
}

int t3x_opt_begin(t3x_problem* h, int mode, double V0, int nLoop, double rMin, double rHat) {
  return guarded(h, [&] {
    if (mode != 0 && mode != 1) { h->err = "Wrong Input!"; return (int)T3X_ERR_INPUT; }
    return opt_begin(*h, mode, V0, nLoop, rMin, rHat);
  });
}

int t3x_opt_step(t3x_problem* h, t3x_iter* it, int* finished) {
  return guarded(h, [&] {
    const int rc = opt_step(*h, it);
    if (rc < 0) return -rc;
    if (finished) *finished = rc;
    return (int)T3X_OK;
  });
}

int t3x_opt_result(t3x_problem* h, t3x_result* out) {
  return guarded(h, [&] { return opt_result(*h, out); });
}

int t3x_kernel_time(t3x_problem* h, int group, long long* count, double* ms) {
  return guarded(h, [&] {
    if (group < 0 || group >= TG_COUNT) return (int)T3X_ERR_INPUT;
    h->tflush(group);
    if (count) *count = h->timers[group].count;
    if (ms) *ms = h->timers[group].ms;
    return (int)T3X_OK;
  });
}

int t3x_footprint(const uint8_t* voxels, int nely, int nelx, int nelz, const t3x_params* params, int mode,
                  long long* bytes) {
  if (!bytes || nelx < 3 || nely < 3 || nelz < 3) return T3X_ERR_INPUT;
  t3x_params P;
  if (params) P = *params; else t3x_default_params(&P);
  long long nv = (long long)nely * nelx * nelz;
  if (voxels) {
    nv = 0;
    for (long long i = 0; i < (long long)nely * nelx * nelz; ++i) nv += voxels[i] ? 1 : 0;
  }
  const LevelDims f = level_dims(nely, nelx, nelz, nv, P.coarsestControl)[0];
  if (!plane_fits(f.ny + 1, f.nx + 1)) return T3X_ERR_INPUT;
  *bytes = footprint_estimate(nely, nelx, nelz, nv, P, mode);
  return T3X_OK;
}

int t3x_stream_copy(long long bytes, int reps, double* msPerCopy) {
  if (bytes < 16 || reps < 1 || !msPerCopy) return T3X_ERR_INPUT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return T3X_ERR_NO_DEVICE;
  const long long n = bytes / 16 * 2;  // doubles, a whole number of 16-byte pairs
  double *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = T3X_OK;
  try {
    a = dalloc<double>(n, s);
    b = dalloc<double>(n, s);
    T3X_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    T3X_HIP(hipEventCreate(&e0));
    T3X_HIP(hipEventCreate(&e1));
    launch_stream_copy(s, b, a, n);  // warm-up
    T3X_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) launch_stream_copy(s, (i & 1) ? a : b, (i & 1) ? b : a, n);
    T3X_HIP(hipEventRecord(e1, s));
    T3X_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    T3X_HIP(hipEventElapsedTime(&ms, e0, e1));
    *msPerCopy = ms / reps;
  } catch (const HipError&) {
    rc = T3X_ERR_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

namespace {

// Uploads the passive cells (h->hPassive) and the design cells (solid and not
// passive, 274-276).
int apply_passive(t3x_problem* h) {
  const Level& F = h->lv[0];
  std::vector<uint8_t> act(F.hEmask);
  int cnt = 0;
  for (long long d = 0; d < F.NN; ++d)
    if (h->hPassive[d]) {
      act[d] = 0;
      ++cnt;
    }
  if (!h->passive) {
    h->passive = h->alloc<uint8_t>(F.NN);
    h->active = h->alloc<uint8_t>(F.NN);
  }
  h->upload(h->passive, h->hPassive.data(), F.NN);
  h->upload(h->active, act.data(), F.NN);
  h->nPassive = cnt;
  return T3X_OK;
}

// One layer of Common_IncludeAdjacentElements (1587-1615): the 3x3x3
// neighbourhood of every listed cell, clipped to the grid and to solid cells.
std::vector<uint8_t> dilate26(const Level& F, const std::vector<uint8_t>& in) {
  std::vector<uint8_t> out(F.NN, 0);
  for (int iz = 0; iz < F.nz; ++iz)
    for (int ix = 0; ix < F.nx; ++ix)
      for (int iy = 0; iy < F.ny; ++iy) {
        const long long d = iy + (long long)F.NY * (ix + (long long)F.NX * iz);
        if (!in[d]) continue;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy) {
              const int y = iy + dy, x = ix + dx, z = iz + dz;
              if (y < 0 || x < 0 || z < 0 || y >= F.ny || x >= F.nx || z >= F.nz) continue;
              const long long e = y + (long long)F.NY * (x + (long long)F.NX * z);
              if (F.hEmask[e]) out[e] = 1;
            }
      }
  return out;
}

}  // namespace

int t3x_set_passive(t3x_problem* h, int n, const int* voxelIds) {
  return guarded(h, [&] {
    if (n < 0 || (n > 0 && !voxelIds)) { h->err = "t3x_set_passive: bad arguments"; return (int)T3X_ERR_INPUT; }
    const Level& F = h->lv[0];
    h->hPassive.assign(F.NN, 0);
    for (int i = 0; i < n; ++i) {
      const long long v = voxelIds[i];
      if (v < 0 || v >= (long long)h->nely * h->nelx * h->nelz) continue;
      const int iy = (int)(v % h->nely), ix = (int)((v / h->nely) % h->nelx), iz = (int)(v / ((long long)h->nely * h->nelx));
      const long long d = iy + (long long)F.NY * (ix + (long long)F.NX * iz);
      if (F.hEmask[d]) h->hPassive[d] = 1;  // void voxels are not elements (527)
    }
    return apply_passive(h);
  });
}

int t3x_set_passive_layers(t3x_problem* h, int numLayerBoundary, int numLayerLoads, int numLayerFixation) {
  return guarded(h, [&] {
    const Level& F = h->lv[0];
    const long long NN = F.NN;
    if (h->hPassive.size() != (size_t)NN) h->hPassive.assign(NN, 0);
    // elementsOnBoundary (1714-1728): solid cells with a node shared by fewer
    // than 8 solid cells
    std::vector<uint8_t> bnode(NN, 0), bcell(NN, 0);
    for (int iz = 0; iz < F.NZ; ++iz)
      for (int ix = 0; ix < F.NX; ++ix)
        for (int iy = 0; iy < F.NY; ++iy) {
          const long long d = iy + (long long)F.NY * (ix + (long long)F.NX * iz);
          if (!F.hNmask[d]) continue;
          int k = 0;
          for (int c = 0; c < 8; ++c) {
            const int y = iy - 1 + (c & 1), x = ix - 1 + ((c >> 1) & 1), z = iz - 1 + (c >> 2);
            if (y < 0 || x < 0 || z < 0 || y >= F.ny || x >= F.nx || z >= F.nz) continue;
            k += F.hEmask[y + (long long)F.NY * (x + (long long)F.NX * z)] ? 1 : 0;
          }
          bnode[d] = k < 8;
        }
    auto cell_touches = [&](long long d, const std::vector<uint8_t>& nodes) {
      for (int j = 0; j < 8; ++j)
        if (nodes[d + kLocDy[j] + (long long)F.NY * (kLocDx[j] + (long long)F.NX * kLocDz[j])]) return true;
      return false;
    };
    for (int iz = 0; iz < F.nz; ++iz)
      for (int ix = 0; ix < F.nx; ++ix)
        for (int iy = 0; iy < F.ny; ++iy) {
          const long long d = iy + (long long)F.NY * (ix + (long long)F.NX * iz);
          bcell[d] = F.hEmask[d] && cell_touches(d, bnode);
        }
    // layer 1 as given, every further layer one neighbourhood step (1521-1575)
    auto layers = [&](std::vector<uint8_t> set, int n) {
      for (int k = 2; k <= n; ++k) set = dilate26(F, set);
      return set;
    };
    auto merge = [&](const std::vector<uint8_t>& set) {
      for (long long d = 0; d < NN; ++d) h->hPassive[d] |= set[d];
    };
    if (numLayerBoundary > 0) merge(layers(bcell, numLayerBoundary));
    // elements near loads / fixations: the boundary cells around the loaded /
    // fixed nodes (nodeStruct_ of 1536-1546 lists boundary cells only)
    for (int pass = 0; pass < 2; ++pass) {
      const int n = pass == 0 ? numLayerLoads : numLayerFixation;
      if (n <= 0) continue;
      const std::vector<uint8_t>& nodes = pass == 0 ? h->hLoaded : h->hFixNode;
      std::vector<uint8_t> set(NN, 0);
      for (long long d = 0; d < NN; ++d) set[d] = bcell[d] && cell_touches(d, nodes);
      merge(layers(set, n));
    }
    return apply_passive(h);
  });
}

int t3x_get_passive(const t3x_problem* h, int cap, int* voxelIds, int* count) {
  if (!h || !count || h->lv.empty()) return T3X_ERR_INPUT;
  const Level& F = h->lv[0];
  int n = 0;
  if (h->hPassive.size() == (size_t)F.NN)
    for (int iz = 0; iz < h->nelz; ++iz)
      for (int ix = 0; ix < h->nelx; ++ix)
        for (int iy = 0; iy < h->nely; ++iy) {
          if (!h->hPassive[iy + (long long)F.NY * (ix + (long long)F.NX * iz)]) continue;
          if (voxelIds && n < cap) voxelIds[n] = iy + h->nely * (ix + h->nelx * iz);
          ++n;
        }
  *count = n;
  return T3X_OK;
}

int t3x_memory(const t3x_problem* h, long long* bytes) {
  if (!h || !bytes) return T3X_ERR_INPUT;
  *bytes = h->bytes;
  return T3X_OK;
}

int t3x_kernel_time_reset(t3x_problem* h) {
  return guarded(h, [&] {
    for (int g = 0; g < TG_COUNT; ++g) {
      h->tflush(g);
      h->timers[g].ms = 0.0;
      h->timers[g].count = 0;
    }
    return (int)T3X_OK;
  });
}

int t3x_set_modulus(t3x_problem* h, const double* E) {
  return guarded(h, [&] {
    upload_cells(*h, E, h->E);
    h->sync();
    return (int)T3X_OK;
  });
}

int t3x_assemble(t3x_problem* h) { return guarded(h, [&] { return assemble(*h); }); }

int t3x_apply_K(t3x_problem* h, const double* u, double* y) {
  return guarded(h, [&] {
    upload_nodal(*h, u, h->p);
    launch_apply_K(h->stream, geo(h->lv[0]), h->sf, h->E, h->fix, h->p, h->q, h->ws, h->sc, nullptr, false);
    download_nodal(*h, h->q, y);
    return (int)T3X_OK;
  });
}

int t3x_vcycle(t3x_problem* h, const double* r, double* x) {
  return guarded(h, [&] {
    if (!h->assembled) { h->err = "t3x_vcycle before t3x_assemble"; return (int)T3X_ERR_INPUT; }
    upload_nodal(*h, r, h->r);
    if (h->mixed()) {
      if (h->standard()) vcycle_standard(*h, h->r, h->z32, h->sc, nullptr, false);
      else vcycle_mixed(*h, h->r, h->z32, h->sc, nullptr, false);
      launch_convert(h->stream, h->z, h->z32, 3 * h->NN());
    } else {
      if (h->standard()) vcycle_standard64(*h, h->r, h->z, h->sc, nullptr, false);
      else vcycle(*h, h->r, h->z, h->sc, nullptr, false);
    }
    download_nodal(*h, h->z, x);
    if (h->cs.chain) {
      T3X_HIP(hipMemcpyAsync(h->hbuf + 255, &h->cs.chain->fail, sizeof(int), hipMemcpyDeviceToHost, h->stream));
      h->sync();
      check_chain(*h);
    }
    return (int)T3X_OK;
  });
}

int t3x_pcg(t3x_problem* h, const double* b, double* x, double tol, int maxIT, int* its, double* resnorm) {
  return guarded(h, [&] {
    if (!h->assembled) { h->err = "t3x_pcg before t3x_assemble"; return (int)T3X_ERR_INPUT; }
    if (!h->bvec) h->bvec = h->alloc<double>(3 * h->NN());
    upload_nodal(*h, b, h->bvec);
    upload_nodal(*h, x, h->U);
    bool warned = false;
    double res = 0.0;
    const int it = pcg_elastic(*h, h->bvec, h->U, tol, maxIT, &res, &warned);
    download_nodal(*h, h->U, x);
    if (its) *its = it;
    if (resnorm) *resnorm = res;
    return warned ? (int)T3X_WARN_MAXIT : (int)T3X_OK;
  });
}

int t3x_compliance(t3x_problem* h, const double* U, double* ce) {
  return guarded(h, [&] {
    upload_nodal(*h, U, h->p);
    launch_compliance_sf(h->stream, geo(h->lv[0]), h->sf, h->lv[0].emask, h->p, h->ce);
    download_cells(*h, h->ce, ce);
    return (int)T3X_OK;
  });
}

int t3x_get_nodal(t3x_problem* h, int which, int loadCase, double* out) {
  return guarded(h, [&] {
    if (!out || which < 0 || which > 1 || loadCase < 0 || loadCase >= h->nCases) {
      h->err = "t3x_get_nodal: which must be 0 (F) or 1 (U) and loadCase within the load cases";
      return (int)T3X_ERR_INPUT;
    }
    download_nodal(*h, which == 0 ? h->Fc[loadCase] : h->Uc[loadCase], out);
    return (int)T3X_OK;
  });
}

int t3x_filter_setup(t3x_problem* h, int slot, double radius) {
  return guarded(h, [&] {
    if (slot < 0 || slot > 1) return (int)T3X_ERR_INPUT;
    filter_setup(*h, slot, radius);
    h->sync();
    return (int)T3X_OK;
  });
}

int t3x_filter(t3x_problem* h, int slot, const double* src, double* dst, int* its) {
  return guarded(h, [&] {
    if (slot < 0 || slot > 1 || !h->filt[slot].ready) return (int)T3X_ERR_INPUT;
    upload_cells(*h, src, h->ca);
    const int it = filter_run(*h, slot, h->ca, h->cb);
    download_cells(*h, h->cb, dst);
    if (its) *its = it;
    return (int)T3X_OK;
  });
}

int t3x_oc(t3x_problem* h, double* x, const double* xPhys, const double* dc, double V0, double* change,
           int* nBisect) {
  return guarded(h, [&] {
    upload_cells(*h, x, h->ca);
    upload_cells(*h, xPhys, h->cb);
    upload_cells(*h, dc, h->cc);
    const double ch = oc_update(*h, h->ca, h->cb, h->cc, V0, nBisect);
    download_cells(*h, h->ca, x);
    if (change) *change = ch;
    return (int)T3X_OK;
  });
}

int t3x_mma(t3x_problem* h, int iter, double* xval, const double* xmin, const double* xmax, double* xold1,
            double* xold2, const double* dfdx, double gx, const double* dgdx, double* low, double* upp,
            int* nNewton) {
  return guarded(h, [&] {
    mma_alloc(*h);
    upload_cells(*h, xval, h->mx);
    upload_cells(*h, xmin, h->mxmin);
    upload_cells(*h, xmax, h->mxmax);
    upload_cells(*h, xold1, h->mxold1);
    upload_cells(*h, xold2, h->mxold2);
    upload_cells(*h, low, h->mlow);
    upload_cells(*h, upp, h->mupp);
    upload_cells(*h, dfdx, h->ca);
    upload_cells(*h, dgdx, h->cb);
    const int nn = mma_update(*h, iter, h->ca, gx, h->cb);
    download_cells(*h, h->mx, xval);
    download_cells(*h, h->mxold1, xold1);
    download_cells(*h, h->mxold2, xold2);
    download_cells(*h, h->mlow, low);
    download_cells(*h, h->mupp, upp);
    if (nNewton) *nNewton = nn;
    return (int)T3X_OK;
  });
}

}  // extern "C"

namespace t3x {
int create_pooled(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                  const uint8_t* fixState, int nLoad, const int* loadNodes, const double* loadVec,
                  const t3x_params* params, const PoolSlot& slot, t3x_problem** out) {
  return create_impl(voxels, nely, nelx, nelz, nFix, fixNodes, fixState, 1, &nLoad, loadNodes, loadVec, params,
                     &slot, out);
}
}  // namespace t3x
