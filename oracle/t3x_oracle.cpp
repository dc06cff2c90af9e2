// =============================================================================
// TEST INFRASTRUCTURE — NOT PART OF THE PRODUCT.
//
// FP64 CPU oracle: a restatement of the TOP3D_XL.m hot path (TO and PIO loop
// bodies and every callee) in C++/OpenMP.  Only the tests, the smoke check and
// the CPU-baseline leg of the benchmark load this library, and only as the
// checker.  The product (libt3x.so, HIP for gfx950) never links or calls it.
//
// It deliberately keeps the reference's own data structures (compact solid
// element/node numbering, eNodMat, elementUpwardMap, transferMat,
// transferMatCoeffi, solidNodeMapCoarser2Finer, per-element Galerkin Ks,
// sparse coarsest matrix) so that the GPU path, which works on a dense padded
// grid with masks and closed-form index math, is compared against an
// independent layout.
//
// Citations "N-M" are lines of TOP3D_XL.m.
//
// Pins (what checks this restatement; see tests/test_oracle_*.py):
//   Ke ............ independent 2x2x2 Gauss quadrature of B^T D B; rigid modes
//   hierarchy ..... brute-force masks/maps; R/P == dense tensor-product hat
//                   functions; partition of unity; linear-field reproduction
//   K*u ........... scipy-assembled global K with the BC modification
//   Galerkin Ks ... dense P^T K_BC P on small grids; diagK; coarsest K
//   coarse solve .. numpy.linalg.solve
//   V-cycle ....... dense operator 2wD^-1 + P(...)R, symmetric positive def.
//   PCG ........... independent numpy PCG on the dense operators
//   compliance .... sum E ce == u^T K u
//   PDE filter .... scipy-assembled Helmholtz system, constants, self-adjoint
//   OC ............ numpy brute force of the bisection
//   Heaviside ..... closed forms (H(0) = 0, H(1) = 1, H(eta) = 1/2, slope at
//                   eta), beta -> 0 / infinity limits, central differences
//   p-norm ........ closed forms (uniform, one non-zero entry), degree-one
//                   homogeneity, central differences of the gradient
//   LOCAL chain ... dfdx and dc (330-350) against central differences of the
//                   exact composition (dense filters, direct elastic solve)
//   MMA ........... first-order agreement of the GenSub model with f and g at
//                   xval, hand-computed asymptotes, KKT residual of the
//                   returned (x, lambda), an independent nested-bisection
//                   solve of the same subproblem, convergence to the analytic
//                   optimum of a separable convex problem
// (tests/test_oracle_pins.py, tests/test_oracle_local_pins.py).  Every
// function of this file is covered by one of those checks.
//
// The only parts not pinnable without MATLAB: bitwise order of MATLAB's
// multithreaded reductions / accumarray and CHOLMOD's ordering.  Those affect
// results at the 1e-15 relative level only.
// =============================================================================

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

namespace {

using i64 = long long;

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---------------------------------------------------------------- settings
// InitialSettings, 21-38.
struct Settings {
  double E0 = 1.0;          // modulus_        24
  double nu = 0.3;          // poissonRatio_   25
  double Emin = 1.0e-6;     // modulusMin_     26
  double penal = 3.0;       // SIMPpenalty_    27
  double cell = 1.0;        // cellSize_       28
  double tol = 1.0e-3;      // tol_            30
  int maxIT = 800;          // maxIT_          31
  double omega = 0.6;       // weightFactorJacobi_ 32
  int coarsestControl = 50000;  // 33
  int nonDyadic = 1;        // 35
  int typeVcycle = 0;       // typeVcycle_ 34: 0 'Adapted', 1 'Standard'
};

// ------------------------------------------------------- deterministic sums
// Fixed-block summation: result independent of the OpenMP thread count.
constexpr i64 kSumBlock = 4096;
template <class F>
double det_sum(i64 n, F f) {
  i64 nb = (n + kSumBlock - 1) / kSumBlock;
  std::vector<double> part(std::max<i64>(nb, 1), 0.0);
#pragma omp parallel for schedule(static)
  for (i64 b = 0; b < nb; ++b) {
    double s = 0.0;
    i64 e = std::min(n, (b + 1) * kSumBlock);
    for (i64 i = b * kSumBlock; i < e; ++i) s += f(i);
    part[b] = s;
  }
  double s = 0.0;
  for (i64 b = 0; b < nb; ++b) s += part[b];
  return s;
}
double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  return det_sum((i64)a.size(), [&](i64 i) { return a[i] * b[i]; });
}
double normv(const std::vector<double>& a) { return std::sqrt(dotv(a, a)); }

// --------------------------------------------------------- element stiffness
// FEA_VoxelBasedElementStiffnessMatrix, 1254-1292: closed form for the unit
// hex8.  The twelve 6x6 blocks are stored as signed indices into C (1-based,
// 0 = zero entry), row-major.
const double kC[9] = {2.0 / 9, 1.0 / 18, 1.0 / 24, 1.0 / 36, 1.0 / 48,
                      5.0 / 72, 1.0 / 3, 1.0 / 6, 1.0 / 12};  // 1258
const signed char kA11[36] = {-1, -3, -3, 2, 3, 3,  -3, -1, -3, -3, -4, -5,
                              -3, -3, -1, -3, -5, -4, 2, -3, -3, -1, 3, 3,
                              3, -4, -5, 3, -1, -3,  3, -5, -4, 3, -3, -1};
const signed char kB11[36] = {7, 0, 0, 0, -8, -8, 0, 7, 0, 8, 0, 0,
                              0, 0, 7, 8, 0, 0,   0, 8, 8, 7, 0, 0,
                              -8, 0, 0, 0, 7, 0,  -8, 0, 0, 0, 0, 7};
const signed char kA22[36] = {-1, -3, 3, 2, 3, -3,  -3, -1, 3, -3, -4, 5,
                              3, 3, -1, 3, 5, -4,   2, -3, 3, -1, 3, -3,
                              3, -4, 5, 3, -1, 3,   -3, 5, -4, -3, 3, -1};
const signed char kB22[36] = {7, 0, 0, 0, -8, 8, 0, 7, 0, 8, 0, 0,
                              0, 0, 7, -8, 0, 0, 0, 8, -8, 7, 0, 0,
                              -8, 0, 0, 0, 7, 0, 8, 0, 0, 0, 0, 7};
const signed char kA12[36] = {6, 3, 5, -4, -3, -5,  3, 6, 5, 3, 2, 3,
                              -5, -5, 4, -5, -3, -4, -4, 3, 5, 6, -3, -5,
                              -3, 2, 3, -3, 6, 5,   5, -3, -4, 5, -5, 4};
const signed char kB12[36] = {-9, 0, -9, 0, 8, 0,  0, -9, -9, -8, 0, -8,
                              9, 9, -9, 0, 8, 0,   0, -8, 0, -9, 0, 9,
                              8, 0, -8, 0, -9, -9, 0, 8, 0, -9, 9, -9};
const signed char kA13[36] = {-4, -5, -3, 6, 5, 3,  -5, -4, -3, -5, 4, -5,
                              3, 3, 2, 3, 5, 6,     6, -5, -3, -4, 5, 3,
                              5, 4, -5, 5, -4, -3,  -3, 5, 6, -3, 3, 2};
const signed char kB13[36] = {0, 0, 8, -9, -9, 0,  0, 0, 8, 9, -9, 9,
                              -8, -8, 0, 0, -9, -9, -9, 9, 0, 0, 0, -8,
                              -9, -9, 9, 0, 0, 8,  0, -9, -9, 8, -8, 0};
const signed char kA14[36] = {2, 5, 5, 4, -5, -5,  5, 2, 5, 5, 6, 3,
                              5, 5, 2, 5, 3, 6,    4, 5, 5, 2, -5, -5,
                              -5, 6, 3, -5, 2, 5,  -5, 3, 6, -5, 5, 2};
const signed char kB14[36] = {-9, 0, 0, -9, 9, 9,  0, -9, 0, -9, -9, 0,
                              0, 0, -9, -9, 0, -9, -9, -9, -9, -9, 0, 0,
                              9, -9, 0, 0, -9, 0,  9, 0, -9, 0, 0, -9};
const signed char kA23[36] = {2, 5, -5, 4, -5, 5,  5, 2, -5, 5, 6, -3,
                              -5, -5, 2, -5, -3, 6, 4, 5, -5, 2, -5, 5,
                              -5, 6, -3, -5, 2, -5, 5, -3, 6, 5, -5, 2};
const signed char kB23[36] = {-9, 0, 0, -9, 9, -9, 0, -9, 0, -9, -9, 0,
                              0, 0, -9, 9, 0, -9,  -9, -9, 9, -9, 0, 0,
                              9, -9, 0, 0, -9, 0,  -9, 0, -9, 0, 0, -9};

double cval(signed char code) {
  if (code == 0) return 0.0;
  return code > 0 ? kC[code - 1] : -kC[-code - 1];
}

// Entry (r,c) of a 6x6 block, optionally transposed.
double blk(const signed char* m, int r, int c, bool tr) {
  return tr ? cval(m[c * 6 + r]) : cval(m[r * 6 + c]);
}

// Returns Ke as 576 doubles, column-major (== row-major: Ke is symmetric).
std::vector<double> element_stiffness(const Settings& S) {
  // Block layout (1289-1290):
  //  [A11  A12  A13  A14 ;  A12' A22 A23 A13' ; A13' A23' A22 A12' ; A14' A13 A12 A11]
  struct Ref { const signed char* a; const signed char* b; bool tr; };
  const Ref lay[4][4] = {
      {{kA11, kB11, false}, {kA12, kB12, false}, {kA13, kB13, false}, {kA14, kB14, false}},
      {{kA12, kB12, true}, {kA22, kB22, false}, {kA23, kB23, false}, {kA13, kB13, true}},
      {{kA13, kB13, true}, {kA23, kB23, true}, {kA22, kB22, false}, {kA12, kB12, true}},
      {{kA14, kB14, true}, {kA13, kB13, false}, {kA12, kB12, false}, {kA11, kB11, false}}};
  const double nu = S.nu;
  const double f = 1.0 / (1.0 + nu) / (2.0 * nu - 1.0);
  std::vector<double> K(576);
  for (int I = 0; I < 4; ++I)
    for (int J = 0; J < 4; ++J)
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          const Ref& R = lay[I][J];
          double v = f * (blk(R.a, r, c, R.tr) + nu * blk(R.b, r, c, R.tr));
          K[(J * 6 + c) * 24 + (I * 6 + r)] = v * S.cell;  // 1291
        }
  return K;
}

// Solving_TrilinearInterpolation, 1785-1798: N(s,t,p) for the 8 local nodes.
void trilinear(double s, double t, double p, double* N) {
  N[0] = 0.125 * (1 - s) * (1 - t) * (1 - p);
  N[1] = 0.125 * (1 + s) * (1 - t) * (1 - p);
  N[2] = 0.125 * (1 + s) * (1 + t) * (1 - p);
  N[3] = 0.125 * (1 - s) * (1 + t) * (1 - p);
  N[4] = 0.125 * (1 - s) * (1 - t) * (1 + p);
  N[5] = 0.125 * (1 + s) * (1 - t) * (1 + p);
  N[6] = 0.125 * (1 + s) * (1 + t) * (1 + p);
  N[7] = 0.125 * (1 - s) * (1 + t) * (1 + p);
}

// Solving_Operator4MultiGridRestrictionAndInterpolation('inNODE'), 1768-1783:
// rows are the (span+1)^3 fine nodes of a coarse cell in column-major
// (y,x,z) order with y running from t=+1 down to t=-1.
std::vector<double> mg_operator(int span) {
  int m = span + 1;
  std::vector<double> RI((size_t)m * m * m * 8);
  for (int lz = 0; lz < m; ++lz)
    for (int lx = 0; lx < m; ++lx)
      for (int ly = 0; ly < m; ++ly) {
        double sy = 1.0 - 2.0 * ly / span;   // linspace(1,-1,span+1)
        double sx = -1.0 + 2.0 * lx / span;  // linspace(-1,1,span+1)
        double sz = -1.0 + 2.0 * lz / span;
        int a = ly + m * (lx + m * lz);
        trilinear(sx, sy, sz, &RI[(size_t)a * 8]);
      }
  return RI;
}

// ----------------------------------------------------------- mesh structures
// Data_CartesianMeshStruct, 1732-1766 (the fields the hot path uses).
struct Level {
  int resX = 0, resY = 0, resZ = 0, span = 0;
  double eleSize = 1.0;
  int numElements = 0, numNodes = 0;
  std::vector<int> eNodMat;        // numElements*8: compact node ids
  std::vector<int> eleMapBack;     // compact -> dense element id of this grid
  std::vector<int> eleMapForward;  // dense -> compact+1 (0 = void)
  std::vector<int> nodMapBack;     // compact -> dense node id of this grid
  std::vector<int> nodMapForward;  // dense -> compact+1
  std::vector<int> numNod2Eles;
  std::vector<int> nodesOnBoundary, elementsOnBoundary;
  // level >= 2 transfer data
  std::vector<int> elementUpwardMap;      // numElements*span^3 : finer compact+1 / 0
  std::vector<int> transferMat;           // numElements*(span+1)^3 : intermediate id
  std::vector<double> transferMatCoeffi;  // intermediateNumNodes
  std::vector<int> solidNodeMapCoarser2Finer;  // finer compact node -> intermediate id
  int intermediateNumNodes = 0;
  std::vector<double> RI;     // (span+1)^3 x 8
  std::vector<double> diagK;  // 3*numNodes
  std::vector<double> Ks;     // numElements*576 (column-major 24x24)
  // node -> incidence list (e*8+j), ordered as accumarray(eNodMat(:)) visits
  std::vector<int> n2eStart, n2eList;
  // intermediate node -> (e*(span+1)^3 + a) list for prolongation
  std::vector<int> i2tStart, i2tList;
  int nY() const { return resY + 1; }
  int nX() const { return resX + 1; }
  int nZ() const { return resZ + 1; }
};

// Dense element (iy,ix,iz) -> its 8 dense node ids (1693-1701 offsets).
inline void dense_elem_nodes(int resY, int resX, int iy, int ix, int iz, int* n) {
  const i64 NY = resY + 1, NX = resX + 1;
  i64 base = (iy + 1) + NY * (ix + NX * (i64)iz);  // local node 1
  const i64 off[8] = {0, NY, NY - 1, -1, NY * NX, NY * NX + NY, NY * NX + NY - 1, NY * NX - 1};
  for (int j = 0; j < 8; ++j) n[j] = (int)(base + off[j]);
}

void build_incidence(Level& L) {
  L.n2eStart.assign(L.numNodes + 1, 0);
  for (i64 k = 0; k < (i64)L.numElements * 8; ++k) L.n2eStart[L.eNodMat[k] + 1]++;
  for (int i = 0; i < L.numNodes; ++i) L.n2eStart[i + 1] += L.n2eStart[i];
  L.n2eList.assign((size_t)L.numElements * 8, 0);
  std::vector<int> fill(L.n2eStart.begin(), L.n2eStart.end() - 1);
  for (int j = 0; j < 8; ++j)  // accumarray(eNodMat(:)) order: j outer, e inner
    for (int e = 0; e < L.numElements; ++e)
      L.n2eList[fill[L.eNodMat[(size_t)e * 8 + j]]++] = e * 8 + j;
}

// ----------------------------------------------------------------- the model
struct Filter {
  double kernel[64];               // iKEF, 8x8 column-major
  std::vector<double> diagPrecond; // numNodes (already inverted)
};

struct Model {
  Settings S;
  int nely = 0, nelx = 0, nelz = 0;  // unpadded voxel dims
  int numLevels = 0;
  std::vector<Level> L;
  std::vector<double> Ke;  // 576
  // boundary conditions (compact finest node ids, 0-based)
  std::vector<int> fixNode;               // sorted unique
  std::vector<unsigned char> fixState;    // 3 per fixed node
  // load cases (loadingCond_{ii}, 447, 481-493): nodes and 3 values per node
  std::vector<std::vector<int>> loadNode;
  std::vector<std::vector<double>> loadVal;
  std::vector<std::vector<double>> F;     // F_ columns, 3*numNodes each (1315-1321)
  std::vector<double> objWeight;          // objWeightingList_ (493, 558)
  std::vector<char> isPassive;            // per compact element (passiveElements_, 81 / 274)
  bool bcOk = false;
  // BC element matrices, 1338-1437
  std::vector<int> mapUniqueKes;          // numElements: 0 or rank+1
  std::vector<double> KesFree, KesFixed;  // nUnique*576
  std::vector<int> l2BC;                  // L2 numElements: 0 or rank+1
  std::vector<std::vector<int>> sonElesWithFixed;  // per rank: local positions
  // per-iteration state
  std::vector<double> eleModulus;  // finest numElements
  // coarsest factor (banded Cholesky with our own node ordering)
  int cN = 0, cBW = 0;
  int orderVariant = 0;           // 0: smallest-bandwidth axis order; k: the k-th best
  std::vector<int> cPerm;         // permuted dof -> compact dof
  std::vector<double> cBand;      // cN*(cBW+1): L(i, i-k) at i*(cBW+1)+k
  bool cholOk = false;
  // PDE filters
  Filter filt[2];
  // workspace
  mutable std::vector<double> scratchK;
  std::string err;
  int warnings = 0;
};

// ------------------------------------------------------ discretisation (L1)
// FEA_VoxelBasedDiscretization, 1629-1730.
bool discretize(Model& M, const unsigned char* vox) {
  const int nely = M.nely, nelx = M.nelx, nelz = M.nelz;
  i64 numVoxels = 0;
  for (i64 i = 0; i < (i64)nely * nelx * nelz; ++i) numVoxels += vox[i] ? 1 : 0;
  int nl = 0;
  while (numVoxels >= M.S.coarsestControl) {  // 1654-1657
    nl++;
    numVoxels = (i64)std::llround((double)numVoxels / 8.0);
  }
  nl = std::max(3, nl);  // 1658
  const int pw = 1 << nl;
  const int ax = (nelx + pw - 1) / pw * pw;  // 1659-1661
  const int ay = (nely + pw - 1) / pw * pw;
  const int az = (nelz + pw - 1) / pw * pw;
  M.numLevels = nl + 1;  // 1662
  M.L.clear();
  M.L.emplace_back();
  Level& L = M.L[0];
  L.resX = ax; L.resY = ay; L.resZ = az; L.span = 0; L.eleSize = 1.0;  // 1674-1681
  const i64 nE = (i64)ax * ay * az;
  L.eleMapForward.assign(nE, 0);
  for (int iz = 0; iz < nelz; ++iz)
    for (int ix = 0; ix < nelx; ++ix)
      for (int iy = 0; iy < nely; ++iy)
        if (vox[iy + (i64)nely * (ix + (i64)nelx * iz)])
          L.eleMapForward[iy + (i64)ay * (ix + (i64)ax * iz)] = 1;
  for (i64 e = 0; e < nE; ++e)  // 1685-1689
    if (L.eleMapForward[e]) {
      L.eleMapBack.push_back((int)e);
      L.eleMapForward[e] = (int)L.eleMapBack.size();
    }
  L.numElements = (int)L.eleMapBack.size();
  // nodes 1693-1711
  const i64 nN = (i64)(ay + 1) * (ax + 1) * (az + 1);
  L.nodMapForward.assign(nN, 0);
  std::vector<int> en((size_t)L.numElements * 8);
  for (int e = 0; e < L.numElements; ++e) {
    int d = L.eleMapBack[e];
    int iy = d % ay, ix = (d / ay) % ax, iz = d / ay / ax;
    dense_elem_nodes(ay, ax, iy, ix, iz, &en[(size_t)e * 8]);
    for (int j = 0; j < 8; ++j) L.nodMapForward[en[(size_t)e * 8 + j]] = 1;
  }
  for (i64 n = 0; n < nN; ++n)
    if (L.nodMapForward[n]) {
      L.nodMapBack.push_back((int)n);
      L.nodMapForward[n] = (int)L.nodMapBack.size();
    }
  L.numNodes = (int)L.nodMapBack.size();
  L.eNodMat.resize(en.size());
  for (size_t k = 0; k < en.size(); ++k) L.eNodMat[k] = L.nodMapForward[en[k]] - 1;
  // boundary info 1714-1728
  L.numNod2Eles.assign(L.numNodes, 0);
  for (size_t k = 0; k < L.eNodMat.size(); ++k) L.numNod2Eles[L.eNodMat[k]]++;
  std::vector<char> onB(L.numNodes, 0);
  for (int n = 0; n < L.numNodes; ++n)
    if (L.numNod2Eles[n] < 8) { L.nodesOnBoundary.push_back(n); onB[n] = 1; }
  for (int e = 0; e < L.numElements; ++e) {
    int t = 0;
    for (int j = 0; j < 8; ++j) t += onB[L.eNodMat[(size_t)e * 8 + j]];
    if (t > 0) L.elementsOnBoundary.push_back(e);
  }
  build_incidence(L);
  return true;
}

// ------------------------------------------------------------ MG hierarchy
// Solving_BuildingMeshHierarchy, 1076-1228.
void build_hierarchy(Model& M) {
  if (M.S.nonDyadic == 1 && M.numLevels >= 4) M.numLevels -= 1;  // 1085-1089
  else M.S.nonDyadic = 0;
  for (int ii = 1; ii < M.numLevels; ++ii) {
    const int span = (ii == 1 && M.S.nonDyadic == 1) ? 4 : 2;  // 1092-1096
    Level& Fn = M.L[ii - 1];
    Level C;
    C.resX = Fn.resX / span; C.resY = Fn.resY / span; C.resZ = Fn.resZ / span;
    C.eleSize = Fn.eleSize * span;
    C.span = span;
    const int nx = C.resX, ny = C.resY, nz = C.resZ;
    const int s3 = span * span * span, m = span + 1, m3 = m * m * m;
    const int FY = Fn.resY, FX = Fn.resX;
    const i64 FNY = FY + 1, FNX = FX + 1;
    // 3.1 raw maps (1111-1139)
    std::vector<int> upAll((size_t)nx * ny * nz * s3);
    std::vector<char> nonEmpty((size_t)nx * ny * nz, 0);
    for (int gz = 0; gz < nz; ++gz)
      for (int gx = 0; gx < nx; ++gx)
        for (int gy = 0; gy < ny; ++gy) {
          i64 ce = gy + (i64)ny * (gx + (i64)nx * gz);
          for (int lz = 0; lz < span; ++lz)
            for (int lx = 0; lx < span; ++lx)
              for (int ly = 0; ly < span; ++ly) {
                i64 fd = (span * gy + ly) + (i64)FY * ((span * gx + lx) + (i64)FX * (span * gz + lz));
                int v = Fn.eleMapForward[fd];
                upAll[ce * s3 + ly + span * (lx + span * lz)] = v;
                if (v) nonEmpty[ce] = 1;
              }
        }
    // unemptyElements (1152), intermediate nodes (1153-1158)
    for (i64 ce = 0; ce < (i64)nx * ny * nz; ++ce)
      if (nonEmpty[ce]) C.eleMapBack.push_back((int)ce);
    C.numElements = (int)C.eleMapBack.size();
    const i64 fineDenseNodes = FNY * FNX * (Fn.resZ + 1);
    std::vector<int> rank(fineDenseNodes, 0);
    auto blockNode = [&](int ce, int a) -> i64 {
      int gy = ce % ny, gx = (ce / ny) % nx, gz = ce / ny / nx;
      int ly = a % m, lx = (a / m) % m, lz = a / m / m;
      return (span * gy + ly) + FNY * ((span * gx + lx) + FNX * (i64)(span * gz + lz));
    };
    for (int e = 0; e < C.numElements; ++e)
      for (int a = 0; a < m3; ++a) rank[blockNode(C.eleMapBack[e], a)] = 1;
    int cnt = 0;
    for (i64 n = 0; n < fineDenseNodes; ++n)
      if (rank[n]) rank[n] = ++cnt;
    C.intermediateNumNodes = cnt;
    // transferMat (1159-1162), transferMatCoeffi (1163-1167)
    C.transferMat.resize((size_t)C.numElements * m3);
    C.transferMatCoeffi.assign(cnt, 0.0);
    for (int e = 0; e < C.numElements; ++e)
      for (int a = 0; a < m3; ++a) {
        int k = rank[blockNode(C.eleMapBack[e], a)] - 1;
        C.transferMat[(size_t)e * m3 + a] = k;
        C.transferMatCoeffi[k] += 1.0;
      }
    // solidNodeMapCoarser2Finer (1168-1173): finer compact node -> intermediate
    C.solidNodeMapCoarser2Finer.resize(Fn.numNodes);
    for (int k = 0; k < Fn.numNodes; ++k)
      C.solidNodeMapCoarser2Finer[k] = rank[Fn.nodMapBack[k]] - 1;
    // 3.3 (1176-1181)
    C.eleMapForward.assign((size_t)nx * ny * nz, 0);
    for (int e = 0; e < C.numElements; ++e) C.eleMapForward[C.eleMapBack[e]] = e + 1;
    C.elementUpwardMap.resize((size_t)C.numElements * s3);
    for (int e = 0; e < C.numElements; ++e)
      std::memcpy(&C.elementUpwardMap[(size_t)e * s3], &upAll[(size_t)C.eleMapBack[e] * s3],
                  sizeof(int) * s3);
    // 4. discretize (1184-1205)
    const i64 nN = (i64)(ny + 1) * (nx + 1) * (nz + 1);
    C.nodMapForward.assign(nN, 0);
    std::vector<int> en((size_t)C.numElements * 8);
    for (int e = 0; e < C.numElements; ++e) {
      int d = C.eleMapBack[e];
      dense_elem_nodes(ny, nx, d % ny, (d / ny) % nx, d / ny / nx, &en[(size_t)e * 8]);
      for (int j = 0; j < 8; ++j) C.nodMapForward[en[(size_t)e * 8 + j]] = 1;
    }
    for (i64 n = 0; n < nN; ++n)
      if (C.nodMapForward[n]) {
        C.nodMapBack.push_back((int)n);
        C.nodMapForward[n] = (int)C.nodMapBack.size();
      }
    C.numNodes = (int)C.nodMapBack.size();
    C.eNodMat.resize(en.size());
    for (size_t k = 0; k < en.size(); ++k) C.eNodMat[k] = C.nodMapForward[en[k]] - 1;
    // 5. R/P operator (1208)
    C.RI = mg_operator(span);
    build_incidence(C);
    // prolongation incidence: intermediate node <- (e, a), order of transferMat(:)
    C.i2tStart.assign(cnt + 1, 0);
    for (size_t k = 0; k < C.transferMat.size(); ++k) C.i2tStart[C.transferMat[k] + 1]++;
    for (int i = 0; i < cnt; ++i) C.i2tStart[i + 1] += C.i2tStart[i];
    C.i2tList.resize(C.transferMat.size());
    std::vector<int> fill(C.i2tStart.begin(), C.i2tStart.end() - 1);
    for (size_t k = 0; k < C.transferMat.size(); ++k) C.i2tList[fill[C.transferMat[k]]++] = (int)k;
    M.L.push_back(std::move(C));
  }
}

// ------------------------------------------------------ boundary conditions
// Built-in cuboid BCs, 536-559: fix the x=1 node face; optLoad=4 load.
void cuboid_bc(Model& M, std::vector<int>& fixDense, std::vector<int>& loadDense) {
  const Level& L = M.L[0];
  const i64 NY = L.resY + 1, NX = L.resX + 1;
  for (int jz = 0; jz <= M.nelz; ++jz)  // 538: (1:nely+1, 1, 1:nelz+1)
    for (int jy = 0; jy <= M.nely; ++jy) fixDense.push_back((int)(jy + NY * (0 + NX * jz)));
  // 552: (1:nely+1, nelx+1, 1:round(nelz/6+1))
  int zTop = (int)std::llround(M.nelz / 6.0 + 1.0);
  for (int jz = 0; jz < zTop; ++jz)
    for (int jy = 0; jy <= M.nely; ++jy) loadDense.push_back((int)(jy + NY * (M.nelx + NX * (i64)jz)));
  std::sort(fixDense.begin(), fixDense.end());
  std::sort(loadDense.begin(), loadDense.end());
}

// FEA_ApplyBoundaryCondition, 1294-1336, over all load cases.
bool apply_bc(Model& M) {
  for (size_t ii = 0; ii < M.loadNode.size(); ++ii) {  // 1301-1312
    // 1303-1306: nodes loaded in this case are not fixed; unique+sort
    std::vector<int> fn; std::vector<unsigned char> fs;
    {
      std::vector<std::pair<int, int>> tmp;
      for (size_t i = 0; i < M.fixNode.size(); ++i) tmp.push_back({M.fixNode[i], (int)i});
      std::sort(tmp.begin(), tmp.end());
      std::vector<int> ln = M.loadNode[ii]; std::sort(ln.begin(), ln.end());
      int last = -1;
      for (auto& p : tmp) {
        if (p.first == last) continue;
        last = p.first;
        if (std::binary_search(ln.begin(), ln.end(), p.first)) continue;
        fn.push_back(p.first);
        for (int c = 0; c < 3; ++c) fs.push_back(M.fixState[(size_t)p.second * 3 + c]);
      }
    }
    M.fixNode = fn; M.fixState = fs;
    {  // 1307-1308: a node listed twice keeps its first load row
      std::vector<std::pair<int, int>> tmp;
      for (size_t i = 0; i < M.loadNode[ii].size(); ++i) tmp.push_back({M.loadNode[ii][i], (int)i});
      std::sort(tmp.begin(), tmp.end());
      std::vector<int> ln; std::vector<double> lv;
      int last = -1;
      for (auto& p : tmp) {
        if (p.first == last) continue;
        last = p.first;
        ln.push_back(p.first);
        for (int c = 0; c < 3; ++c) lv.push_back(M.loadVal[ii][(size_t)p.second * 3 + c]);
      }
      M.loadNode[ii] = ln; M.loadVal[ii] = lv;
    }
    if (M.loadNode[ii].empty()) { M.err = "No Loads!"; return false; }    // 1309
    if (M.fixNode.empty()) { M.err = "No Fixations!"; return false; }     // 1310
  }
  const Level& L = M.L[0];
  M.F.assign(M.loadNode.size(), std::vector<double>((size_t)L.numNodes * 3, 0.0));  // 1315-1321
  for (size_t ii = 0; ii < M.loadNode.size(); ++ii)
    for (size_t i = 0; i < M.loadNode[ii].size(); ++i)
      for (int c = 0; c < 3; ++c) M.F[ii][(size_t)M.loadNode[ii][i] * 3 + c] = M.loadVal[ii][i * 3 + c];
  return true;
}

// Solving_SetupKeWithFixedDOFs + Solving_ApplyBConEleStiffMat_B, 1338-1437.
bool setup_bc_ke(Model& M) {
  const Level& L = M.L[0];
  std::vector<char> isBE(L.numElements, 0);
  for (int e : L.elementsOnBoundary) isBE[e] = 1;
  // element lists per boundary node, ascending element order (1356-1366)
  std::vector<std::vector<int>> nodeEles(L.numNodes);
  std::vector<char> onB(L.numNodes, 0);
  for (int n : L.nodesOnBoundary) onB[n] = 1;
  for (int e = 0; e < L.numElements; ++e)
    if (isBE[e])
      for (int j = 0; j < 8; ++j) {
        int n = L.eNodMat[(size_t)e * 8 + j];
        if (onB[n]) nodeEles[n].push_back(e);
      }
  // 1369-1374.  Fixed nodes must be boundary nodes: the reference intersects
  // fixingCond_ with nodesOnBoundary but then pairs the ii-th common node's
  // element list with fixingCond_'s ii-th row (1379-1383), so a fixation
  // inside the domain both goes missing and shifts every later pairing.
  // That case is refused here rather than restated.
  std::vector<int> allFixedEles;
  for (int n : M.fixNode) {
    if (nodeEles[n].empty()) {
      M.err = "fixation at a node inside the domain (not on its boundary): not restated";
      return false;
    }
    for (int e : nodeEles[n]) allFixedEles.push_back(e);
  }
  std::sort(allFixedEles.begin(), allFixedEles.end());
  allFixedEles.erase(std::unique(allFixedEles.begin(), allFixedEles.end()), allFixedEles.end());
  M.mapUniqueKes.assign(L.numElements, 0);
  for (size_t r = 0; r < allFixedEles.size(); ++r) M.mapUniqueKes[allFixedEles[r]] = (int)r + 1;
  const size_t nU = allFixedEles.size();
  M.KesFree.resize(nU * 576);
  M.KesFixed.assign(nU * 576, 0.0);
  for (size_t u = 0; u < nU; ++u) std::memcpy(&M.KesFree[u * 576], M.Ke.data(), 576 * sizeof(double));
  for (size_t ii = 0; ii < M.fixNode.size(); ++ii) {  // 1379-1395
    int node = M.fixNode[ii];
    const std::vector<int>& eles = nodeEles[node];
    const double nAdj = (double)eles.size();
    for (int e : eles) {
      int jl = -1;
      for (int j = 0; j < 8; ++j)
        if (L.eNodMat[(size_t)e * 8 + j] == node) jl = j;
      size_t u = M.mapUniqueKes[e] - 1;
      double* Fr = &M.KesFree[u * 576];
      double* Fx = &M.KesFixed[u * 576];
      for (int c = 0; c < 3; ++c) {
        if (!M.fixState[ii * 3 + c]) continue;
        int d = 3 * jl + c;
        for (int k = 0; k < 24; ++k) { Fr[d * 24 + k] = 0.0; Fr[k * 24 + d] = 0.0; }  // 1430-1431
        Fx[d * 24 + d] = 1.0 / nAdj;  // 1433
      }
    }
  }
  for (double& v : M.KesFixed) v = v * M.S.E0 * M.S.cell;  // 1396
  // level-2 bookkeeping 1397-1422
  const Level& C = M.L[1];
  const int s3 = C.span * C.span * C.span;
  M.l2BC.assign(C.numElements, 0);
  M.sonElesWithFixed.clear();
  for (int jj = 0; jj < C.numElements; ++jj) {
    std::vector<int> loc;
    for (int ss = 0; ss < s3; ++ss) {
      int v = C.elementUpwardMap[(size_t)jj * s3 + ss];
      if (v && M.mapUniqueKes[v - 1]) loc.push_back(ss);
    }
    if (!loc.empty()) {
      M.sonElesWithFixed.push_back(loc);
      M.l2BC[jj] = (int)M.sonElesWithFixed.size();
    }
  }
  return true;
}

// ------------------------------------------------------- Galerkin assembly
// Local dof index inside a (span+1)^3 block for sub-element ss, local dof k
// (Solving_SubEleNodMat 1617-1627 + the reOrdering at 925/931-932).
void sub_elem_dofs(int span, int ss, int* dof) {
  const int m = span + 1;
  int ly = ss % span, lx = (ss / span) % span, lz = ss / span / span;
  const int off[8][3] = {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 0},
                         {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {0, 0, 1}};  // (dy,dx,dz)
  for (int j = 0; j < 8; ++j) {
    int a = (ly + off[j][0]) + m * ((lx + off[j][1]) + m * (lz + off[j][2]));
    for (int c = 0; c < 3; ++c) dof[3 * j + c] = 3 * a + c;
  }
}

// out(24x24) = P^T * K * P with K dense nd x nd (column-major) and
// P = kron(RI, I3) (nd x 24).  Matches interpolatingKe' * tmpK * interpolatingKe.
void project(int span, const std::vector<double>& RI, const double* K, int nd, double* out) {
  std::vector<double> KP((size_t)nd * 24, 0.0);  // K*P, column-major nd x 24
  for (int q = 0; q < 24; ++q) {
    int jq = q / 3, cq = q % 3;
    for (int l = 0; l < nd; ++l) {
      if (l % 3 != cq) continue;
      double p = RI[(size_t)(l / 3) * 8 + jq];
      if (p == 0.0) continue;
      const double* Kc = K + (size_t)l * nd;
      double* o = &KP[(size_t)q * nd];
      for (int k = 0; k < nd; ++k) o[k] += Kc[k] * p;
    }
  }
  (void)span;
  for (int q = 0; q < 24; ++q)
    for (int p = 0; p < 24; ++p) {
      int jp = p / 3, cp = p % 3;
      double s = 0.0;
      for (int k = cp; k < nd; k += 3) s += RI[(size_t)(k / 3) * 8 + jp] * KP[(size_t)q * nd + k];
      out[q * 24 + p] = s;
    }
}

// Solving_AssembleFEAstencil, 919-1066.
bool assemble(Model& M) {
  const int NL = M.numLevels;
  std::vector<double> prevKs;
  for (int ii = 1; ii < NL; ++ii) {
    Level& C = M.L[ii];
    const int span = C.span, m = span + 1, m3 = m * m * m, s3 = span * span * span;
    const int nd = 3 * m3;
    std::vector<double> Ks((size_t)C.numElements * 576, 0.0);
    if (ii == 1) {
      // templates (950-954)
      std::vector<double> T((size_t)s3 * 576);
      for (int ss = 0; ss < s3; ++ss) {
        std::vector<double> Kloc((size_t)nd * nd, 0.0);
        int dof[24];
        sub_elem_dofs(span, ss, dof);
        for (int q = 0; q < 24; ++q)
          for (int p = 0; p < 24; ++p) Kloc[(size_t)dof[q] * nd + dof[p]] += M.Ke[q * 24 + p];
        project(span, C.RI, Kloc.data(), nd, &T[(size_t)ss * 576]);
      }
      const std::vector<double>& E = M.eleModulus;
#pragma omp parallel for schedule(dynamic, 16)
      for (int jj = 0; jj < C.numElements; ++jj) {
        const int* sons = &C.elementUpwardMap[(size_t)jj * s3];
        double* out = &Ks[(size_t)jj * 576];
        int idx = M.l2BC[jj];
        if (idx > 0) {  // 961-971
          std::vector<double> Kloc((size_t)nd * nd, 0.0);
          const std::vector<int>& loc = M.sonElesWithFixed[idx - 1];
          for (int ss = 0; ss < s3; ++ss) {
            if (!sons[ss]) continue;
            int e = sons[ss] - 1;
            bool bc = std::find(loc.begin(), loc.end(), ss) != loc.end();
            const double* Fr = bc ? &M.KesFree[(size_t)(M.mapUniqueKes[e] - 1) * 576] : M.Ke.data();
            const double* Fx = bc ? &M.KesFixed[(size_t)(M.mapUniqueKes[e] - 1) * 576] : nullptr;
            int dof[24];
            sub_elem_dofs(span, ss, dof);
            for (int q = 0; q < 24; ++q)
              for (int p = 0; p < 24; ++p) {
                double v = Fr[q * 24 + p] * E[e];
                if (bc) v = v + Fx[q * 24 + p];
                Kloc[(size_t)dof[q] * nd + dof[p]] += v;
              }
          }
          project(span, C.RI, Kloc.data(), nd, out);
        } else {  // 973-976: sum over ss of E_ss * T_ss
          for (int k = 0; k < 576; ++k) {
            double s = 0.0;
            for (int ss = 0; ss < s3; ++ss) {
              double Ev = sons[ss] ? E[sons[ss] - 1] : 0.0;
              s += T[(size_t)ss * 576 + k] * Ev;
            }
            out[k] = s;
          }
        }
      }
    } else {  // 980-995
#pragma omp parallel for schedule(dynamic, 16)
      for (int jj = 0; jj < C.numElements; ++jj) {
        const int* sons = &C.elementUpwardMap[(size_t)jj * s3];
        std::vector<double> Kloc((size_t)nd * nd, 0.0);
        for (int ss = 0; ss < s3; ++ss) {
          if (!sons[ss]) continue;
          const double* Kc = &prevKs[(size_t)(sons[ss] - 1) * 576];
          int dof[24];
          sub_elem_dofs(span, ss, dof);
          for (int q = 0; q < 24; ++q)
            for (int p = 0; p < 24; ++p) Kloc[(size_t)dof[q] * nd + dof[p]] += Kc[q * 24 + p];
        }
        project(span, C.RI, Kloc.data(), nd, &Ks[(size_t)jj * 576]);
      }
    }
    // diagonals on coarser levels (1001-1018)
    if (ii < NL - 1) {
      C.diagK.assign((size_t)C.numNodes * 3, 0.0);
#pragma omp parallel for schedule(static)
      for (int n = 0; n < C.numNodes; ++n)
        for (int t = C.n2eStart[n]; t < C.n2eStart[n + 1]; ++t) {
          int e = C.n2eList[t] / 8, j = C.n2eList[t] % 8;
          for (int c = 0; c < 3; ++c) C.diagK[(size_t)n * 3 + c] += Ks[(size_t)e * 576 + (3 * j + c) * 25];
        }
    }
    C.Ks.swap(Ks);
    prevKs = C.Ks;
  }
  // finest diagonal (1021-1050)
  {
    Level& F = M.L[0];
    F.diagK.assign((size_t)F.numNodes * 3, 0.0);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < F.numNodes; ++n)
      for (int t = F.n2eStart[n]; t < F.n2eStart[n + 1]; ++t) {
        int e = F.n2eList[t] / 8, j = F.n2eList[t] % 8;
        for (int c = 0; c < 3; ++c) {
          int d = 3 * j + c;
          double v;
          if (M.mapUniqueKes[e]) {
            size_t u = M.mapUniqueKes[e] - 1;
            v = M.KesFree[u * 576 + d * 25] * M.eleModulus[e] + M.KesFixed[u * 576 + d * 25];
          } else {
            v = M.Ke[d * 25] * M.eleModulus[e];
          }
          F.diagK[(size_t)n * 3 + c] += v;
        }
      }
  }
  // coarsest assembly + Cholesky (1053-1065), banded with our own ordering
  {
    Level& C = M.L[NL - 1];
    const int nn = C.numNodes;
    // choose node ordering: dense coordinates with the axis permutation that
    // minimises the bandwidth
    std::vector<int> cy(nn), cx(nn), cz(nn);
    const int NY = C.resY + 1, NX = C.resX + 1;
    for (int n = 0; n < nn; ++n) {
      int d = C.nodMapBack[n];
      cy[n] = d % NY; cx[n] = (d / NY) % NX; cz[n] = d / NY / NX;
    }
    const int dims[3] = {C.resY + 1, C.resX + 1, C.resZ + 1};
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<std::pair<int, std::vector<int>>> cands;
    for (auto& pm : perms) {  // pm[0] fastest
      std::vector<i64> key(nn);
      for (int n = 0; n < nn; ++n) {
        int co[3] = {cy[n], cx[n], cz[n]};
        key[n] = co[pm[0]] + (i64)dims[pm[0]] * (co[pm[1]] + (i64)dims[pm[1]] * co[pm[2]]);
      }
      std::vector<int> ord(nn);
      for (int n = 0; n < nn; ++n) ord[n] = n;
      std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
      std::vector<int> pos(nn);
      for (int k = 0; k < nn; ++k) pos[ord[k]] = k;
      int bw = 0;
      for (int e = 0; e < C.numElements; ++e) {
        int lo = nn, hi = -1;
        for (int j = 0; j < 8; ++j) {
          int p = pos[C.eNodMat[(size_t)e * 8 + j]];
          lo = std::min(lo, p); hi = std::max(hi, p);
        }
        bw = std::max(bw, hi - lo);
      }
      cands.push_back({bw, ord});
    }
    std::stable_sort(cands.begin(), cands.end(),
                     [](const std::pair<int, std::vector<int>>& a, const std::pair<int, std::vector<int>>& b) {
                       return a.first < b.first;
                     });
    const size_t pick = std::min<size_t>((size_t)M.orderVariant, cands.size() - 1);
    const int bestBW = cands[pick].first;
    const std::vector<int>& bestOrd = cands[pick].second;
    std::vector<int> pos(nn);
    for (int k = 0; k < nn; ++k) pos[bestOrd[k]] = k;
    M.cN = 3 * nn;
    M.cBW = 3 * bestBW + 2;
    const int W = M.cBW + 1;
    M.cPerm.resize(M.cN);
    for (int k = 0; k < nn; ++k)
      for (int c = 0; c < 3; ++c) M.cPerm[3 * k + c] = 3 * bestOrd[k] + c;
    M.cBand.assign((size_t)M.cN * W, 0.0);
    // sparse(iK,jK,sK): sum element contributions, keep lower band
    for (int e = 0; e < C.numElements; ++e) {
      const double* Ke = &C.Ks[(size_t)e * 576];
      int pd[24];
      for (int j = 0; j < 8; ++j)
        for (int c = 0; c < 3; ++c) pd[3 * j + c] = 3 * pos[C.eNodMat[(size_t)e * 8 + j]] + c;
      for (int q = 0; q < 24; ++q)
        for (int p = 0; p < 24; ++p) {
          int r = pd[p], cc = pd[q];
          if (r < cc) continue;
          M.cBand[(size_t)r * W + (r - cc)] += Ke[q * 24 + p];
        }
    }
    // banded Cholesky: A = L L^T, L stored row-wise in band form
    std::vector<double>& B = M.cBand;
    const int N = M.cN, bw = M.cBW;
    for (int j = 0; j < N; ++j) {
      double d = B[(size_t)j * W + 0];
      for (int k = std::max(0, j - bw); k < j; ++k) {
        double l = B[(size_t)j * W + (j - k)];
        d -= l * l;
      }
      if (!(d > 0.0)) { M.err = "coarsest matrix not positive definite"; M.cholOk = false; return false; }
      d = std::sqrt(d);
      B[(size_t)j * W] = d;
      const int iEnd = std::min(N - 1, j + bw);
#pragma omp parallel for schedule(static) if (iEnd - j > 256)
      for (int i = j + 1; i <= iEnd; ++i) {
        double s = B[(size_t)i * W + (i - j)];
        const int k0 = std::max(0, i - bw);
        for (int k = k0; k < j; ++k) s -= B[(size_t)i * W + (i - k)] * B[(size_t)j * W + (j - k)];
        B[(size_t)i * W + (i - j)] = s / d;
      }
    }
    M.cholOk = true;
  }
  return true;
}

// x = S (L^T \ (L \ (S^T r)))   (767)
void coarse_solve(const Model& M, const double* r, double* x) {
  const int N = M.cN, bw = M.cBW, W = bw + 1;
  const std::vector<double>& B = M.cBand;
  std::vector<double> y(N);
  for (int i = 0; i < N; ++i) {
    double s = r[M.cPerm[i]];
    for (int k = std::max(0, i - bw); k < i; ++k) s -= B[(size_t)i * W + (i - k)] * y[k];
    y[i] = s / B[(size_t)i * W];
  }
  for (int i = N - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k <= std::min(N - 1, i + bw); ++k) s -= B[(size_t)k * W + (k - i)] * y[k];
    y[i] = s / B[(size_t)i * W];
  }
  for (int i = 0; i < N; ++i) x[M.cPerm[i]] = y[i];
}

// ------------------------------------------------------------ K*u (level 1)
// Solving_KbyU_MatrixFree, 641-726 (the live branch 684-725).
void apply_K(const Model& M, const double* u, double* Y) {
  const Level& L = M.L[0];
  const int nE = L.numElements;
  // per element products (697-715)
  std::vector<double>& prod = M.scratchK;
  prod.resize((size_t)nE * 24);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nE; ++e) {
    double ue[24];
    for (int j = 0; j < 8; ++j) {
      int n = L.eNodMat[(size_t)e * 8 + j];
      for (int c = 0; c < 3; ++c) ue[3 * j + c] = u[(size_t)n * 3 + c];
    }
    double* o = &prod[(size_t)e * 24];
    const double Ee = M.eleModulus[e];
    double s[24];
    for (int k = 0; k < 24; ++k) s[k] = 0.0;
    if (M.mapUniqueKes[e]) {  // 708-713: u_e * (KeFree*E + KeFixed)
      size_t uq = M.mapUniqueKes[e] - 1;
      const double* Fr = &M.KesFree[uq * 576];
      const double* Fx = &M.KesFixed[uq * 576];
      for (int b = 0; b < 24; ++b)
        for (int k = 0; k < 24; ++k) s[k] += ue[b] * (Fr[k * 24 + b] * Ee + Fx[k * 24 + b]);
      for (int k = 0; k < 24; ++k) o[k] = s[k];
    } else {  // 714: (u_e * Ke) .* E   (Ke is exactly symmetric: row b == column b)
      const double* K = M.Ke.data();
      for (int b = 0; b < 24; ++b)
        for (int k = 0; k < 24; ++k) s[k] += ue[b] * K[b * 24 + k];
      for (int k = 0; k < 24; ++k) o[k] = s[k] * Ee;
    }
  }
  // accumarray scatter (716-721) as an ordered gather
#pragma omp parallel for schedule(static)
  for (int n = 0; n < L.numNodes; ++n) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int t = L.n2eStart[n]; t < L.n2eStart[n + 1]; ++t) {
      int e = L.n2eList[t] / 8, j = L.n2eList[t] % 8;
      const double* o = &prod[(size_t)e * 24 + 3 * j];
      s0 += o[0]; s1 += o[1]; s2 += o[2];
    }
    Y[(size_t)n * 3 + 0] = s0; Y[(size_t)n * 3 + 1] = s1; Y[(size_t)n * 3 + 2] = s2;
  }
}

// ---------------------------------------------------------------- V-cycle
// Solving_RestrictResidual, 783-798.
void restrict_residual(const Level& C, const double* rF, double* rC) {
  const int m = C.span + 1, m3 = m * m * m;
  std::vector<double> r1((size_t)C.intermediateNumNodes * 3, 0.0);
  const int nF = (int)C.solidNodeMapCoarser2Finer.size();
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nF; ++k)
    for (int c = 0; c < 3; ++c) r1[(size_t)C.solidNodeMapCoarser2Finer[k] * 3 + c] = rF[(size_t)k * 3 + c];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < C.intermediateNumNodes; ++i)
    for (int c = 0; c < 3; ++c) r1[(size_t)i * 3 + c] = r1[(size_t)i * 3 + c] / C.transferMatCoeffi[i];
  std::vector<double> contrib((size_t)C.numElements * 24);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < C.numElements; ++e)
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < 8; ++j) {
        double s = 0.0;
        for (int a = 0; a < m3; ++a)
          s += r1[(size_t)C.transferMat[(size_t)e * m3 + a] * 3 + c] * C.RI[(size_t)a * 8 + j];
        contrib[(size_t)e * 24 + 3 * j + c] = s;
      }
#pragma omp parallel for schedule(static)
  for (int n = 0; n < C.numNodes; ++n) {
    double s[3] = {0, 0, 0};
    for (int t = C.n2eStart[n]; t < C.n2eStart[n + 1]; ++t) {
      int e = C.n2eList[t] / 8, j = C.n2eList[t] % 8;
      for (int c = 0; c < 3; ++c) s[c] += contrib[(size_t)e * 24 + 3 * j + c];
    }
    for (int c = 0; c < 3; ++c) rC[(size_t)n * 3 + c] = s[c];
  }
}

// Solving_InterpolationDeviation, 800-814.
void interpolate(const Level& C, const double* xC, double* xF) {
  const int m = C.span + 1, m3 = m * m * m;
  std::vector<double> vals((size_t)C.numElements * m3 * 3);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < C.numElements; ++e)
    for (int a = 0; a < m3; ++a)
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int j = 0; j < 8; ++j) s += C.RI[(size_t)a * 8 + j] * xC[(size_t)C.eNodMat[(size_t)e * 8 + j] * 3 + c];
        vals[((size_t)e * m3 + a) * 3 + c] = s;
      }
  const int nF = (int)C.solidNodeMapCoarser2Finer.size();
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nF; ++k) {
    int i = C.solidNodeMapCoarser2Finer[k];
    double s[3] = {0, 0, 0};
    for (int t = C.i2tStart[i]; t < C.i2tStart[i + 1]; ++t)
      for (int c = 0; c < 3; ++c) s[c] += vals[(size_t)C.i2tList[t] * 3 + c];
    for (int c = 0; c < 3; ++c) xF[(size_t)k * 3 + c] = s[c] / C.transferMatCoeffi[i];
  }
}

// Solving_Vcycle ('Adapted'), 747-781.
void vcycle(const Model& M, const double* r, double* x) {
  const int NL = M.numLevels;
  std::vector<std::vector<double>> R(NL), X(NL);
  R[0].assign(r, r + (size_t)M.L[0].numNodes * 3);
  for (int ii = 0; ii < NL - 1; ++ii) {  // 757-766
    const Level& L = M.L[ii];
    const size_t n = (size_t)L.numNodes * 3;
    X[ii].resize(n);
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; ++k) X[ii][k] = M.S.omega * (R[ii][k] / L.diagK[k]);  // 758
    R[ii + 1].resize((size_t)M.L[ii + 1].numNodes * 3);
    restrict_residual(M.L[ii + 1], R[ii].data(), R[ii + 1].data());  // 761
  }
  X[NL - 1].resize(R[NL - 1].size());
  coarse_solve(M, R[NL - 1].data(), X[NL - 1].data());  // 767
  for (int ii = NL - 2; ii >= 0; --ii) {  // 769-778
    const Level& L = M.L[ii];
    const size_t n = (size_t)L.numNodes * 3;
    std::vector<double> px(n);
    interpolate(M.L[ii + 1], X[ii + 1].data(), px.data());  // 770
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; ++k) {
      X[ii][k] = X[ii][k] + px[k];
      X[ii][k] = X[ii][k] + M.S.omega * R[ii][k] / L.diagK[k];  // 773
    }
  }
  std::memcpy(x, X[0].data(), X[0].size() * sizeof(double));
}

// Solving_KbyU_MatrixFree for iLevel > 1 (727-744): per-element products with
// the stored Galerkin matrices Ks, accumulated per node (accumarray 736-741).
void apply_K_level(const Model& M, int level, const double* u, double* Y) {
  const Level& L = M.L[level];
  std::vector<double> prod((size_t)L.numElements * 24);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < L.numElements; ++e) {
    double ue[24];
    for (int j = 0; j < 8; ++j)
      for (int c = 0; c < 3; ++c) ue[3 * j + c] = u[(size_t)L.eNodMat[(size_t)e * 8 + j] * 3 + c];
    const double* K = &L.Ks[(size_t)e * 576];  // column-major 24x24
    for (int k = 0; k < 24; ++k) {  // 734: (u_e * Ks)(k) = sum_b u_e(b) Ks(b, k)
      double s = 0.0;
      for (int b = 0; b < 24; ++b) s += ue[b] * K[k * 24 + b];
      prod[(size_t)e * 24 + k] = s;
    }
  }
#pragma omp parallel for schedule(static)
  for (int n = 0; n < L.numNodes; ++n) {
    double s[3] = {0, 0, 0};
    for (int t = L.n2eStart[n]; t < L.n2eStart[n + 1]; ++t) {
      int e = L.n2eList[t] / 8, j = L.n2eList[t] % 8;
      for (int c = 0; c < 3; ++c) s[c] += prod[(size_t)e * 24 + 3 * j + c];
    }
    for (int c = 0; c < 3; ++c) Y[(size_t)n * 3 + c] = s[c];
  }
}

// K*u on any level: the fine operator with E and the BC rows on level 0,
// the stored Galerkin matrices above it.
void apply_K_any(const Model& M, int level, const double* u, double* Y) {
  if (level == 0) apply_K(M, u, Y);
  else apply_K_level(M, level, u, Y);
}

// Solving_Vcycle ('Standard'), 747-781: damped-Jacobi pre-smoothing from
// zero, defect restriction (759-762), coarsest solve (767), coarse-grid
// correction (770) and one damped-Jacobi post-smoothing step (775-776).
void vcycle_standard(const Model& M, const double* r, double* x) {
  const int NL = M.numLevels;
  const double w = M.S.omega;
  std::vector<std::vector<double>> R(NL), X(NL);
  R[0].assign(r, r + (size_t)M.L[0].numNodes * 3);
  for (int ii = 0; ii < NL - 1; ++ii) {
    const Level& L = M.L[ii];
    const size_t n = (size_t)L.numNodes * 3;
    X[ii].resize(n);
    for (size_t k = 0; k < n; ++k) X[ii][k] = w * (R[ii][k] / L.diagK[k]);  // 758
    std::vector<double> d(n);
    apply_K_any(M, ii, X[ii].data(), d.data());
    for (size_t k = 0; k < n; ++k) d[k] = R[ii][k] - d[k];  // 764
    R[ii + 1].resize((size_t)M.L[ii + 1].numNodes * 3);
    restrict_residual(M.L[ii + 1], d.data(), R[ii + 1].data());  // 765
  }
  X[NL - 1].resize(R[NL - 1].size());
  coarse_solve(M, R[NL - 1].data(), X[NL - 1].data());  // 767
  for (int ii = NL - 2; ii >= 0; --ii) {
    const Level& L = M.L[ii];
    const size_t n = (size_t)L.numNodes * 3;
    std::vector<double> px(n), kx(n);
    interpolate(M.L[ii + 1], X[ii + 1].data(), px.data());  // 770
    for (size_t k = 0; k < n; ++k) X[ii][k] += px[k];
    apply_K_any(M, ii, X[ii].data(), kx.data());
    for (size_t k = 0; k < n; ++k) X[ii][k] += w * (R[ii][k] - kx[k]) / L.diagK[k];  // 775-776
  }
  std::memcpy(x, X[0].data(), X[0].size() * sizeof(double));
}

void precondition(const Model& M, const double* r, double* x) {
  if (M.S.typeVcycle == 1) vcycle_standard(M, r, x);
  else vcycle(M, r, x);
}

// ------------------------------------------------------------------- PCG
// Solving_PCG, 600-639.  Returns its; *res = final relative residual.
template <class AOp, class POp>
int pcg(AOp A, POp P, const std::vector<double>& b, std::vector<double>& y, double tol, int maxIT,
        double* resOut, bool* warned) {
  const size_t n = b.size();
  const double normB = normv(b);  // 607
  std::vector<double> r(n), z(n), p(n), q(n);
  A(y.data(), q.data());
  for (size_t k = 0; k < n; ++k) r[k] = b[k] - q[k];  // 609
  P(r.data(), z.data());                            // 610
  p = z;                                            // 611
  double x1 = dotv(z, r);                           // 612
  int its = 0;
  double resnorm = 0.0;
  for (its = 1; its <= maxIT; ++its) {  // 613
    A(p.data(), q.data());              // 614
    const double lambda = x1 / dotv(p, q);  // 615
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; ++k) {
      y[k] = y[k] + lambda * p[k];  // 616
      r[k] = r[k] - lambda * q[k];  // 617
    }
    resnorm = normv(r) / normB;  // 618
    if (resnorm < tol) break;     // 622
    P(r.data(), z.data());        // 628
    const double x2 = dotv(z, r); // 629
    const double beta = x2 / x1;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];  // 630
    x1 = x2;
  }
  if (its > maxIT) its = maxIT;
  if (warned) *warned = (its == maxIT);  // 634-637
  if (resOut) *resOut = resnorm;
  return its;
}

// ---------------------------------------------------------------- compliance
// TopOpti_ComputeUnitCompliance, 579-598, one load case (the columns of
// ceList, 594).
void compliance(const Model& M, const double* U, double* ce) {
  const Level& L = M.L[0];
#pragma omp parallel for schedule(static)
  for (int e = 0; e < L.numElements; ++e) {
    double ue[24];
    for (int j = 0; j < 8; ++j)
      for (int c = 0; c < 3; ++c) ue[3 * j + c] = U[(size_t)L.eNodMat[(size_t)e * 8 + j] * 3 + c];
    double s = 0.0;
    for (int k = 0; k < 24; ++k) {
      double t = 0.0;
      for (int b = 0; b < 24; ++b) t += ue[b] * M.Ke[k * 24 + b];
      s += t * ue[k];
    }
    ce[e] = s;
  }
}

// ceList * objWeightingList_(:) (597): the weighted sum over the load cases.
void compliance_cases(const Model& M, const std::vector<std::vector<double>>& U, double* ce) {
  const int nE = M.L[0].numElements;
  std::vector<double> cj(nE);
  for (int e = 0; e < nE; ++e) ce[e] = 0.0;
  for (size_t j = 0; j < U.size(); ++j) {
    compliance(M, U[j].data(), cj.data());
    for (int e = 0; e < nE; ++e) ce[e] += cj[e] * M.objWeight[j];
  }
}

// ---------------------------------------------------------------- PDE filter
// TopOpti_SetupPDEfilter_matrixFree, 816-868.
void setup_filter(Model& M, double rMin, Filter& Fl) {
  const double g = 1.0 / std::sqrt(3.0);
  const double s[8] = {-g, g, g, -g, -g, g, g, -g};   // 820
  const double t[8] = {-g, -g, g, g, -g, -g, g, g};   // 821
  const double p[8] = {-g, -g, -g, -g, g, g, g, g};   // 822
  double N[8][8];
  for (int q = 0; q < 8; ++q) trilinear(s[q], t[q], p[q], N[q]);  // 829
  double dS[24][8];                                   // 832-841
  for (int q = 0; q < 8; ++q) {
    const double S_ = s[q], T = t[q], P = p[q];
    const double ds[8] = {-0.125 * (1 - T) * (1 - P), 0.125 * (1 - T) * (1 - P), 0.125 * (1 + T) * (1 - P),
                          -0.125 * (1 + T) * (1 - P), -0.125 * (1 - T) * (1 + P), 0.125 * (1 - T) * (1 + P),
                          0.125 * (1 + T) * (1 + P), -0.125 * (1 + T) * (1 + P)};
    const double dt[8] = {-0.125 * (1 - S_) * (1 - P), -0.125 * (1 + S_) * (1 - P), 0.125 * (1 + S_) * (1 - P),
                          0.125 * (1 - S_) * (1 - P), -0.125 * (1 - S_) * (1 + P), -0.125 * (1 + S_) * (1 + P),
                          0.125 * (1 + S_) * (1 + P), 0.125 * (1 - S_) * (1 + P)};
    const double dp[8] = {-0.125 * (1 - S_) * (1 - T), -0.125 * (1 + S_) * (1 - T), -0.125 * (1 + S_) * (1 + T),
                          -0.125 * (1 - S_) * (1 + T), 0.125 * (1 - S_) * (1 - T), 0.125 * (1 + S_) * (1 - T),
                          0.125 * (1 + S_) * (1 + T), 0.125 * (1 - S_) * (1 + T)};
    for (int a = 0; a < 8; ++a) { dS[3 * q][a] = ds[a]; dS[3 * q + 1][a] = dt[a]; dS[3 * q + 2][a] = dp[a]; }
  }
  const double detJ = 1.0 / 8.0;  // 845-847
  const double iRmin = (rMin * M.L[0].eleSize) / 2.0 / std::sqrt(3.0);  // 850
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      double k0 = 0.0, k1 = 0.0;
      for (int r = 0; r < 24; ++r) k0 += dS[r][a] * dS[r][b];           // 848
      for (int q = 0; q < 8; ++q) k1 += N[q][a] * (1.0 * detJ) * N[q][b];  // 849
      Fl.kernel[b * 8 + a] = iRmin * iRmin * k0 + k1;                     // 851
    }
  const Level& L = M.L[0];
  Fl.diagPrecond.assign(L.numNodes, 0.0);  // 855-867
  for (int n = 0; n < L.numNodes; ++n) {
    double sd = 0.0;
    for (int tt = L.n2eStart[n]; tt < L.n2eStart[n + 1]; ++tt) {
      int j = L.n2eList[tt] % 8;
      sd += Fl.kernel[j * 9];
    }
    Fl.diagPrecond[n] = 1.0 / sd;
  }
}

// MatTimesVec_matrixFree_B, 904-917.
void filter_apply(const Model& M, const Filter& Fl, const double* u, double* Y) {
  const Level& L = M.L[0];
  std::vector<double> prod((size_t)L.numElements * 8);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < L.numElements; ++e) {
    double ue[8];
    for (int j = 0; j < 8; ++j) ue[j] = u[L.eNodMat[(size_t)e * 8 + j]];
    for (int k = 0; k < 8; ++k) {
      double s = 0.0;
      for (int b = 0; b < 8; ++b) s += ue[b] * Fl.kernel[k * 8 + b];
      prod[(size_t)e * 8 + k] = s;
    }
  }
#pragma omp parallel for schedule(static)
  for (int n = 0; n < L.numNodes; ++n) {
    double s = 0.0;
    for (int t = L.n2eStart[n]; t < L.n2eStart[n + 1]; ++t) s += prod[L.n2eList[t]];
    Y[n] = s;
  }
}

// TopOpti_ConductPDEFiltering_matrixFree, 870-902.
int filter_run(const Model& M, const Filter& Fl, const double* src, double* tar) {
  const Level& L = M.L[0];
  std::vector<double> b(L.numNodes, 0.0);  // 880-886
  for (int n = 0; n < L.numNodes; ++n) {
    double s = 0.0;
    for (int t = L.n2eStart[n]; t < L.n2eStart[n + 1]; ++t) s += src[L.n2eList[t] / 8] * (1.0 / 8.0);
    b[n] = s;
  }
  std::vector<double> x(L.numNodes, 0.0);
  auto A = [&](const double* u, double* y) { filter_apply(M, Fl, u, y); };
  auto P = [&](const double* r, double* z) {
    for (int n = 0; n < L.numNodes; ++n) z[n] = Fl.diagPrecond[n] * r[n];  // 889
  };
  int its = pcg(A, P, b, x, M.S.tol / 10.0, M.S.maxIT, nullptr, nullptr);  // 891
#pragma omp parallel for schedule(static)
  for (int e = 0; e < L.numElements; ++e) {  // 894-901
    double s = 0.0;
    for (int j = 0; j < 8; ++j) s += x[L.eNodMat[(size_t)e * 8 + j]];
    tar[e] = s * (1.0 / 8.0);
  }
  return its;
}

// ------------------------------------------------------------------ solver
int solve_fea(Model& M, int lc, std::vector<double>& U, bool* warned, double* res) {
  auto A = [&](const double* u, double* y) { apply_K(M, u, y); };
  auto P = [&](const double* r, double* z) { precondition(M, r, z); };
  return pcg(A, P, M.F[lc], U, M.S.tol, M.S.maxIT, res, warned);
}

// --------------------------------------------------------------------- OC
// 145-155.  Returns change; x updated in place.
double oc_update(const Model& M, std::vector<double>& x, const std::vector<double>& xPhys,
                 const std::vector<double>& dc, double V0, int* nBisect) {
  const i64 n = (i64)x.size();
  const double fval = det_sum(n, [&](i64 i) { return xPhys[i]; }) / (double)n - V0;  // 145
  double l1 = 0.0, l2 = 1e9;
  const double move = 0.2;  // 146
  std::vector<double> xnew(n);
  int nb = 0;
  while ((l2 - l1) / (l1 + l2) > 1e-6) {  // 147
    const double lmid = 0.5 * (l2 + l1);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {  // 149
      double v = x[i] * std::sqrt(-dc[i] / 1.0 / lmid);
      v = std::min(x[i] + move, v);
      v = std::min(1.0, v);
      v = std::max(x[i] - move, v);
      xnew[i] = std::max(0.0, v);
      if (!M.isPassive.empty() && M.isPassive[i]) xnew[i] = 1.0;  // 150
    }
    const double gt = fval + det_sum(n, [&](i64 i) { return 1.0 * (xnew[i] - x[i]); });  // 151
    if (gt > 0) l1 = lmid; else l2 = lmid;  // 152
    nb++;
  }
  double change = 0.0;  // 154
  for (i64 i = 0; i < n; ++i) change = std::max(change, std::fabs(xnew[i] - x[i]));
  x.swap(xnew);
  if (nBisect) *nBisect = nb;
  return change;
}

// -------------------------------------------------------------------- MMA
// MMAseq / Update / GenSub / SolveDIP and helpers, 1804-2049, for m = 1.
struct MMA {
  int n = 0;
  std::vector<double> x, xold1, xold2, low, upp, alpha, beta, p0, q0, pij, qij;
  double a = 0.0, c = 1000.0, d = 0.0;  // 1814-1816
  double y = 0.0, z = 0.0, lam = 0.0, mu = 0.0, b = 0.0, grad = 0.0, hess = 0.0, s1 = 0.0, s2 = 0.0;
};

void mma_xyz(MMA& S) {  // XYZofLAMBDA 1964-1979
  if (S.lam < 0.0) S.lam = 0.0;
  S.y = S.lam - S.c; if (S.y < 0.0) S.y = 0.0;
  const double lamai = S.lam * S.a;
  S.z = std::max(0.0, 10.0 * (lamai - 1.0));
  const double lam = S.lam;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < S.n; ++i) {
    const double pj = S.p0[i] + S.pij[i] * lam, qj = S.q0[i] + S.qij[i] * lam;
    double xv = (std::sqrt(pj) * S.low[i] + std::sqrt(qj) * S.upp[i]) / (std::sqrt(pj) + std::sqrt(qj));
    if (xv - S.alpha[i] < 0) xv = S.alpha[i];
    if (xv - S.beta[i] > 0) xv = S.beta[i];
    S.x[i] = xv;
  }
}
void mma_grad(MMA& S) {  // DualGrad 1981-1986
  S.grad = -S.b - S.a * S.z - S.y;
  S.grad = S.grad + det_sum(S.n, [&](i64 i) { return S.pij[i] / (S.upp[i] - S.x[i]) + S.qij[i] / (S.x[i] - S.low[i]); });
}
void mma_hess(MMA& S) {  // DualHess 1988-2022
  const double lam = S.lam;
  double H = det_sum(S.n, [&](i64 i) {
    const double pj = S.p0[i] + S.pij[i] * lam, qj = S.q0[i] + S.qij[i] * lam;
    const double ux = S.upp[i] - S.x[i], xl = S.x[i] - S.low[i];
    const double PQ = S.pij[i] / (ux * ux) - S.qij[i] / (xl * xl);
    double df2 = -1.0 / (2.0 * pj / (ux * ux * ux) + 2.0 * qj / (xl * xl * xl));
    const double xp = (std::sqrt(pj) * S.low[i] + std::sqrt(qj) * S.upp[i]) / (std::sqrt(pj) + std::sqrt(qj));
    if (xp - S.alpha[i] < 0) df2 = 0.0;
    if (xp - S.beta[i] > 0) df2 = 0.0;
    return PQ * df2 * PQ;
  });
  if (S.lam < 0.0) S.lam = 0.0;
  const double lamai = S.lam * S.a;
  if (S.lam > S.c) H = H - 1.0;
  H = H - S.mu / S.lam;
  if (lamai > 1.0) H = H - 10.0 * S.a * S.a;
  double corr = 1e-4 * H / 1.0;  // HessTrace/m
  if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
  S.hess = H + corr;
}
void mma_linesearch(MMA& S) {  // DualLineSearch 2024-2034
  double theta = 1.005;
  if (theta < -1.01 * S.s1 / S.lam) theta = -1.01 * S.s1 / S.lam;
  if (theta < -1.01 * S.s2 / S.mu) theta = -1.01 * S.s2 / S.mu;
  theta = 1.0 / theta;
  S.lam = S.lam + theta * S.s1;
  S.mu = S.mu + theta * S.s2;
}
double mma_residual(MMA& S, double epsi) {  // DualResidual 2036-2049
  double r1 = -S.b - S.a * S.z - S.y + S.mu;
  const double sp = det_sum(S.n, [&](i64 i) { return S.pij[i] / (S.upp[i] - S.x[i]); });
  const double sq = det_sum(S.n, [&](i64 i) { return S.qij[i] / (S.x[i] - S.low[i]); });
  r1 = r1 + sp + sq;
  const double r2 = S.mu * S.lam - epsi;
  return std::max(std::fabs(r1), std::fabs(r2));
}

// One MMAseq call (1804-1860).  dfdx, dgdx: n; gx scalar.
void mma_update(MMA& S, int iter, const double* xval, const double* xmin, const double* xmax,
                const double* dfdx, double gx, const double* dgdx, int* nNewton) {
  const int n = S.n;
  S.x.assign(xval, xval + n);
  // GenSub (1873-1928)
  const double asyinit = 0.5, asydec = 0.7, asyinc = 1.2, feps = 1.0e-6;
  if (iter < 3 || (int)S.low.size() < n || (int)S.upp.size() < n) {
    S.low.resize(n); S.upp.resize(n);
    for (int i = 0; i < n; ++i) {
      S.low[i] = S.x[i] - asyinit * (xmax[i] - xmin[i]);
      S.upp[i] = S.x[i] + asyinit * (xmax[i] - xmin[i]);
    }
  } else {
    for (int i = 0; i < n; ++i) {
      const double zzz = (S.x[i] - S.xold1[i]) * (S.xold1[i] - S.xold2[i]);
      double gamma = 1.0;
      if (zzz > 0) gamma = asyinc;
      if (zzz < 0) gamma = asydec;
      double lo = S.x[i] - gamma * (S.xold1[i] - S.low[i]);
      double up = S.x[i] + gamma * (S.upp[i] - S.xold1[i]);
      const double lowmin = S.x[i] - 10.0 * (xmax[i] - xmin[i]);
      const double lowmax = S.x[i] - 0.01 * (xmax[i] - xmin[i]);
      const double uppmin = S.x[i] + 0.01 * (xmax[i] - xmin[i]);
      const double uppmax = S.x[i] + 10.0 * (xmax[i] - xmin[i]);
      lo = std::max(lo, lowmin); lo = std::min(lo, lowmax);
      up = std::min(up, uppmax); up = std::max(up, uppmin);
      S.low[i] = lo; S.upp[i] = up;
    }
  }
  S.alpha.resize(n); S.beta.resize(n); S.p0.resize(n); S.q0.resize(n); S.pij.resize(n); S.qij.resize(n);
  for (int i = 0; i < n; ++i) {
    double al = 0.9 * S.low[i] + 0.1 * S.x[i];
    if (al - xmin[i] < 0) al = xmin[i];
    double be = 0.9 * S.upp[i] + 0.1 * S.x[i];
    if (be - xmax[i] > 0) be = xmax[i];
    S.alpha[i] = al; S.beta[i] = be;
    const double ux = S.upp[i] - S.x[i], xl = S.x[i] - S.low[i], ul = S.upp[i] - S.low[i];
    const double fp = std::max(0.0, dfdx[i]), fm = std::max(0.0, -1.0 * dfdx[i]);
    S.p0[i] = ux * ux * (fp + 0.001 * std::fabs(dfdx[i]) + 0.5 * feps / ul);
    S.q0[i] = xl * xl * (fm + 0.001 * std::fabs(dfdx[i]) + 0.5 * feps / ul);
    const double gp = std::max(0.0, dgdx[i]), gm = std::max(0.0, -1.0 * dgdx[i]);
    const double extra = 0.001 * std::fabs(dgdx[i]) + 0.5 * feps / ul;
    S.pij[i] = ux * ux * (gp + extra);
    S.qij[i] = xl * xl * (gm + extra);
  }
  S.b = -gx + det_sum(n, [&](i64 i) { return S.pij[i] / (S.upp[i] - S.x[i]) + S.qij[i] / (S.x[i] - S.low[i]); });
  // Update: xold2 = xold1; xold1 = xval (1867-1868)
  S.xold2 = S.xold1;
  S.xold1 = S.x;
  // SolveDIP (1930-1962)
  S.lam = S.c / 2.0;
  S.mu = 1.0;
  const double tol = 1.0e-9 * std::sqrt(1.0 + n);
  double epsi = 1.0;
  int total = 0;
  while (epsi > tol) {
    int loop = 0;
    double err = INFINITY;
    while (err > 0.9 * epsi && loop < 100) {
      loop++;
      mma_xyz(S);
      mma_grad(S);
      S.grad = -1.0 * S.grad - epsi / S.lam;
      mma_hess(S);
      S.s1 = S.grad / S.hess;
      S.s2 = -S.mu + epsi / S.lam - S.s1 * S.mu / S.lam;
      mma_linesearch(S);
      mma_xyz(S);
      err = mma_residual(S, epsi);
      total++;
    }
    epsi = epsi * 0.1;
  }
  if (nNewton) *nNewton = total;
}

// ------------------------------------------------------------ run drivers
struct RunOut {
  double* xPhys = nullptr;     // numElements
  double* cHist = nullptr;     // nLoop
  double* volHist = nullptr;
  double* consHist = nullptr;
  double* sharpHist = nullptr;
  int* lssIts = nullptr;       // nLoop*nCases (lssIts(end+1), 121: one entry per solve)
  double* tHist = nullptr;     // nLoop*6
  double* xSnap = nullptr;     // nSnap*numElements
  int nSnap = 0;
  double* cSolid = nullptr;
  int* loops = nullptr;
  int* filterIts = nullptr;    // nLoop (sum of filter CG iterations)
  double* resHist = nullptr;   // nLoop*nCases: final relative residual of each elasticity solve
};

inline double pow16(double v) { double v2 = v * v, v4 = v2 * v2, v8 = v4 * v4; return v8 * v8; }
inline double pow15(double v) { double v2 = v * v, v4 = v2 * v2, v8 = v4 * v4; return v8 * v4 * v2 * v; }

// Compliance sensitivity with respect to the physical densities,
// dc = -DeSIMP(xPhys) .* ce / cSolid (108, 129 / 330).
void compliance_sensitivity(const Model& M, i64 n, const double* xPhys, const double* ce, double cSolid, double* dc) {
  const double E0 = M.S.E0, Emin = M.S.Emin;
  for (i64 e = 0; e < n; ++e) dc[e] = -(M.S.penal * (E0 - Emin) * (xPhys[e] * xPhys[e])) * (ce[e] / cSolid);
}

// Heaviside projection (380) and its derivative (340) at xTilde for one beta;
// either output may be null.
void heaviside(i64 n, double beta, double eta, const double* xt, double* xPhys, double* dx) {
  const double t0 = std::tanh(beta * eta);
  const double den = t0 + std::tanh(beta * (1 - eta));
  for (i64 e = 0; e < n; ++e) {
    const double th = std::tanh(beta * (xt[e] - eta));
    if (xPhys) xPhys[e] = (t0 + th) / den;
    if (dx) dx[e] = beta * (1 - th * th) / den;
  }
}

// p-norm local volume with p = 16 (pPIO, 235): the constraint value
// fval = mean((xhat/volMax)^p)^(1/p) - 1 (360) and its gradient with respect
// to xhat (335).
double pnorm_volume(i64 n, const double* xhat, const double* volMax, double* dfdx_pde) {
  const double p = 16.0;
  const double sp = det_sum(n, [&](i64 i) { return pow16(xhat[i]) / pow16(volMax[i]); }) / (double)n;
  const double fac = std::pow(sp, 1.0 / p - 1.0);
  for (i64 e = 0; e < n; ++e) dfdx_pde[e] = fac * (pow15(xhat[e]) / pow16(volMax[e])) / (double)n;  // 335
  return std::pow(sp, 1.0 / p) - 1.0;  // 360
}

// The LOCAL-mode sensitivity chain of one design iteration (334-350), filters
// in slots 0 (rMin) and 1 (rHat) set up:
//   xhat = F_rHat(xPhys); fval, dfdx_pde (p-norm); dx = H'(xTilde);
//   dc <- F_rMin(dc .* dx); dfdx = F_rMin(F_rHat(dfdx_pde) .* dx).
// Returns fval; t[0] += local-volume seconds, t[1] += filtering seconds.
double local_sensitivities(const Model& M, double beta, const double* xPhys, const double* xTilde,
                           const double* volMax, double* dc, double* dfdx, int* fIts, double* t) {
  const i64 nE = M.L[0].numElements;
  const double eta = 0.5;  // etaPIO, 234
  std::vector<double> xhat(nE), dfdx_pde(nE), dx(nE), tmp(nE);
  double t1 = now_s();
  *fIts += filter_run(M, M.filt[1], xPhys, xhat.data());  // 334
  const double fval = pnorm_volume(nE, xhat.data(), volMax, dfdx_pde.data());
  double t2 = now_s();
  t[0] += t2 - t1;
  heaviside(nE, beta, eta, xTilde, nullptr, dx.data());  // 340
  for (i64 e = 0; e < nE; ++e) tmp[e] = dc[e] * dx[e];
  *fIts += filter_run(M, M.filt[0], tmp.data(), dc);  // 341
  double t3 = now_s();
  t[1] += t3 - t2;
  *fIts += filter_run(M, M.filt[1], dfdx_pde.data(), dfdx);  // 346
  double t4 = now_s();
  t[0] += t4 - t3;
  for (i64 e = 0; e < nE; ++e) tmp[e] = dfdx[e] * dx[e];
  *fIts += filter_run(M, M.filt[0], tmp.data(), dfdx);  // 350
  t[1] += now_s() - t4;
  return fval;
}

int run(Model& M, int mode, double V0, int nLoop, double rMin, double rHat, RunOut& O) {
  const Level& L = M.L[0];
  const int nE = L.numElements;
  const double E0 = M.S.E0, Emin = M.S.Emin;
  setup_filter(M, rMin, M.filt[0]);
  if (mode == 1) setup_filter(M, rHat, M.filt[1]);
  std::vector<double> x(nE, V0), xPhys(nE, V0), xTilde(nE, V0), volMax(nE, V0);
  const int nCases = (int)M.F.size();
  std::vector<std::vector<double>> U(nCases, std::vector<double>((size_t)L.numNodes * 3, 0.0));  // 95 / 295
  std::vector<double> ce(nE), dc(nE), dv(nE, 1.0);
  // solid reference solve (95-106 / 295-306)
  M.eleModulus.assign(nE, E0);
  if (!assemble(M)) return -1;
  bool warned = false;
  for (int lc = 0; lc < nCases; ++lc) {  // 101 / 301
    solve_fea(M, lc, U[lc], &warned, nullptr);
    if (warned) M.warnings++;
  }
  compliance_cases(M, U, ce.data());
  const double cSolid = det_sum(nE, [&](i64 i) { return M.eleModulus[i] * ce[i]; });
  if (O.cSolid) *O.cSolid = cSolid;
  auto SIMP = [&](double xp) { return Emin + (xp * xp * xp) * (E0 - Emin); };       // 107
  double change = 1.0, sharp = 1.0, beta = 1.0;
  const double eta = 0.5;
  int loop = 0, loopbeta = 0;
  MMA mma;
  std::vector<double> tmp(nE), dfdx(nE);
  while (loop < nLoop && change > 0.0001 && sharp > 0.01) {  // 111 / 311
    const double t0 = now_s();
    loop++; loopbeta++;
    for (int e = 0; e < nE; ++e) M.eleModulus[e] = SIMP(xPhys[e]);  // 116 / 316
    const double ta = now_s();
    if (!assemble(M)) return -1;  // 118
    const double tb = now_s();
    for (int lc = 0; lc < nCases; ++lc) {  // 121 / 321: one warm-started solve per F_ column
      double resFinal = 0.0;
      const int its = solve_fea(M, lc, U[lc], &warned, &resFinal);
      if (warned) M.warnings++;
      if (O.lssIts) O.lssIts[(size_t)(loop - 1) * nCases + lc] = its;
      if (O.resHist) O.resHist[(size_t)(loop - 1) * nCases + lc] = resFinal;
    }
    const double tc = now_s();
    compliance_cases(M, U, ce.data());  // 124
    double cObj = det_sum(nE, [&](i64 i) { return M.eleModulus[i] * (ce[i] / cSolid); });  // 125-126
    const double cDesign = cObj * cSolid;  // 127
    const double V = det_sum(nE, [&](i64 i) { return xPhys[i]; }) / (double)nE;  // 128
    compliance_sensitivity(M, nE, xPhys.data(), ce.data(), cSolid, dc.data());  // 129 / 330
    double tOpt = now_s() - tc, tFilt = 0.0, tLVF = 0.0, fval = 0.0;
    int fIts = 0;
    if (mode == 0) {
      double t1 = now_s();
      for (int e = 0; e < nE; ++e) tmp[e] = x[e] * dc[e];  // 137
      fIts += filter_run(M, M.filt[0], tmp.data(), dc.data());
      for (int e = 0; e < nE; ++e) dc[e] = dc[e] / std::max(1e-3, x[e]);
      double t2 = now_s();
      tFilt += t2 - t1;
      fval = det_sum(nE, [&](i64 i) { return xPhys[i]; }) / (double)nE - V0;  // 145
      change = oc_update(M, x, xPhys, dc, V0, nullptr);
      xPhys = x;  // 160
      if (!M.isPassive.empty())
        for (int e = 0; e < nE; ++e)
          if (M.isPassive[e]) xPhys[e] = 1.0;  // 165
      double t3 = now_s();
      tOpt += t3 - t2;
    } else {
      double tt[2] = {0.0, 0.0};
      const double fvalP = local_sensitivities(M, beta, xPhys.data(), xTilde.data(), volMax.data(), dc.data(),
                                               dfdx.data(), &fIts, tt);
      tLVF += tt[0];
      tFilt += tt[1];
      double t5 = now_s();
      // MMA (354-369) over the active (non-passive) elements, 274-276
      std::vector<int> act;
      for (int e = 0; e < nE; ++e)
        if (M.isPassive.empty() || !M.isPassive[e]) act.push_back(e);
      const int na = (int)act.size();
      std::vector<double> xa(na), dca(na), dfa(na);
      for (int k = 0; k < na; ++k) { xa[k] = x[act[k]]; dca[k] = dc[act[k]]; dfa[k] = dfdx[act[k]]; }
      if (loop == 1) { mma.n = na; mma.xold1 = xa; mma.xold2 = xa; mma.low.assign(na, 0.0); mma.upp.assign(na, 0.0); }
      fval = fvalP;  // 360
      std::vector<double> xmin(na), xmax(na);
      for (int k = 0; k < na; ++k) { xmin[k] = std::max(0.0, xa[k] - 0.1); xmax[k] = std::min(1.0, xa[k] + 0.1); }
      mma_update(mma, loop, xa.data(), xmin.data(), xmax.data(), dca.data(), fval, dfa.data(), nullptr);
      change = 0.0;
      for (int k = 0; k < na; ++k) change = std::max(change, std::fabs(mma.x[k] - xa[k]));  // 368
      if (na < nE) x.assign(nE, 1.0);  // 366: x = onesArrSingle; x(activeEles) = xmma
      for (int k = 0; k < na; ++k) x[act[k]] = mma.x[k];  // 369
      double t6 = now_s();
      tOpt += t6 - t5;
      fIts += filter_run(M, M.filt[0], x.data(), xTilde.data());  // 374
      if (loopbeta >= 40) { beta = std::min(2 * beta, 128.0); loopbeta = 0; }  // 375-379
      heaviside(nE, beta, eta, xTilde.data(), xPhys.data(), nullptr);  // 380
      if (!M.isPassive.empty())
        for (int e = 0; e < nE; ++e)
          if (M.isPassive[e]) xPhys[e] = 1.0;  // 381
      tFilt += now_s() - t6;
    }
    sharp = 4.0 * det_sum(nE, [&](i64 i) { return xPhys[i] * (1.0 - xPhys[i]); }) / (double)nE;  // 166 / 382
    const double tTot = now_s() - t0;
    const int k = loop - 1;
    if (O.cHist) O.cHist[k] = cDesign;
    if (O.volHist) O.volHist[k] = V;
    if (O.consHist) O.consHist[k] = fval;
    if (O.sharpHist) O.sharpHist[k] = sharp;
    if (O.filterIts) O.filterIts[k] = fIts;
    if (O.tHist) {
      double* th = &O.tHist[(size_t)k * 6];
      th[0] = tb - ta; th[1] = tc - tb; th[2] = tOpt; th[3] = tFilt; th[4] = tLVF; th[5] = tTot;
    }
    if (O.xSnap && k < O.nSnap) std::memcpy(&O.xSnap[(size_t)k * nE], xPhys.data(), nE * sizeof(double));
  }
  if (O.xPhys) std::memcpy(O.xPhys, xPhys.data(), nE * sizeof(double));
  if (O.loops) *O.loops = loop;
  return 0;
}

}  // namespace

// ============================================================== C interface
extern "C" {

// Builds the model.  voxels: uint8 nely*nelx*nelz column-major (y fastest).
// nFix < 0 selects the built-in cuboid BCs (536-559, one load case, weight
// 1).  Otherwise fixNodes / loadNodes are 0-based node ids of the UNPADDED
// grid (nely+1, nelx+1, nelz+1), column-major; fixState is nFix*3 flags;
// the nCases load cases (loadingCond_, 481-493) are concatenated, case ii
// holding nPerCase[ii] nodes and 3 values per node; objective weights
// 1/nCases (493).
void* orc_create_multi(const unsigned char* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                       const unsigned char* fixState, int nCases, const int* nPerCase, const int* loadNodes,
                       const double* loadVec) {
  Model* M = new Model();
  M->nely = nely; M->nelx = nelx; M->nelz = nelz;
  if (nelx < 3 || nely < 3 || nelz < 3) { M->err = "Inappropriate Input Model!"; return M; }  // 532-534
  discretize(*M, voxels);
  M->Ke = element_stiffness(M->S);
  const Level& L0 = M->L[0];
  std::vector<int> fixDense;
  std::vector<std::vector<int>> loadDense;
  std::vector<std::vector<double>> lv;
  std::vector<unsigned char> fs;
  if (nFix < 0) {
    loadDense.resize(1);
    lv.resize(1);
    cuboid_bc(*M, fixDense, loadDense[0]);
    fs.assign(fixDense.size() * 3, 1);
    const double q = 1.0 / (double)loadDense[0].size();  // 556
    for (size_t i = 0; i < loadDense[0].size(); ++i) {
      lv[0].push_back(0.0 * q); lv[0].push_back(0.0 * q); lv[0].push_back(-1.0 * q);
    }
    M->objWeight.assign(1, 1.0);  // 558
  } else {
    // AdaptBCExternalMdl (1439-1489): unpadded -> padded node ids
    const i64 uy = nely + 1, ux = nelx + 1;
    const i64 PY = L0.resY + 1, PX = L0.resX + 1;
    auto adapt = [&](int id) {
      i64 jy = id % uy, jx = (id / uy) % ux, jz = id / uy / ux;
      return (int)(jy + PY * (jx + PX * jz));
    };
    for (int i = 0; i < nFix; ++i) { fixDense.push_back(adapt(fixNodes[i])); for (int c = 0; c < 3; ++c) fs.push_back(fixState[i * 3 + c]); }
    if (nCases < 1) { M->err = "No Loads!"; return M; }
    loadDense.resize(nCases);
    lv.resize(nCases);
    int off = 0;
    for (int ii = 0; ii < nCases; ++ii)
      for (int i = 0; i < nPerCase[ii]; ++i, ++off) {
        loadDense[ii].push_back(adapt(loadNodes[off]));
        for (int c = 0; c < 3; ++c) lv[ii].push_back(loadVec[(size_t)off * 3 + c]);
      }
    M->objWeight.assign(nCases, 1.0 / nCases);  // 493
  }
  for (size_t i = 0; i < fixDense.size(); ++i) {
    int k = L0.nodMapForward[fixDense[i]] - 1;
    if (k < 0) continue;
    M->fixNode.push_back(k);
    for (int c = 0; c < 3; ++c) M->fixState.push_back(fs[i * 3 + c]);
  }
  M->loadNode.assign(loadDense.size(), {});
  M->loadVal.assign(loadDense.size(), {});
  for (size_t ii = 0; ii < loadDense.size(); ++ii)
    for (size_t i = 0; i < loadDense[ii].size(); ++i) {
      int k = L0.nodMapForward[loadDense[ii][i]] - 1;
      if (k < 0) continue;
      M->loadNode[ii].push_back(k);
      for (int c = 0; c < 3; ++c) M->loadVal[ii].push_back(lv[ii][i * 3 + c]);
    }
  if (!apply_bc(*M)) return M;
  build_hierarchy(*M);
  if (!setup_bc_ke(*M)) return M;
  M->bcOk = true;
  return M;
}

// One load case.
void* orc_create(const unsigned char* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                 const unsigned char* fixState, int nLoad, const int* loadNodes, const double* loadVec) {
  return orc_create_multi(voxels, nely, nelx, nelz, nFix, fixNodes, fixState, 1, &nLoad, loadNodes, loadVec);
}

void orc_destroy(void* h) { delete (Model*)h; }
const char* orc_last_error(void* h) { return ((Model*)h)->err.c_str(); }
int orc_ok(void* h) { return ((Model*)h)->bcOk ? 1 : 0; }
int orc_warnings(void* h) { return ((Model*)h)->warnings; }

// info[0] = numLevels; per level l (offset 1+8l): resX,resY,resZ,span,
// numElements,numNodes,intermediateNumNodes,0
int orc_info(void* h, int* info) {
  Model* M = (Model*)h;
  info[0] = M->numLevels;
  for (int l = 0; l < M->numLevels; ++l) {
    const Level& L = M->L[l];
    int* o = info + 1 + 8 * l;
    o[0] = L.resX; o[1] = L.resY; o[2] = L.resZ; o[3] = L.span;
    o[4] = L.numElements; o[5] = L.numNodes; o[6] = L.intermediateNumNodes; o[7] = 0;
  }
  return 0;
}

// Copies a named array out.  Returns its length (elements) or -1.
long long orc_get(void* h, const char* name, int level, void* out, long long cap) {
  Model* M = (Model*)h;
  std::string n(name);
  auto cp = [&](const void* src, size_t cnt, size_t sz) -> long long {
    if (out && (long long)cnt <= cap) std::memcpy(out, src, cnt * sz);
    return (long long)cnt;
  };
  if (level < 0 || level >= (int)M->L.size()) return -1;
  const Level& L = M->L[level];
  if (n == "eNodMat") return cp(L.eNodMat.data(), L.eNodMat.size(), sizeof(int));
  if (n == "eleMapBack") return cp(L.eleMapBack.data(), L.eleMapBack.size(), sizeof(int));
  if (n == "nodMapBack") return cp(L.nodMapBack.data(), L.nodMapBack.size(), sizeof(int));
  if (n == "elementUpwardMap") return cp(L.elementUpwardMap.data(), L.elementUpwardMap.size(), sizeof(int));
  if (n == "transferMat") return cp(L.transferMat.data(), L.transferMat.size(), sizeof(int));
  if (n == "transferMatCoeffi") return cp(L.transferMatCoeffi.data(), L.transferMatCoeffi.size(), sizeof(double));
  if (n == "solidNodeMap") return cp(L.solidNodeMapCoarser2Finer.data(), L.solidNodeMapCoarser2Finer.size(), sizeof(int));
  if (n == "diagK") return cp(L.diagK.data(), L.diagK.size(), sizeof(double));
  if (n == "Ks") return cp(L.Ks.data(), L.Ks.size(), sizeof(double));
  if (n == "RI") return cp(L.RI.data(), L.RI.size(), sizeof(double));
  if (n == "Ke") return cp(M->Ke.data(), M->Ke.size(), sizeof(double));
  if (n == "F") {  // level: the load case
    if (level < 0 || level >= (int)M->F.size()) return -1;
    return cp(M->F[level].data(), M->F[level].size(), sizeof(double));
  }
  if (n == "objWeight") return cp(M->objWeight.data(), M->objWeight.size(), sizeof(double));
  if (n == "fixNode") return cp(M->fixNode.data(), M->fixNode.size(), sizeof(int));
  if (n == "isPassive") return cp(M->isPassive.data(), M->isPassive.size(), 1);
  if (n == "elementsOnBoundary") return cp(L.elementsOnBoundary.data(), L.elementsOnBoundary.size(), sizeof(int));
  if (n == "fixState") return cp(M->fixState.data(), M->fixState.size(), 1);
  if (n == "mapUniqueKes") return cp(M->mapUniqueKes.data(), M->mapUniqueKes.size(), sizeof(int));
  if (n == "KesFree") return cp(M->KesFree.data(), M->KesFree.size(), sizeof(double));
  if (n == "KesFixed") return cp(M->KesFixed.data(), M->KesFixed.size(), sizeof(double));
  if (n == "filterKernel0") return cp(M->filt[0].kernel, 64, sizeof(double));
  if (n == "filterDiag0") return cp(M->filt[0].diagPrecond.data(), M->filt[0].diagPrecond.size(), sizeof(double));
  if (n == "nodesOnBoundary") return cp(L.nodesOnBoundary.data(), L.nodesOnBoundary.size(), sizeof(int));
  return -1;
}

int orc_set_modulus(void* h, const double* E) {
  Model* M = (Model*)h;
  M->eleModulus.assign(E, E + M->L[0].numElements);
  return 0;
}
int orc_assemble(void* h) { return assemble(*(Model*)h) ? 0 : -1; }
int orc_apply_K(void* h, const double* u, double* y) { apply_K(*(Model*)h, u, y); return 0; }
int orc_vcycle(void* h, const double* r, double* x) { precondition(*(Model*)h, r, x); return 0; }
// typeVcycle_ (34): 0 'Adapted', 1 'Standard'.  Ks is kept on every coarse
// level either way (997-999 keeps it for 'Standard').
int orc_set_vcycle(void* h, int type) {
  if (type != 0 && type != 1) return 1;
  ((Model*)h)->S.typeVcycle = type;
  return 0;
}
// tol_ (30) and maxIT_ (31) of the elasticity solves.
int orc_set_tol(void* h, double tol, int maxIT) {
  ((Model*)h)->S.tol = tol;
  ((Model*)h)->S.maxIT = maxIT;
  return 0;
}
// weightFactorJacobi_ (32).
int orc_set_omega(void* h, double w) {
  ((Model*)h)->S.omega = w;
  return 0;
}
// K*u on one level (0: 641-726, >0: 727-744), compact node order.
int orc_apply_K_level(void* h, int level, const double* u, double* y) {
  Model* M = (Model*)h;
  if (level < 0 || level >= M->numLevels) return 1;
  apply_K_any(*M, level, u, y);
  return 0;
}
int orc_coarse_solve(void* h, const double* r, double* x) { coarse_solve(*(Model*)h, r, x); return 0; }
int orc_restrict(void* h, int level, const double* rF, double* rC) {
  restrict_residual(((Model*)h)->L[level], rF, rC); return 0;
}
int orc_interpolate(void* h, int level, const double* xC, double* xF) {
  interpolate(((Model*)h)->L[level], xC, xF); return 0;
}
// PCG on the elasticity system; x holds x0 on entry.
int orc_pcg(void* h, const double* b, double* x, double tol, int maxIT, double* res) {
  Model* M = (Model*)h;
  const size_t n = (size_t)M->L[0].numNodes * 3;
  std::vector<double> bb(b, b + n), y(x, x + n);
  auto A = [&](const double* u, double* yy) { apply_K(*M, u, yy); };
  auto P = [&](const double* r, double* z) { precondition(*M, r, z); };
  int its = pcg(A, P, bb, y, tol, maxIT, res, nullptr);
  std::memcpy(x, y.data(), n * sizeof(double));
  return its;
}
int orc_compliance(void* h, const double* U, double* ce) { compliance(*(Model*)h, U, ce); return 0; }
int orc_filter_setup(void* h, int slot, double r) { Model* M = (Model*)h; setup_filter(*M, r, M->filt[slot]); return 0; }
int orc_filter_apply(void* h, int slot, const double* u, double* y) {
  Model* M = (Model*)h; filter_apply(*M, M->filt[slot], u, y); return 0;
}
int orc_filter(void* h, int slot, const double* src, double* dst) {
  Model* M = (Model*)h; return filter_run(*M, M->filt[slot], src, dst);
}
double orc_oc(void* h, double* x, const double* xPhys, const double* dc, double V0, int* nBisect) {
  Model* M = (Model*)h;
  const int n = M->L[0].numElements;
  std::vector<double> xx(x, x + n), xp(xPhys, xPhys + n), d(dc, dc + n);
  double ch = oc_update(*M, xx, xp, d, V0, nBisect);
  std::memcpy(x, xx.data(), n * sizeof(double));
  return ch;
}
// One MMA update (m = 1) on n variables.  State arrays (xold1, xold2, low,
// upp) are read and written.  Returns the number of Newton steps.
// sub (may be null): 6n + 3 doubles receiving the subproblem GenSub built
// (alpha, beta, p0, q0, pij, qij: n each; then b) and the final dual
// variables (lambda, mu) of SolveDIP.
int orc_mma_ex(int n, int iter, double* xval, const double* xmin, const double* xmax, double* xold1, double* xold2,
               const double* dfdx, double gx, const double* dgdx, double* low, double* upp, double* sub) {
  MMA S;
  S.n = n;
  S.xold1.assign(xold1, xold1 + n); S.xold2.assign(xold2, xold2 + n);
  S.low.assign(low, low + n); S.upp.assign(upp, upp + n);
  int nn = 0;
  mma_update(S, iter, xval, xmin, xmax, dfdx, gx, dgdx, &nn);
  std::memcpy(xval, S.x.data(), n * sizeof(double));
  std::memcpy(xold1, S.xold1.data(), n * sizeof(double));
  std::memcpy(xold2, S.xold2.data(), n * sizeof(double));
  std::memcpy(low, S.low.data(), n * sizeof(double));
  std::memcpy(upp, S.upp.data(), n * sizeof(double));
  if (sub) {
    const std::vector<double>* parts[6] = {&S.alpha, &S.beta, &S.p0, &S.q0, &S.pij, &S.qij};
    for (int k = 0; k < 6; ++k) std::memcpy(sub + (size_t)k * n, parts[k]->data(), n * sizeof(double));
    sub[6 * (size_t)n] = S.b;
    sub[6 * (size_t)n + 1] = S.lam;
    sub[6 * (size_t)n + 2] = S.mu;
  }
  return nn;
}
int orc_mma(int n, int iter, double* xval, const double* xmin, const double* xmax, double* xold1, double* xold2,
            const double* dfdx, double gx, const double* dgdx, double* low, double* upp) {
  return orc_mma_ex(n, iter, xval, xmin, xmax, xold1, xold2, dfdx, gx, dgdx, low, upp, nullptr);
}

// Heaviside projection (380) and derivative (340); xPhys or dx may be null.
void orc_heaviside(int n, double beta, double eta, const double* xTilde, double* xPhys, double* dx) {
  heaviside(n, beta, eta, xTilde, xPhys, dx);
}
// p-norm local-volume constraint (360) and its gradient in xhat (335).
double orc_pnorm(int n, const double* xhat, const double* volMax, double* dfdx_pde) {
  return pnorm_volume(n, xhat, volMax, dfdx_pde);
}
// dc = -DeSIMP(xPhys) .* ce / cSolid (129 / 330).
int orc_compliance_sens(void* h, const double* xPhys, const double* ce, double cSolid, double* dc) {
  Model* M = (Model*)h;
  compliance_sensitivity(*M, M->L[0].numElements, xPhys, ce, cSolid, dc);
  return 0;
}
// The LOCAL sensitivity chain (334-350) with filter slots 0 (rMin) and 1
// (rHat) set up and volMax = Ve0 (278): dc is filtered in place, dfdx
// receives the constraint gradient; returns fval (360).
double orc_local_sens(void* h, double beta, const double* xPhys, const double* xTilde, double Ve0, double* dc,
                      double* dfdx, int* filterIts) {
  Model* M = (Model*)h;
  const int nE = M->L[0].numElements;
  std::vector<double> volMax(nE, Ve0);
  double t[2] = {0.0, 0.0};
  int its = 0;
  const double f = local_sensitivities(*M, beta, xPhys, xTilde, volMax.data(), dc, dfdx, &its, t);
  if (filterIts) *filterIts = its;
  return f;
}

// Whole optimisation (mode 0 = GLOBAL / TOP3D_XL_TO, 1 = LOCAL / TOP3D_XL_PIO).
int orc_run(void* h, int mode, double V0, int nLoop, double rMin, double rHat, double* xPhys, double* cHist,
            double* volHist, double* consHist, double* sharpHist, int* lssIts, double* tHist, double* xSnap,
            int nSnap, double* cSolid, int* loops, int* filterIts, double* resHist) {
  Model* M = (Model*)h;
  if (!M->bcOk) return -1;
  RunOut O;
  O.xPhys = xPhys; O.cHist = cHist; O.volHist = volHist; O.consHist = consHist; O.sharpHist = sharpHist;
  O.lssIts = lssIts; O.tHist = tHist; O.xSnap = xSnap; O.nSnap = nSnap; O.cSolid = cSolid; O.loops = loops;
  O.filterIts = filterIts;
  O.resHist = resHist;
  return run(*M, mode, V0, nLoop, rMin, rHat, O);
}

int orc_num_threads(void) { return omp_get_max_threads(); }

// Selects the k-th best coarsest-level node ordering (0 = smallest bandwidth).
// Every ordering gives the same factor in exact arithmetic, as CHOLMOD's
// fill-reducing permutation does; it only changes FP64 rounding.
void orc_set_order_variant(void* h, int k) { ((Model*)h)->orderVariant = k; }

// passiveElements_ (472, 524-527): 0-based voxel ids of the unpadded
// nely x nelx x nelz grid, mapped to compact elements; void voxels ignored.
void orc_set_passive(void* h, int n, const int* voxelIds) {
  Model* M = (Model*)h;
  const Level& L = M->L[0];
  M->isPassive.assign(L.numElements, 0);
  for (int i = 0; i < n; ++i) {
    const int v = voxelIds[i];
    if (v < 0 || v >= M->nely * M->nelx * M->nelz) continue;
    const int iy = v % M->nely, ix = (v / M->nely) % M->nelx, iz = v / (M->nely * M->nelx);
    const int dense = iy + L.resY * (ix + L.resX * iz);
    const int c = L.eleMapForward[dense];
    if (c > 0) M->isPassive[c - 1] = 1;
  }
}

// TopOpti_SetPassiveElements, 1511-1584, on the compact lists: layer 1 from
// elementsOnBoundary (1526) or from the boundary elements around the loaded
// / fixed nodes (nodeStruct_, 1536-1546; the nodes are taken as the loaded
// and fixed compact nodes themselves), further layers by
// Common_IncludeAdjacentElements (1587-1615): the 3x3x3 neighbourhood of
// each element on the dense grid, kept where it is solid.
void orc_set_passive_layers(void* h, int nB, int nLd, int nFx) {
  Model* M = (Model*)h;
  const Level& L = M->L[0];
  if (M->isPassive.size() != (size_t)L.numElements) M->isPassive.assign(L.numElements, 0);
  auto adjacent = [&](const std::vector<int>& in) {  // 1587-1615
    std::vector<int> out;
    for (int e : in) {
      const int d = L.eleMapBack[e];
      const int iy = d % L.resY, ix = (d / L.resY) % L.resX, iz = d / (L.resY * L.resX);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy) {
            const int y = iy + dy, x = ix + dx, z = iz + dz;
            if (y < 0 || x < 0 || z < 0 || y >= L.resY || x >= L.resX || z >= L.resZ) continue;
            const int c = L.eleMapForward[y + L.resY * (x + L.resX * z)];
            if (c > 0) out.push_back(c - 1);
          }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  };
  auto grow = [&](std::vector<int> set, int n) {
    for (int k = 2; k <= n; ++k) set = adjacent(set);
    return set;
  };
  auto mark = [&](const std::vector<int>& set) { for (int e : set) M->isPassive[e] = 1; };
  if (nB > 0) mark(grow(L.elementsOnBoundary, nB));
  if (nLd > 0 || nFx > 0) {
    std::vector<char> isBE(L.numElements, 0);
    for (int e : L.elementsOnBoundary) isBE[e] = 1;
    auto near = [&](const std::vector<char>& nodeFlag) {
      std::vector<int> out;
      for (int e = 0; e < L.numElements; ++e) {
        if (!isBE[e]) continue;
        for (int j = 0; j < 8; ++j)
          if (nodeFlag[L.eNodMat[(size_t)e * 8 + j]]) { out.push_back(e); break; }
      }
      return out;
    };
    if (nLd > 0) {
      std::vector<char> f(L.numNodes, 0);
      for (const auto& c : M->loadNode) for (int n : c) f[n] = 1;
      mark(grow(near(f), nLd));
    }
    if (nFx > 0) {
      std::vector<char> f(L.numNodes, 0);
      for (int n : M->fixNode) f[n] = 1;
      mark(grow(near(f), nFx));
    }
  }
}

}  // extern "C"
