// Internal declarations shared by the libt3x translation units.
//
// Layout on the device (every level of the multigrid hierarchy):
//   * a dense, padded structured grid of ny x nx x nz hex8 cells and
//     (ny+1) x (nx+1) x (nz+1) nodes, y fastest, then x, then z — the order
//     of TOP3D_XL.m:1684-1711, so the reference's compact numbering is a
//     masked subsequence of ours;
//   * node vectors are structure-of-arrays: component c of node n lives at
//     c*NN + n, NN = NY*NX*NZ;
//   * cell (element) arrays reuse the node indexing: cell (iy,ix,iz) lives at
//     iy + NY*(ix + NX*iz), so a cell and its "lower" corner node share an
//     index and the K*u stencil needs a single index space;
//   * cells outside the design domain carry E = 0 and nodes outside carry
//     mask 0; vectors are kept exactly 0 there.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace t3x {

// Local hex8 node j <-> (dy, dx, dz) offset from the cell's (iy,ix,iz) node.
// Node 1 sits one row down (TOP3D_XL.m:1697, diagram 1635-1648).
constexpr int kLocDy[8] = {1, 1, 0, 0, 1, 1, 0, 0};
constexpr int kLocDx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
constexpr int kLocDz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// (dy,dx,dz) -> local index
constexpr int loc_index(int dy, int dx, int dz) {
  return dz * 4 + (dy ? (dx ? 1 : 0) : (dx ? 2 : 3));
}

// Upper-triangle packing of a symmetric 24x24 block: p <= q.
constexpr int kTri = 300;
__host__ __device__ inline int tri_index(int p, int q) {
  if (p > q) { int t = p; p = q; q = t; }
  return q * (q + 1) / 2 + p;
}

struct Level {
  int ny = 0, nx = 0, nz = 0;  // cells
  int NY = 0, NX = 0, NZ = 0;  // nodes
  long long NN = 0;            // NY*NX*NZ
  int span = 0;                // coarsening factor from the finer level (0 on the finest)
  int nCells = 0, nNodes = 0;  // solid counts
  std::vector<uint8_t> hEmask, hNmask;
  uint8_t* emask = nullptr;  // NN (cell indexing)
  uint8_t* nmask = nullptr;  // NN
  double* diag = nullptr;    // 3*NN, Jacobi diagonal (levels < L)
  double* r = nullptr;       // 3*NN, V-cycle residual (levels >= 2)
  double* x = nullptr;       // 3*NN, V-cycle correction (levels >= 2)
  double* Ks = nullptr;      // kTri*NN, Galerkin cell matrices (levels >= 2), entry-major
  float* rf = nullptr;       // mixed-precision V-cycle vectors
  float* xf = nullptr;
  double* t = nullptr;       // FP64 'Standard' cycle: defect / smoother output (levels 1..L-2)
  float* tf = nullptr;       // FP32 'Standard' cycle: defect / smoother output (levels 1..L-2)
  float* diagf = nullptr;
  float* Ksf = nullptr;      // FP32 copy of Ks (levels 1..L-2, 'Standard' cycle)
  __half* Ksh = nullptr;     // FP16 copy instead (precision 2): Ks / Kscale per cell
  float* Kscale = nullptr;   // its per-cell power-of-two scales
  int* sfTiles = nullptr;      // finest level: K*u tiles with a solid cell (null: all)
  int nSfTiles = 0;
};

struct Filter {
  double kernel[64];        // 8x8 Helmholtz cell matrix, column-major
  double* diagInv = nullptr;  // NN
  bool ready = false;
  double r = 0.0;
};

struct Problem;

// ------------------------------------------------------------ error helper
#define T3X_HIP(call)                                                            \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      throw t3x::HipError(std::string(#call) + ": " + hipGetErrorString(e_));    \
    }                                                                            \
  } while (0)

struct HipError {
  std::string msg;
  explicit HipError(std::string m) : msg(std::move(m)) {}
};

}  // namespace t3x
