// Coarsest-level banded Cholesky factor + block-bidiagonal solves (see t3x_coarse.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace t3x {

// Device words of the one-launch recurrences (k_solve_chain), zeroed at set-up:
// per work-group the tag its next launch starts from (equal across the
// groups between launches), and a failure word set when a wait expired
// (cleared by the host).
constexpr int kChainMaxSB = 2048;            // largest solve block of the one-launch recurrences
constexpr int kChainMaxWG = kChainMaxSB / 4;  // their work-groups (one row per wave, 4 waves)
struct ChainState {
  unsigned tag[kChainMaxWG];
  int fail;
  int pad[3];
};

struct CoarseSolver {
  int n = 0;     // true order
  int bw = 0;    // bandwidth
  int SB = 0;    // solve block (multiple of 64, >= bw + 1)
  int np = 0;    // padded order (multiple of SB)
  int nblk = 0;  // np / SB
  ChainState* chain = nullptr;  // device words of the one-launch recurrences (null: a launch per block)
  bool one_launch() const { return chain && SB <= kChainMaxSB; }
  // parts per million of the current device's work-group slots (its
  // occupancy for the recurrences' registers and LDS) that one launch of the
  // one-launch recurrences needs; above 1e6 (or on error) they cannot run
  long long chain_share() const;
  void plan(int n, int bw);
  size_t band_len() const;
  size_t tinv_len() const { return (size_t)np * 64; }
  // two panels of row tiles (64x64) below a diagonal tile: the factorization's scratch
  size_t panel_len() const { return 2 * (size_t)((bw + 63) / 64 + 1) * 64 * 64; }
  size_t blk_len() const { return (size_t)nblk * SB * SB; }
  void pad_identity(hipStream_t s, double* band);
  void factor(hipStream_t s, double* band, double* tinv, double* pan, double* Dinv, double* DT, double* Gm,
              double* H, int* status);
  // FP64 solve; the right-hand side is gathered from / the solution scattered
  // to dense level vectors of either precision.  va, vb: np doubles each (a
  // launch per block) or 3 np (one launch per recurrence: np doubles, then
  // 2 np tagged words).
  void solve(hipStream_t s, const long long* map, const double* rdense, double* xdense, const double* Dinv,
             const double* DT, const double* Gm, const double* H, double* va, double* vb, const int* fl) const;
  void solve(hipStream_t s, const long long* map, const float* rdense, float* xdense, const double* Dinv,
             const double* DT, const double* Gm, const double* H, double* va, double* vb, const int* fl) const;
};

}  // namespace t3x
