/*
 * libt3x serving engine: many topology optimisations of uneven sizes on one
 * GPU with continuous admission.
 *
 * No counterpart in the reference (TOP3D_XL.m runs one optimisation per
 * MATLAB call).  Each request is a full TOP3D_XL call on a cuboid
 * (nely, nelx, nelz) with the built-in cantilever supports (x = 0 face fixed,
 * 536-541) and a load of chosen direction and magnitude on the built-in load
 * patch (x = nelx face, lowest round(nelz/6+1) node layers, 552-556).
 *
 * Scheduling: requests wait in arrival order; a request is admitted when a
 * worker slot is free and its device footprint (t3x_footprint) fits in the
 * memory budget next to the requests already running.  Each admitted request
 * owns a problem handle with its own HIP stream and device state, so the kernels of
 * concurrent requests overlap on the device; a finished request frees its
 * memory and slot for the next one at once.  A request's arithmetic does not
 * depend on what else runs: its results are bitwise the same alone and in a
 * mix.
 */
#ifndef T3X_SERVE_H_
#define T3X_SERVE_H_

#include "t3x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct t3x_server t3x_server;

typedef struct t3x_server_config {
  long long memBudget; /* device bytes for admitted requests; <= 0: 90% of free memory */
  int maxActive;       /* concurrent requests (worker threads), >= 1 */
  t3x_params params;   /* solver parameters of every request */
} t3x_server_config;

typedef struct t3x_request {
  long long tag;       /* caller's id, returned in the completion */
  int nely, nelx, nelz;
  int mode;            /* 0 GLOBAL, 1 LOCAL */
  double V0;           /* volume fraction (GLOBAL) / local volume bound (LOCAL) */
  int nLoop;
  double rMin, rHat;
  double loadDir[3];   /* load direction (normalised by the server) */
  double loadMag;      /* total load magnitude, split evenly over the patch nodes */
} t3x_request;

typedef struct t3x_completion {
  long long tag;
  int status;          /* T3X_OK, T3X_WARN_MAXIT or an error */
  int loops, pcgIts, nE;
  double tSubmit, tStart, tEnd; /* seconds since t3x_server_create */
  double cFinal, cSolid;
  long long footprint; /* admitted device bytes */
  const double* xPhys; /* nE values; valid until t3x_server_release(tag) */
  const double* cHist; /* loops values; same lifetime */
} t3x_completion;

int t3x_server_create(const t3x_server_config* cfg, t3x_server** out);
/* Queues a request (returns at once).  T3X_ERR_INPUT if it can never fit the
 * budget, or if its tag names a request that is queued, running, or finished
 * and not yet released (a tag is unique until t3x_server_release). */
int t3x_server_submit(t3x_server* s, const t3x_request* req);
/* Up to max finished requests not returned before; waits up to waitMs for the first. */
int t3x_server_poll(t3x_server* s, t3x_completion* out, int max, int waitMs, int* n);
/* Frees the result arrays of a returned completion. */
int t3x_server_release(t3x_server* s, long long tag);
/* queued, running, finished-not-polled counts and device bytes in use. */
int t3x_server_stats(t3x_server* s, int* queued, int* running, int* ready, long long* memInUse);
/* Waits for every queued and running request, then stops the workers. */
void t3x_server_destroy(t3x_server* s);

#ifdef __cplusplus
}
#endif

#endif /* T3X_SERVE_H_ */
