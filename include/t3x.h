/*
 * libt3x — MI355X (gfx950) implementation of the TOP3D_XL topology-
 * optimisation hot path behind a plain C ABI.
 *
 * Every entry point replaces a function (or a loop body) of the reference
 * MATLAB file TOP3D_XL.m; the cited lines are that file's.  Host buffers are
 * caller-owned and use the reference's compact ordering:
 *   - cells   : the solid voxels of the padded grid in column-major (y,x,z)
 *               order (eleMapBack, TOP3D_XL.m:1684-1689), 0-based;
 *   - nodes   : the nodes touching a solid voxel, same order (nodMapBack,
 *               1693-1707); nodal vectors interleave (ux,uy,uz) per node (1320).
 * Device state is library-owned behind the opaque handle.  Functions return a
 * T3X_* status and never throw; t3x_last_error() gives the message.
 * A handle is not thread-safe; each handle owns one HIP stream.
 */
#ifndef T3X_H_
#define T3X_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct t3x_problem t3x_problem;

enum {
  T3X_OK = 0,
  T3X_ERR_INPUT = 1,      /* error('Inappropriate Input Model!') 533, error('Wrong Input!') 16 */
  T3X_ERR_NO_LOADS = 2,   /* warning('No Loads!') 1309 */
  T3X_ERR_NO_FIX = 3,     /* warning('No Fixations!') 1310 */
  T3X_ERR_HIP = 4,        /* HIP runtime failure (message in t3x_last_error) */
  T3X_ERR_NOT_PD = 5,     /* coarsest matrix not positive definite (chol, 1065) */
  T3X_ERR_NO_DEVICE = 6,  /* no GPU visible: there is no CPU fallback */
  T3X_WARN_MAXIT = 100    /* warning('Exceed the maximum iterate numbers') 634-637 */
};

/* InitialSettings (TOP3D_XL.m:21-38) plus build options. */
typedef struct t3x_params {
  double E0;             /* modulus_            (24) 1.0 */
  double nu;             /* poissonRatio_       (25) 0.3 */
  double Emin;           /* modulusMin_         (26) 1e-6 */
  double penal;          /* SIMPpenalty_        (27) 3 */
  double tol;            /* tol_                (30) 1e-3 */
  int maxIT;             /* maxIT_              (31) 800 */
  double omega;          /* weightFactorJacobi_ (32) 0.6 */
  int coarsestControl;   /* coarsestResolutionControl_ (33) 50000; <= 0: the
                            count for 50000, deepened while the coarsest band
                            Cholesky would cost more than 1e11 flop (an option
                            of this library) */
  int precision;         /* 0 = FP64 plain path;
                            1 = mixed: FP64 CG (K*p, vectors, dots, step
                                lengths) preconditioned by the V-cycle with FP32
                                level vectors, diagonals and smoother output
                                around the FP64 coarsest Cholesky;
                            2 = as 1, with the 'Standard' cycle's coarse-level
                                Galerkin matrices stored in FP16 with a
                                power-of-two scale per cell (FP32 products) */
  int checkEvery;        /* PCG convergence flag read-back interval (>=1) */
  int timing;            /* 1 = time kernel groups with HIP events (t3x_kernel_time);
                            2 = only the fine K*p group (fewer event packets) */
  int vcycle;            /* typeVcycle_ (34): 0 = 'Adapted' (747-781); 1 = 'Standard'
                            (V(1,1) damped Jacobi with K*u on every level, 759-777;
                            coarse levels through the stored Galerkin matrices,
                            727-744): FP64 with precision = 0, FP32 level
                            vectors and matrices with precision = 1 */
  int graphs;            /* 1 = each PCG iteration is captured once as a HIP graph and
                            replayed (not while timing = 1); 0 (default) =
                            kernel-by-kernel launches.  Results are identical
                            either way; on MI355X the speed is too. */
} t3x_params;

/* Fills the reference defaults (InitialSettings, 21-38). */
void t3x_default_params(t3x_params* p);

/*
 * Builds the voxel FEA model, boundary conditions and multigrid hierarchy.
 * Replaces CreateVoxelFEAmodel (442-561), FEA_VoxelBasedDiscretization
 * (1629-1730), FEA_ApplyBoundaryCondition (1294-1336), FEA_SetupVoxelBased
 * (1238-1252) and Solving_SetupKeWithFixedDOFs (1338-1437).
 *   voxels : uint8 nely*nelx*nelz, column-major (y fastest) — inputModel
 *   nFix < 0 : the built-in cuboid BCs (536-559)
 *   otherwise fixNodes/loadNodes are 0-based node ids of the UNPADDED
 *   (nely+1)x(nelx+1)x(nelz+1) grid (the .TopVoxel convention, 474-481,
 *   adapted as in AdaptBCExternalMdl 1439-1489), fixState nFix*3 flags,
 *   loadVec nLoad*3 forces.
 * Every listed fixed DOF is fixed (P_free K P_free + I on the fixed rows).
 * The reference applies a fixation only at boundary nodes (it intersects
 * fixingCond_ with nodesOnBoundary, 1369) and then pairs the ii-th common
 * node's elements with fixingCond_'s ii-th row (1379-1383), so a fixation
 * at a node inside the domain is dropped and misaligns the later ones; for
 * fixations on the boundary (supports on the surface, the built-in cuboid)
 * the two agree.
 */
int t3x_create(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
               const uint8_t* fixState, int nLoad, const int* loadNodes, const double* loadVec,
               const t3x_params* params, t3x_problem** out);
/*
 * t3x_create with several load cases (loadingCond_{1..n}, the .TopVoxel
 * "Additional Loads" of 483-493; F_ columns, 1315-1321).  The cases are
 * concatenated in loadNodes/loadVec, case ii holding nPerCase[ii] nodes.
 * A node loaded in any case is released from the fixings (1301-1305).
 * Each design iteration solves one warm-started PCG per case (101, 121,
 * 301, 321) and the element compliance is sum_j ce_j / nCases
 * (objWeightingList_, 493, 597).  nFix < 0 (the built-in cuboid) ignores
 * the load arguments and has one case of weight 1 (557-558).
 * t3x_result.lssIts and .resHist then hold nLoop*nCases entries, one per
 * solve, case fastest (lssIts(end+1), 121); t3x_iter.pcgIts is the sum over
 * the cases and t3x_iter.res the largest final residual.
 */
int t3x_create_multi(const uint8_t* voxels, int nely, int nelx, int nelz, int nFix, const int* fixNodes,
                     const uint8_t* fixState, int nCases, const int* nPerCase, const int* loadNodes,
                     const double* loadVec, const t3x_params* params, t3x_problem** out);
int t3x_num_load_cases(const t3x_problem* h);
/* On failure, *out still receives a handle (unless out is NULL) so that
 * t3x_last_error() can report why; the caller destroys it either way. */
void t3x_destroy(t3x_problem* h);
const char* t3x_last_error(const t3x_problem* h);
/* Passive elements (passiveElements_, 472, 524-527): voxel ids (0-based,
 * column-major, unpadded nely x nelx x nelz grid) held solid: OC sets them to
 * 1 (150, 165), LOCAL excludes them from MMA and sets x and xPhys to 1
 * (274-276, 366, 381).  Call before t3x_run / t3x_opt_begin; void voxels are
 * ignored; replaces any earlier set. */
int t3x_set_passive(t3x_problem* h, int n, const int* voxelIds);
/* TopOpti_SetPassiveElements (1511-1584, called at 79 / 272 with
 * specifyPassiveRegions_ = [#LayersBoundary, #LayersLoads, #LayersFixations],
 * 37): adds to the passive set numLayerBoundary layers of cells from the
 * domain boundary inwards, and numLayerLoads / numLayerFixation layers
 * around the loaded (any case) / fixed nodes.  Layer 1 is the boundary cells
 * (a node shared by fewer than 8 solid cells, 1714-1728), the ones touching
 * those nodes; each further layer adds the 3x3x3 neighbourhood within the
 * solid cells (Common_IncludeAdjacentElements, 1587-1615).  Call after
 * t3x_set_passive (which replaces the set) and before t3x_run. */
int t3x_set_passive_layers(t3x_problem* h, int numLayerBoundary, int numLayerLoads, int numLayerFixation);
/* The passive set (passiveElements_): *count voxel ids, ascending, 0-based on
 * the unpadded grid; at most cap are written to voxelIds (which may be NULL
 * to query the count). */
int t3x_get_passive(const t3x_problem* h, int cap, int* voxelIds, int* count);
/* info[0]=numLevels; per level l at 1+8l: resX,resY,resZ,span,numElements,numNodes,a,0;
   a = 1 on the coarsest level when its solve runs as one launch per block
   recurrence (granted while the process's handles fit the device's
   work-group slots), else 0 */
int t3x_info(const t3x_problem* h, int* info);

/* Results of one optimisation (TOP3D_XL_TO 40-223 / TOP3D_XL_PIO 225-440).
 * Any pointer may be NULL.  Arrays of nLoop entries; tHist is nLoop x 6:
 * [assembling, CG, optimisation, filtering, local-volume, total] seconds. */
typedef struct t3x_result {
  double* xPhys;       /* numElements: densityLayout_ (186, 402) */
  double* cHist;       /* compliance per iteration (170, 395) */
  double* volHist;
  double* consHist;
  double* sharpHist;
  int* lssIts;         /* PCG iterations of each solve (121): nLoop x nCases, case fastest */
  double* tHist;
  double* xSnap;       /* nSnap x numElements: xPhys after each of the first nSnap iterations */
  int nSnap;
  double* resHist;     /* final relative PCG residual of each elasticity solve (618), as lssIts */
  double cSolid;       /* out: compliance of the fully solid domain (104) */
  int loops;           /* out: iterations run */
  int warnings;        /* out: PCG max-iteration warnings */
} t3x_result;

/* mode 0 = 'GLOBAL' (TOP3D_XL_TO), 1 = 'LOCAL' (TOP3D_XL_PIO; V0 is the local
 * volume bound and rHat the influence radius, 12-18). */
int t3x_run(t3x_problem* h, int mode, double V0, int nLoop, double rMin, double rHat, t3x_result* out);

/* One design iteration's record (the fprintf of 178-182 / 387-392). */
typedef struct t3x_iter {
  int loop;            /* iteration number (113 / 313) */
  int pcgIts;          /* lssIts entry (121) */
  int filterIts;       /* PDE-filter CG iterations summed over the iteration */
  double c;            /* cDesign (127) */
  double vol;          /* V (128) */
  double cons;         /* fval (145 / 360) */
  double sharp;        /* sharpness (166 / 382) */
  double change;       /* change (154 / 368) */
  double res;          /* final relative PCG residual (618) */
  double t[6];         /* tHist row */
} t3x_iter;

/* The optimisation loop one design iteration at a time: t3x_run ==
 * t3x_opt_begin; t3x_opt_step until *finished; t3x_opt_result.
 * opt_begin does the set-up and the fully-solid solve (94-106 / 294-306). */
int t3x_opt_begin(t3x_problem* h, int mode, double V0, int nLoop, double rMin, double rHat);
int t3x_opt_step(t3x_problem* h, t3x_iter* it, int* finished);
int t3x_opt_result(t3x_problem* h, t3x_result* out);

/* HIP-event time of a kernel group over all its launches since the last reset
 * (params.timing must be 1): 0 = fine-level K*u inside PCG, 1 = V-cycle,
 * 2 = coarsest solve, 3 = assembly (Galerkin + diagonals + Cholesky).
 * No counterpart in the reference: the measurement hook of the benchmark. */
int t3x_kernel_time(t3x_problem* h, int group, long long* count, double* ms);
int t3x_kernel_time_reset(t3x_problem* h);

/* Device bytes a problem of this domain allocates over a whole run (an upper
 * bound: coarsest band at the full-box bandwidth); voxels may be NULL for a
 * full cuboid.  No counterpart in the reference: the admission rule of the
 * serving engine (t3x_serve.h). */
int t3x_footprint(const uint8_t* voxels, int nely, int nelx, int nelz, const t3x_params* params, int mode,
                  long long* bytes);
/* Device bytes the handle holds now.  No counterpart in the reference. */
int t3x_memory(const t3x_problem* h, long long* bytes);
/* Times reps device-to-device copies of `bytes` bytes by this library's
 * streaming-copy kernel (HIP events); *msPerCopy per copy.  No counterpart in
 * the reference: the practical-bandwidth and byte-counter calibration
 * reference of the benchmark. */
int t3x_stream_copy(long long bytes, int reps, double* msPerCopy);

/* ---- step-level entry points (compact host buffers) ---------------------- */
/* meshHierarchy_(1).eleModulus = E (116) */
int t3x_set_modulus(t3x_problem* h, const double* E);
/* Solving_AssembleFEAstencil (919-1066) */
int t3x_assemble(t3x_problem* h);
/* Solving_KbyU_MatrixFree, finest level (641-726); always FP64 */
int t3x_apply_K(t3x_problem* h, const double* u, double* y);
/* Solving_Vcycle ('Adapted', 747-781); with params.precision = 1 the FP32
 * V-cycle (FP64 coarsest Cholesky) */
int t3x_vcycle(t3x_problem* h, const double* r, double* x);
/* Solving_PCG(@Solving_KbyU_MatrixFree, @Solving_Vcycle, ...) (600-639); x holds x0 on entry.
 * With params.precision = 1 the preconditioner is the FP32 V-cycle. */
int t3x_pcg(t3x_problem* h, const double* b, double* x, double tol, int maxIT, int* its, double* resnorm);
/* TopOpti_ComputeUnitCompliance (579-598) */
int t3x_compliance(t3x_problem* h, const double* U, double* ce);
/* The load vector F_ (1315-1321; which = 0) or the displacement U_ of the
 * last elasticity solve (121; which = 1; t3x_pcg's solution after a
 * t3x_pcg call) of load case loadCase, compact nodal order (3 x numNodes,
 * interleaved per node). */
int t3x_get_nodal(t3x_problem* h, int which, int loadCase, double* out);
/* TopOpti_SetupPDEfilter_matrixFree (816-868); slot 0 = rMin filter, 1 = rHat filter */
int t3x_filter_setup(t3x_problem* h, int slot, double radius);
/* TopOpti_ConductPDEFiltering_matrixFree (870-902) */
int t3x_filter(t3x_problem* h, int slot, const double* src, double* dst, int* its);
/* OC bisection (145-155): x updated in place, returns change and bisection count */
int t3x_oc(t3x_problem* h, double* x, const double* xPhys, const double* dc, double V0, double* change,
           int* nBisect);
/* One MMAseq update for m = 1 (1804-2049) on the n cells of the handle:
 * xval, xold1, xold2, low, upp updated in place. */
int t3x_mma(t3x_problem* h, int iter, double* xval, const double* xmin, const double* xmax, double* xold1,
            double* xold2, const double* dfdx, double gx, const double* dgdx, double* low, double* upp,
            int* nNewton);

/* ---- allocation-free entry point ------------------------------------------ */
/* No counterpart in the reference: t3x_oc for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same OC bisection (145-155) on caller-owned
 * DEVICE arrays of nCells compact design cells (x updated in place; xPhys,
 * dc read; passive: NULL or one byte per cell, nonzero = held at 1), run on
 * `stream` (a hipStream_t; NULL = the null stream) and waited for; *change
 * and *nBisect are host outputs.  It calls no hipMalloc/hipFree: its
 * reduction partials and bisection state live in a fixed device array of
 * T3X_OC_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  Calls
 * that overlap on one GPU share that array and race.  nCells beyond
 * T3X_OC_NOALLOC_MAX_CELLS (the reference's int32 cell numbering) is refused
 * with T3X_ERR_INPUT; err (errLen bytes, may be NULL) receives the message. */
#define T3X_OC_NOALLOC_MAX_CELLS 2147483647LL
#define T3X_OC_NOALLOC_SCRATCH_BYTES 262272
int t3x_oc_noalloc(long long nCells, double* x, const double* xPhys, const double* dc, const uint8_t* passive,
                   double V0, void* stream, double* change, int* nBisect, char* err, int errLen);

#ifdef __cplusplus
}
#endif

#endif /* T3X_H_ */
