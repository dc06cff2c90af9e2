// MEX gateway between the MATLAB drop-in TOP3D_XL.m and libt3x's C ABI.
//
// Build where MATLAB is installed (MATLAB's mex.h is not part of this tree):
//   mex -R2018a t3x_mex.cpp -I../include -L../top3d_xl_amd -lt3x
//
// Commands (handles are uint64 scalars owned by this gateway and released at
// 'destroy' or when MATLAB clears the MEX file):
//   h = t3x_mex('create', voxels)                      built-in cuboid BCs (536-559)
//   h = t3x_mex('create', voxels, fixNodes, fixState, loadNodes, loadVec)
//                                                      explicit BCs, 1-based node ids of the
//                                                      unpadded (nely+1)x(nelx+1)x(nelz+1) grid
//   h = t3x_mex(..., params)                           optional struct with t3x_params fields
//   h = t3x_mex('create', voxels, fixNodes, fixState, loadNodes, loadVec, params, extraLoads)
//                                                      further load cases (loadingCond_{2..}, 483-493):
//                                                      a cell array of n x 4 [node Fx Fy Fz] matrices
//   t3x_mex('passive', h, voxelIds)                   passive elements, 1-based voxel ids
//   out = t3x_mex('run', h, consType, V0, nLoop, rMin [, rHat])
//   res = t3x_mex('resolution', h)                    padded cell grid [resY resX resZ]
//   t3x_mex('destroy', h)
//
// Status codes become the reference's errors and warnings:
//   T3X_ERR_INPUT -> error('Inappropriate Input Model!' / 'Wrong Input!') (16, 533)
//   T3X_WARN_MAXIT -> warning('Exceed the maximum iterate numbers') (634-637)
#include <cstdint>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "mex.h"
#include "t3x.h"

namespace {

std::set<t3x_problem*> g_handles;

void release_all() {
  for (t3x_problem* h : g_handles) t3x_destroy(h);
  g_handles.clear();
}

t3x_problem* handle_arg(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("t3x:handle", "bad handle");
  t3x_problem* h = reinterpret_cast<t3x_problem*>(static_cast<uintptr_t>(*static_cast<uint64_t*>(mxGetData(a))));
  if (!g_handles.count(h)) mexErrMsgIdAndTxt("t3x:handle", "unknown or destroyed handle");
  return h;
}

void fail(int rc, t3x_problem* h) {
  const char* msg = h ? t3x_last_error(h) : "t3x error";
  switch (rc) {
    case T3X_ERR_INPUT: mexErrMsgIdAndTxt("t3x:input", "%s", msg); break;
    case T3X_ERR_NO_LOADS: mexErrMsgIdAndTxt("t3x:noLoads", "No Loads!"); break;
    case T3X_ERR_NO_FIX: mexErrMsgIdAndTxt("t3x:noFix", "No Fixations!"); break;
    case T3X_ERR_NOT_PD: mexErrMsgIdAndTxt("t3x:chol", "%s", msg); break;
    case T3X_ERR_NO_DEVICE: mexErrMsgIdAndTxt("t3x:device", "%s", msg); break;
    default: mexErrMsgIdAndTxt("t3x:hip", "%s", msg); break;
  }
}

void read_params(const mxArray* s, t3x_params* p) {
  t3x_default_params(p);
  if (!s || !mxIsStruct(s)) return;
  auto num = [&](const char* f, double* dst) {
    const mxArray* v = mxGetField(s, 0, f);
    if (v) *dst = mxGetScalar(v);
  };
  auto inum = [&](const char* f, int* dst) {
    const mxArray* v = mxGetField(s, 0, f);
    if (v) *dst = static_cast<int>(mxGetScalar(v));
  };
  num("E0", &p->E0); num("nu", &p->nu); num("Emin", &p->Emin); num("penal", &p->penal); num("tol", &p->tol);
  inum("maxIT", &p->maxIT); num("omega", &p->omega); inum("coarsestControl", &p->coarsestControl);
  inum("precision", &p->precision); inum("checkEvery", &p->checkEvery); inum("timing", &p->timing);
  inum("vcycle", &p->vcycle); inum("graphs", &p->graphs);
}

mxArray* dvec(const double* v, int n) {
  mxArray* a = mxCreateDoubleMatrix(n, 1, mxREAL);
  if (n) std::memcpy(mxGetDoubles(a), v, sizeof(double) * n);
  return a;
}

void cmd_create(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 2) mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");
  const mxArray* vox = prhs[1];
  if (mxGetNumberOfDimensions(vox) != 3) mexErrMsgIdAndTxt("t3x:input", "Inappropriate Input Model!");
  const mwSize* d = mxGetDimensions(vox);
  const int nely = (int)d[0], nelx = (int)d[1], nelz = (int)d[2];
  const size_t n = (size_t)nely * nelx * nelz;
  std::vector<uint8_t> v(n);
  if (mxIsLogical(vox)) {
    const mxLogical* L = mxGetLogicals(vox);
    for (size_t i = 0; i < n; ++i) v[i] = L[i] ? 1 : 0;
  } else {
    const double* D = mxGetDoubles(vox);
    if (!D) mexErrMsgIdAndTxt("t3x:input", "Inappropriate Input Model!");
    for (size_t i = 0; i < n; ++i) v[i] = D[i] != 0.0 ? 1 : 0;
  }
  t3x_params P;
  const bool explicitBC = nrhs >= 6;
  read_params(nrhs == 3 ? prhs[2] : (nrhs >= 7 ? prhs[6] : nullptr), &P);
  t3x_problem* h = nullptr;
  int rc;
  if (explicitBC) {
    // MATLAB's 1-based node ids -> the ABI's 0-based ids.  fixNodes / loadNodes
    // are double vectors, fixState an nf x 3 and loadVec an nl x 3 double
    // matrix: checked before any of them is read.
    for (int k = 2; k <= 5; ++k)
      if (!mxIsDouble(prhs[k]) || mxIsComplex(prhs[k]) || mxIsSparse(prhs[k]))
        mexErrMsgIdAndTxt("t3x:input", "boundary conditions must be real full double arrays");
    const size_t nf = mxGetNumberOfElements(prhs[2]), nl = mxGetNumberOfElements(prhs[4]);
    if (mxGetM(prhs[3]) != nf || mxGetN(prhs[3]) != 3)
      mexErrMsgIdAndTxt("t3x:input", "fixState must be numel(fixNodes) x 3");
    if (mxGetM(prhs[5]) != nl || mxGetN(prhs[5]) != 3)
      mexErrMsgIdAndTxt("t3x:input", "loadVec must be numel(loadNodes) x 3");
    std::vector<int> fn(nf), ln(nl);
    std::vector<uint8_t> fs(3 * nf);
    std::vector<double> lv(3 * nl);
    const double* f = mxGetDoubles(prhs[2]);
    const double* fsd = mxGetDoubles(prhs[3]);
    const double* l = mxGetDoubles(prhs[4]);
    const double* lvd = mxGetDoubles(prhs[5]);
    for (size_t i = 0; i < nf; ++i) {
      fn[i] = (int)f[i] - 1;
      for (int c = 0; c < 3; ++c) fs[3 * i + c] = fsd[i + nf * c] != 0.0;  // nf x 3, column-major
    }
    for (size_t i = 0; i < nl; ++i) {
      ln[i] = (int)l[i] - 1;
      for (int c = 0; c < 3; ++c) lv[3 * i + c] = lvd[i + nl * c];       // nl x 3
    }
    std::vector<int> npc(1, (int)nl);
    if (nrhs >= 8 && mxIsCell(prhs[7])) {
      for (size_t k = 0; k < mxGetNumberOfElements(prhs[7]); ++k) {
        const mxArray* c = mxGetCell(prhs[7], k);
        if (!c || !mxIsDouble(c) || mxIsComplex(c) || mxIsSparse(c) || mxGetN(c) != 4)
          mexErrMsgIdAndTxt("t3x:input", "Incompatible Mesh Data Format!");
        const size_t m = mxGetM(c);
        const double* a = mxGetDoubles(c);  // m x 4, column-major
        for (size_t i = 0; i < m; ++i) {
          ln.push_back((int)a[i] - 1);
          for (int cc = 0; cc < 3; ++cc) lv.push_back(a[i + m * (cc + 1)]);
        }
        npc.push_back((int)m);
      }
    }
    rc = t3x_create_multi(v.data(), nely, nelx, nelz, (int)nf, fn.data(), fs.data(), (int)npc.size(), npc.data(),
                          ln.data(), lv.data(), &P, &h);
  } else {
    rc = t3x_create(v.data(), nely, nelx, nelz, -1, nullptr, nullptr, 0, nullptr, nullptr, &P, &h);
  }
  if (rc != T3X_OK) {
    std::string msg = h ? t3x_last_error(h) : "t3x_create failed";
    if (h) t3x_destroy(h);
    if (rc == T3X_ERR_INPUT) mexErrMsgIdAndTxt("t3x:input", "%s", msg.c_str());
    mexErrMsgIdAndTxt("t3x:create", "%s", msg.c_str());
  }
  g_handles.insert(h);
  plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *static_cast<uint64_t*>(mxGetData(plhs[0])) = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(h));
  (void)nlhs;
}

void cmd_run(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 6) mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");
  t3x_problem* h = handle_arg(prhs[1]);
  char cons[16] = {0};
  mxGetString(prhs[2], cons, sizeof(cons));
  int mode;
  if (!std::strcmp(cons, "GLOBAL")) {
    mode = 0;
  } else if (!std::strcmp(cons, "LOCAL")) {
    if (nrhs != 7) mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");  // 16
    mode = 1;
  } else {
    mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");
    return;
  }
  const double V0 = mxGetScalar(prhs[3]);
  const int nLoop = (int)mxGetScalar(prhs[4]);
  const double rMin = mxGetScalar(prhs[5]);
  const double rHat = mode ? mxGetScalar(prhs[6]) : 0.0;
  int info[1 + 8 * 16];
  t3x_info(h, info);
  const int nE = info[1 + 4];
  const int nc = t3x_num_load_cases(h);
  std::vector<double> xPhys(nE), cH(nLoop), vH(nLoop), conH(nLoop), sH(nLoop), tH(6 * (size_t)nLoop),
      rH((size_t)nLoop * nc);
  std::vector<int> its((size_t)nLoop * nc);  // one entry per solve (lssIts(end+1), 121)
  t3x_result R{};
  R.xPhys = xPhys.data(); R.cHist = cH.data(); R.volHist = vH.data(); R.consHist = conH.data();
  R.sharpHist = sH.data(); R.lssIts = its.data(); R.tHist = tH.data(); R.resHist = rH.data();
  const int rc = t3x_run(h, mode, V0, nLoop, rMin, rHat, &R);
  if (rc == T3X_WARN_MAXIT) mexWarnMsgIdAndTxt("t3x:maxit", "Exceed the maximum iterate numbers");
  else if (rc != T3X_OK) fail(rc, h);
  const int k = R.loops;
  const char* names[] = {"xPhys", "cHist", "volHist", "consHist", "sharpHist", "lssIts", "tHist", "cSolid"};
  mxArray* out = mxCreateStructMatrix(1, 1, 8, names);
  mxSetField(out, 0, "xPhys", dvec(xPhys.data(), nE));
  mxSetField(out, 0, "cHist", dvec(cH.data(), k));
  mxSetField(out, 0, "volHist", dvec(vH.data(), k));
  mxSetField(out, 0, "consHist", dvec(conH.data(), k));
  mxSetField(out, 0, "sharpHist", dvec(sH.data(), k));
  std::vector<double> itsd(its.begin(), its.begin() + (size_t)k * nc);
  mxSetField(out, 0, "lssIts", dvec(itsd.data(), k * nc));
  mxArray* T = mxCreateDoubleMatrix(k, 6, mxREAL);  // row-major nLoop x 6 -> MATLAB column-major
  for (int i = 0; i < k; ++i)
    for (int c = 0; c < 6; ++c) mxGetDoubles(T)[i + (size_t)k * c] = tH[6 * (size_t)i + c];
  mxSetField(out, 0, "tHist", T);
  mxSetField(out, 0, "cSolid", mxCreateDoubleScalar(R.cSolid));
  plhs[0] = out;
  (void)nlhs;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static bool registered = false;
  if (!registered) {
    mexAtExit(release_all);
    registered = true;
  }
  if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("t3x:input", "first argument: command");
  char cmd[16] = {0};
  mxGetString(prhs[0], cmd, sizeof(cmd));
  if (!std::strcmp(cmd, "create")) {
    cmd_create(nlhs, plhs, nrhs, prhs);
  } else if (!std::strcmp(cmd, "run")) {
    cmd_run(nlhs, plhs, nrhs, prhs);
  } else if (!std::strcmp(cmd, "passive")) {
    if (nrhs < 3) mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");
    t3x_problem* h = handle_arg(prhs[1]);
    const size_t n = mxGetNumberOfElements(prhs[2]);
    const double* v = mxGetDoubles(prhs[2]);
    std::vector<int> ids(n);
    for (size_t i = 0; i < n; ++i) ids[i] = (int)v[i] - 1;  // 1-based -> 0-based
    const int rc = t3x_set_passive(h, (int)n, ids.data());
    if (rc != T3X_OK) fail(rc, h);
  } else if (!std::strcmp(cmd, "resolution")) {
    // [resY resX resZ]: the padded cell grid (1651-1671), for the design export
    if (nrhs < 2) mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");
    t3x_problem* h = handle_arg(prhs[1]);
    int info[1 + 8 * 16];
    t3x_info(h, info);
    const double r[3] = {(double)info[2], (double)info[1], (double)info[3]};
    mxArray* a = mxCreateDoubleMatrix(1, 3, mxREAL);
    std::memcpy(mxGetDoubles(a), r, sizeof(r));
    plhs[0] = a;
  } else if (!std::strcmp(cmd, "destroy")) {
    if (nrhs < 2) mexErrMsgIdAndTxt("t3x:input", "Wrong Input!");
    t3x_problem* h = handle_arg(prhs[1]);
    g_handles.erase(h);
    t3x_destroy(h);
  } else {
    mexErrMsgIdAndTxt("t3x:input", "unknown command '%s'", cmd);
  }
}
