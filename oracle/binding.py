"""ctypes binding to liboracle.so (test infrastructure; see package docstring)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build_oracle(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (g++/OpenMP)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build_oracle()
    L = C.CDLL(_LIB)
    vp = C.c_void_p
    L.orc_create.restype = vp
    L.orc_create.argtypes = [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                             C.c_int, C.c_void_p, C.c_void_p]
    L.orc_create_multi.restype = vp
    L.orc_create_multi.argtypes = [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_destroy.argtypes = [vp]
    L.orc_last_error.restype = C.c_char_p
    L.orc_last_error.argtypes = [vp]
    L.orc_ok.argtypes = [vp]
    L.orc_warnings.argtypes = [vp]
    L.orc_info.argtypes = [vp, _ip]
    L.orc_get.restype = C.c_longlong
    L.orc_get.argtypes = [vp, C.c_char_p, C.c_int, C.c_void_p, C.c_longlong]
    L.orc_set_modulus.argtypes = [vp, _dp]
    L.orc_assemble.argtypes = [vp]
    L.orc_apply_K.argtypes = [vp, _dp, _dp]
    L.orc_vcycle.argtypes = [vp, _dp, _dp]
    L.orc_coarse_solve.argtypes = [vp, _dp, _dp]
    L.orc_restrict.argtypes = [vp, C.c_int, _dp, _dp]
    L.orc_interpolate.argtypes = [vp, C.c_int, _dp, _dp]
    L.orc_pcg.argtypes = [vp, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.orc_compliance.argtypes = [vp, _dp, _dp]
    L.orc_filter_setup.argtypes = [vp, C.c_int, C.c_double]
    L.orc_filter_apply.argtypes = [vp, C.c_int, _dp, _dp]
    L.orc_filter.argtypes = [vp, C.c_int, _dp, _dp]
    L.orc_oc.restype = C.c_double
    L.orc_oc.argtypes = [vp, _dp, _dp, _dp, C.c_double, C.POINTER(C.c_int)]
    L.orc_mma.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
    L.orc_mma_ex.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
    L.orc_heaviside.argtypes = [C.c_int, C.c_double, C.c_double, _dp, C.c_void_p, C.c_void_p]
    L.orc_pnorm.restype = C.c_double
    L.orc_pnorm.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_compliance_sens.argtypes = [vp, _dp, _dp, C.c_double, _dp]
    L.orc_local_sens.restype = C.c_double
    L.orc_local_sens.argtypes = [vp, C.c_double, _dp, _dp, C.c_double, _dp, _dp, C.POINTER(C.c_int)]
    L.orc_run.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                          _dp, _dp, _dp, _dp, _dp, _ip, _dp, C.c_void_p, C.c_int,
                          C.POINTER(C.c_double), C.POINTER(C.c_int), _ip, _dp]
    L.orc_num_threads.restype = C.c_int
    L.orc_set_order_variant.argtypes = [vp, C.c_int]
    L.orc_set_passive.argtypes = [vp, C.c_int, _ip]
    L.orc_set_passive_layers.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.orc_set_vcycle.argtypes = [vp, C.c_int]
    L.orc_set_omega.argtypes = [vp, C.c_double]
    L.orc_set_tol.argtypes = [vp, C.c_double, C.c_int]
    L.orc_apply_K_level.argtypes = [vp, C.c_int, _dp, _dp]
    _lib = L
    return L


def cuboid_bc_oracle(nely: int, nelx: int, nelz: int):
    """Marker: use the reference's built-in cuboid BCs (TOP3D_XL.m:536-559)."""
    return None


class Oracle:
    """One oracle model (TOP3D_XL.m state after CreateVoxelFEAmodel + setup)."""

    def __init__(self, voxels: np.ndarray, fix=None, load=None, extra_loads=()):
        """extra_loads: further load cases (nodes, vec) after `load`
        (loadingCond_{2..}, 483-493)."""
        L = _load()
        vox = np.asfortranarray(voxels.astype(np.uint8))
        nely, nelx, nelz = vox.shape
        flat = np.ascontiguousarray(vox.reshape(-1, order="F"))
        self._keep = [flat]
        if fix is None:
            h = L.orc_create(flat, nely, nelx, nelz, -1, None, None, 0, None, None)
        else:
            fn, fs = fix
            cases = [load] + list(extra_loads)
            fn = np.ascontiguousarray(fn, dtype=np.int32)
            fs = np.ascontiguousarray(fs, dtype=np.uint8).reshape(-1)
            npc = np.array([len(c[0]) for c in cases], dtype=np.int32)
            ln = np.ascontiguousarray(np.concatenate([np.asarray(c[0]).reshape(-1) for c in cases]), dtype=np.int32)
            lv = np.ascontiguousarray(np.concatenate([np.asarray(c[1], dtype=np.float64).reshape(-1)
                                                      for c in cases]), dtype=np.float64)
            self._keep += [fn, fs, npc, ln, lv]
            h = L.orc_create_multi(flat, nely, nelx, nelz, len(fn), fn.ctypes.data, fs.ctypes.data,
                                   len(cases), npc.ctypes.data, ln.ctypes.data, lv.ctypes.data)
        self.h = h
        self.L = L
        if not L.orc_ok(h):
            msg = L.orc_last_error(h).decode()
            L.orc_destroy(h)
            self.h = None
            raise ValueError(msg)
        info = np.zeros(1 + 8 * 16, dtype=np.int32)
        L.orc_info(h, info)
        self.num_levels = int(info[0])
        self.levels = []
        for l in range(self.num_levels):
            o = info[1 + 8 * l: 9 + 8 * l]
            self.levels.append(dict(resX=int(o[0]), resY=int(o[1]), resZ=int(o[2]), span=int(o[3]),
                                    numElements=int(o[4]), numNodes=int(o[5]),
                                    intermediateNumNodes=int(o[6])))
        self.nE = self.levels[0]["numElements"]
        self.nN = self.levels[0]["numNodes"]
        self.num_cases = 1 + (len(extra_loads) if fix is not None else 0)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    # ------------------------------------------------------------- accessors
    _types = {"transferMatCoeffi": np.float64, "diagK": np.float64, "Ks": np.float64, "RI": np.float64,
              "Ke": np.float64, "F": np.float64, "objWeight": np.float64, "KesFree": np.float64,
              "KesFixed": np.float64,
              "filterKernel0": np.float64, "filterDiag0": np.float64, "fixState": np.uint8,
              "isPassive": np.uint8}

    def get(self, name: str, level: int = 0) -> np.ndarray:
        n = self.L.orc_get(self.h, name.encode(), level, None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(int(n), dtype=self._types.get(name, np.int32))
        self.L.orc_get(self.h, name.encode(), level, out.ctypes.data, int(n))
        return out

    # --------------------------------------------------------------- steps
    def set_modulus(self, E):
        self.L.orc_set_modulus(self.h, np.ascontiguousarray(E, dtype=np.float64))

    def assemble(self):
        if self.L.orc_assemble(self.h) != 0:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())

    def apply_K(self, u):
        y = np.zeros(3 * self.nN)
        self.L.orc_apply_K(self.h, np.ascontiguousarray(u, dtype=np.float64), y)
        return y

    def set_vcycle(self, kind):
        """typeVcycle_ (34): 'Adapted' or 'Standard'."""
        if self.L.orc_set_vcycle(self.h, {"Adapted": 0, "Standard": 1}[kind]) != 0:
            raise ValueError(kind)

    def set_tol(self, tol, max_it=800):
        """tol_ (30) and maxIT_ (31) of the elasticity solves."""
        self.L.orc_set_tol(self.h, float(tol), int(max_it))

    def set_omega(self, w):
        """weightFactorJacobi_ (32)."""
        self.L.orc_set_omega(self.h, float(w))

    def apply_K_level(self, level, u):
        y = np.zeros(3 * self.levels[level]["numNodes"])
        if self.L.orc_apply_K_level(self.h, level, np.ascontiguousarray(u, dtype=np.float64), y) != 0:
            raise ValueError(level)
        return y

    def vcycle(self, r):
        x = np.zeros(3 * self.nN)
        self.L.orc_vcycle(self.h, np.ascontiguousarray(r, dtype=np.float64), x)
        return x

    def coarse_solve(self, r):
        x = np.zeros_like(r, dtype=np.float64)
        self.L.orc_coarse_solve(self.h, np.ascontiguousarray(r, dtype=np.float64), x)
        return x

    def restrict(self, level, rF):
        rC = np.zeros(3 * self.levels[level]["numNodes"])
        self.L.orc_restrict(self.h, level, np.ascontiguousarray(rF, dtype=np.float64), rC)
        return rC

    def interpolate(self, level, xC):
        xF = np.zeros(3 * self.levels[level - 1]["numNodes"])
        self.L.orc_interpolate(self.h, level, np.ascontiguousarray(xC, dtype=np.float64), xF)
        return xF

    def pcg(self, b, x0=None, tol=1e-3, maxit=800):
        x = np.zeros(3 * self.nN) if x0 is None else np.array(x0, dtype=np.float64)
        res = C.c_double(0.0)
        its = self.L.orc_pcg(self.h, np.ascontiguousarray(b, dtype=np.float64), x, tol, maxit, C.byref(res))
        return x, its, res.value

    def compliance(self, U):
        ce = np.zeros(self.nE)
        self.L.orc_compliance(self.h, np.ascontiguousarray(U, dtype=np.float64), ce)
        return ce

    def filter_setup(self, slot, r):
        self.L.orc_filter_setup(self.h, slot, float(r))

    def filter_apply(self, slot, u):
        y = np.zeros(self.nN)
        self.L.orc_filter_apply(self.h, slot, np.ascontiguousarray(u, dtype=np.float64), y)
        return y

    def filter(self, slot, src):
        dst = np.zeros(self.nE)
        its = self.L.orc_filter(self.h, slot, np.ascontiguousarray(src, dtype=np.float64), dst)
        return dst, its

    def oc(self, x, xPhys, dc, V0):
        xx = np.array(x, dtype=np.float64)
        nb = C.c_int(0)
        ch = self.L.orc_oc(self.h, xx, np.ascontiguousarray(xPhys, dtype=np.float64),
                           np.ascontiguousarray(dc, dtype=np.float64), float(V0), C.byref(nb))
        return xx, ch, nb.value

    @staticmethod
    def mma(iter_, xval, xmin, xmax, xold1, xold2, dfdx, gx, dgdx, low, upp):
        L = _load()
        n = len(xval)
        arrs = [np.array(a, dtype=np.float64) for a in (xval, xmin, xmax, xold1, xold2, dfdx, dgdx, low, upp)]
        xv, xmn, xmx, xo1, xo2, df, dg, lo, up = arrs
        nn = L.orc_mma(n, iter_, xv, xmn, xmx, xo1, xo2, df, float(gx), dg, lo, up)
        return dict(x=xv, xold1=xo1, xold2=xo2, low=lo, upp=up, newton=nn)

    @staticmethod
    def mma_ex(iter_, xval, xmin, xmax, xold1, xold2, dfdx, gx, dgdx, low, upp):
        """One MMAseq update that also returns GenSub's subproblem (alpha,
        beta, p0, q0, pij, qij, b) and SolveDIP's final dual (lam, mu)."""
        L = _load()
        n = len(xval)
        arrs = [np.array(a, dtype=np.float64) for a in (xval, xmin, xmax, xold1, xold2, dfdx, dgdx, low, upp)]
        xv, xmn, xmx, xo1, xo2, df, dg, lo, up = arrs
        sub = np.zeros(6 * n + 3)
        nn = L.orc_mma_ex(n, iter_, xv, xmn, xmx, xo1, xo2, df, float(gx), dg, lo, up, sub)
        parts = sub[:6 * n].reshape(6, n)
        return dict(x=xv, xold1=xo1, xold2=xo2, low=lo, upp=up, newton=nn, alpha=parts[0], beta=parts[1],
                    p0=parts[2], q0=parts[3], pij=parts[4], qij=parts[5], b=sub[6 * n], lam=sub[6 * n + 1],
                    mu=sub[6 * n + 2])

    @staticmethod
    def heaviside(beta, eta, x_tilde):
        """Heaviside projection (380) and its derivative (340)."""
        L = _load()
        xt = np.ascontiguousarray(x_tilde, dtype=np.float64)
        xp, dx = np.zeros_like(xt), np.zeros_like(xt)
        L.orc_heaviside(len(xt), float(beta), float(eta), xt, xp.ctypes.data, dx.ctypes.data)
        return xp, dx

    @staticmethod
    def pnorm(xhat, vol_max):
        """p-norm local-volume constraint value (360) and gradient (335)."""
        L = _load()
        xh = np.ascontiguousarray(xhat, dtype=np.float64)
        vm = np.ascontiguousarray(vol_max, dtype=np.float64)
        g = np.zeros_like(xh)
        return L.orc_pnorm(len(xh), xh, vm, g), g

    def compliance_sens(self, xPhys, ce, c_solid):
        """dc = -DeSIMP(xPhys) .* ce / cSolid (129 / 330)."""
        dc = np.zeros(self.nE)
        self.L.orc_compliance_sens(self.h, np.ascontiguousarray(xPhys, dtype=np.float64),
                                   np.ascontiguousarray(ce, dtype=np.float64), float(c_solid), dc)
        return dc

    def local_sens(self, beta, xPhys, xTilde, ve0, dc):
        """The LOCAL sensitivity chain (334-350): returns (fval, dc filtered,
        dfdx, filter CG iterations).  Filter slots 0 (rMin) and 1 (rHat)
        must be set up."""
        d = np.array(dc, dtype=np.float64)
        dfdx = np.zeros(self.nE)
        its = C.c_int(0)
        f = self.L.orc_local_sens(self.h, float(beta), np.ascontiguousarray(xPhys, dtype=np.float64),
                                  np.ascontiguousarray(xTilde, dtype=np.float64), float(ve0), d, dfdx,
                                  C.byref(its))
        return f, d, dfdx, its.value

    def run(self, mode, V0, nLoop, rMin, rHat=0.0, nSnap=0):
        nE = self.nE
        xPhys = np.zeros(nE)
        cH, vH, conH, sH = (np.zeros(max(nLoop, 1)) for _ in range(4))
        nc = self.num_cases
        its = np.zeros(max(nLoop, 1) * nc, dtype=np.int32)
        fits = np.zeros(max(nLoop, 1), dtype=np.int32)
        resH = np.zeros(max(nLoop, 1) * nc)
        tH = np.zeros(max(nLoop, 1) * 6)
        snap = np.zeros(max(nSnap, 1) * nE)
        cS = C.c_double(0.0)
        loops = C.c_int(0)
        m = {"GLOBAL": 0, "LOCAL": 1}[mode] if isinstance(mode, str) else int(mode)
        rc = self.L.orc_run(self.h, m, float(V0), int(nLoop), float(rMin), float(rHat), xPhys, cH, vH, conH, sH,
                            its, tH, snap.ctypes.data, int(nSnap), C.byref(cS), C.byref(loops), fits, resH)
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())
        k = loops.value
        return dict(xPhys=xPhys, cHist=cH[:k], volHist=vH[:k], consHist=conH[:k], sharpHist=sH[:k],
                    lssIts=its[:k * nc], filterIts=fits[:k], resHist=resH[:k * nc], tHist=tH[:6 * k].reshape(k, 6),
                    xSnap=snap[:min(nSnap, k) * nE].reshape(-1, nE), cSolid=cS.value, loops=k,
                    warnings=self.L.orc_warnings(self.h))

    def set_order_variant(self, k):
        """k-th best coarsest-level ordering: same factor, different FP64 rounding."""
        self.L.orc_set_order_variant(self.h, int(k))

    def set_passive(self, voxel_ids):
        """passiveElements_: 0-based voxel ids of the unpadded grid."""
        ids = np.ascontiguousarray(voxel_ids, dtype=np.int32)
        self.L.orc_set_passive(self.h, len(ids), ids)

    def set_passive_layers(self, n_boundary, n_loads, n_fixations):
        """specifyPassiveRegions_ (37) through TopOpti_SetPassiveElements."""
        self.L.orc_set_passive_layers(self.h, int(n_boundary), int(n_loads), int(n_fixations))

    @staticmethod
    def num_threads():
        return _load().orc_num_threads()
