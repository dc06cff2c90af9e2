"""ctypes binding to libt3x.so and the Python mirror of the reference entry point.

``TOP3D_XL(inputModel, consType, V0, nLoop, rMin, *varargin)`` keeps the
signature of TOP3D_XL.m:4-19: 'GLOBAL' runs the compliance minimisation loop
(TOP3D_XL_TO), 'LOCAL' the porous-infill loop (TOP3D_XL_PIO) and needs a 6th
argument rHat.  Instead of writing files it returns the densities and the
histories the reference writes to ./out/ (185-204, 402-421).

All computation runs in the HIP library; there is no CPU path.  If the shared
library is missing or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libt3x.so")

T3X_OK = 0
T3X_WARN_MAXIT = 100
_STATUS = {1: "input", 2: "no loads", 3: "no fixations", 4: "HIP error", 5: "coarsest matrix not PD",
           6: "no GPU device"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("E0", C.c_double), ("nu", C.c_double), ("Emin", C.c_double), ("penal", C.c_double),
                ("tol", C.c_double), ("maxIT", C.c_int), ("omega", C.c_double), ("coarsestControl", C.c_int),
                ("precision", C.c_int), ("checkEvery", C.c_int), ("timing", C.c_int), ("vcycle", C.c_int),
                ("graphs", C.c_int)]


class Iter(C.Structure):
    _fields_ = [("loop", C.c_int), ("pcgIts", C.c_int), ("filterIts", C.c_int), ("c", C.c_double),
                ("vol", C.c_double), ("cons", C.c_double), ("sharp", C.c_double), ("change", C.c_double),
                ("res", C.c_double), ("t", C.c_double * 6)]


class Result(C.Structure):
    _fields_ = [("xPhys", C.c_void_p), ("cHist", C.c_void_p), ("volHist", C.c_void_p),
                ("consHist", C.c_void_p), ("sharpHist", C.c_void_p), ("lssIts", C.c_void_p),
                ("tHist", C.c_void_p), ("xSnap", C.c_void_p), ("nSnap", C.c_int), ("resHist", C.c_void_p),
                ("cSolid", C.c_double),
                ("loops", C.c_int), ("warnings", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("memBudget", C.c_longlong), ("maxActive", C.c_int), ("params", Params)]


class Request(C.Structure):
    _fields_ = [("tag", C.c_longlong), ("nely", C.c_int), ("nelx", C.c_int), ("nelz", C.c_int),
                ("mode", C.c_int), ("V0", C.c_double), ("nLoop", C.c_int), ("rMin", C.c_double),
                ("rHat", C.c_double), ("loadDir", C.c_double * 3), ("loadMag", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_longlong), ("status", C.c_int), ("loops", C.c_int), ("pcgIts", C.c_int),
                ("nE", C.c_int), ("tSubmit", C.c_double), ("tStart", C.c_double), ("tEnd", C.c_double),
                ("cFinal", C.c_double), ("cSolid", C.c_double), ("footprint", C.c_longlong),
                ("xPhys", C.POINTER(C.c_double)), ("cHist", C.POINTER(C.c_double))]


_lib = None


def load_library():
    """Loads libt3x.so (built by top3d_xl_amd/build.py); raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run top3d_xl_amd/build.py (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.t3x_default_params.argtypes = [C.POINTER(Params)]
    L.t3x_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                             C.c_void_p, C.c_void_p, C.POINTER(Params), C.POINTER(vp)]
    L.t3x_create_multi.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Params), C.POINTER(vp)]
    L.t3x_num_load_cases.argtypes = [vp]
    L.t3x_destroy.argtypes = [vp]
    L.t3x_last_error.restype = C.c_char_p
    L.t3x_last_error.argtypes = [vp]
    L.t3x_info.argtypes = [vp, C.c_void_p]
    L.t3x_run.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(Result)]
    L.t3x_opt_begin.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double]
    L.t3x_opt_step.argtypes = [vp, C.POINTER(Iter), C.POINTER(C.c_int)]
    L.t3x_opt_result.argtypes = [vp, C.POINTER(Result)]
    L.t3x_kernel_time.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
    L.t3x_kernel_time_reset.argtypes = [vp]
    L.t3x_set_modulus.argtypes = [vp, _dp]
    L.t3x_assemble.argtypes = [vp]
    L.t3x_apply_K.argtypes = [vp, _dp, _dp]
    L.t3x_vcycle.argtypes = [vp, _dp, _dp]
    L.t3x_pcg.argtypes = [vp, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.t3x_compliance.argtypes = [vp, _dp, _dp]
    L.t3x_get_nodal.argtypes = [vp, C.c_int, C.c_int, _dp]
    L.t3x_filter_setup.argtypes = [vp, C.c_int, C.c_double]
    L.t3x_filter.argtypes = [vp, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
    L.t3x_oc.argtypes = [vp, _dp, _dp, _dp, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.t3x_mma.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp,
                          C.POINTER(C.c_int)]
    L.t3x_footprint.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(Params), C.c_int,
                                C.POINTER(C.c_longlong)]
    L.t3x_memory.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.t3x_set_passive.argtypes = [vp, C.c_int, C.c_void_p]
    L.t3x_set_passive_layers.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.t3x_get_passive.argtypes = [vp, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
    L.t3x_stream_copy.argtypes = [C.c_longlong, C.c_int, C.POINTER(C.c_double)]
    L.t3x_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
    L.t3x_server_submit.argtypes = [vp, C.POINTER(Request)]
    L.t3x_server_poll.argtypes = [vp, C.POINTER(Completion), C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.t3x_server_release.argtypes = [vp, C.c_longlong]
    L.t3x_server_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_longlong)]
    L.t3x_server_destroy.argtypes = [vp]
    L.t3x_oc_noalloc.argtypes = [C.c_longlong, vp, vp, vp, vp, C.c_double, vp, C.POINTER(C.c_double),
                                 C.POINTER(C.c_int), C.c_char_p, C.c_int]
    _lib = L
    return L


EXPORTED = ["t3x_default_params", "t3x_create", "t3x_create_multi", "t3x_num_load_cases", "t3x_destroy", "t3x_last_error", "t3x_info", "t3x_run",
            "t3x_opt_begin", "t3x_opt_step", "t3x_opt_result", "t3x_kernel_time", "t3x_kernel_time_reset",
            "t3x_set_modulus", "t3x_assemble", "t3x_apply_K", "t3x_vcycle", "t3x_pcg", "t3x_compliance", "t3x_get_nodal",
            "t3x_filter_setup", "t3x_filter", "t3x_oc", "t3x_mma", "t3x_footprint", "t3x_memory", "t3x_stream_copy", "t3x_set_passive",
            "t3x_set_passive_layers", "t3x_get_passive",
            "t3x_server_create", "t3x_server_submit", "t3x_server_poll", "t3x_server_release",
            "t3x_server_stats", "t3x_server_destroy", "t3x_oc_noalloc"]


def default_params(**kw) -> Params:
    p = Params()
    load_library().t3x_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class T3XError(RuntimeError):
    pass


class Problem:
    """One voxel model on the GPU (state after CreateVoxelFEAmodel + FEA set-up)."""

    def __init__(self, voxels, fix=None, load=None, extra_loads=(), **params):
        """fix = (nodes, state n x 3), load = (nodes, vec n x 3) on the unpadded
        node grid (0-based); extra_loads: further load cases (nodes, vec)
        (loadingCond_{2..}, 483-493).  fix None: the built-in cuboid."""
        L = load_library()
        vox = np.asfortranarray(np.asarray(voxels).astype(np.uint8))
        if vox.ndim != 3:
            raise T3XError("Inappropriate Input Model!")
        nely, nelx, nelz = vox.shape
        flat = np.ascontiguousarray(vox.reshape(-1, order="F"))
        self._keep = [flat]
        P = default_params(**params)
        h = C.c_void_p()
        if fix is None:
            rc = L.t3x_create(flat.ctypes.data, nely, nelx, nelz, -1, None, None, 0, None, None, C.byref(P),
                              C.byref(h))
        else:
            fn = np.ascontiguousarray(fix[0], dtype=np.int32)
            fs = np.ascontiguousarray(fix[1], dtype=np.uint8).reshape(-1)
            cases = [load] + list(extra_loads)
            npc = np.array([len(np.asarray(c[0]).reshape(-1)) for c in cases], dtype=np.int32)
            ln = np.ascontiguousarray(np.concatenate([np.asarray(c[0]).reshape(-1) for c in cases]), dtype=np.int32)
            lv = np.ascontiguousarray(np.concatenate([np.asarray(c[1], dtype=np.float64).reshape(-1)
                                                      for c in cases]), dtype=np.float64)
            self._keep += [fn, fs, npc, ln, lv]
            rc = L.t3x_create_multi(flat.ctypes.data, nely, nelx, nelz, len(fn), fn.ctypes.data, fs.ctypes.data,
                                    len(cases), npc.ctypes.data, ln.ctypes.data, lv.ctypes.data, C.byref(P),
                                    C.byref(h))
        self.L = L
        self.h = h
        if rc != T3X_OK:
            msg = L.t3x_last_error(h).decode() if h else _STATUS.get(rc, str(rc))
            if h:
                L.t3x_destroy(h)
            self.h = None
            raise T3XError(f"t3x_create failed ({_STATUS.get(rc, rc)}): {msg}")
        self.num_cases = L.t3x_num_load_cases(h)
        info = np.zeros(1 + 8 * 16, dtype=np.int32)
        L.t3x_info(h, info.ctypes.data)
        self.num_levels = int(info[0])
        self.levels = []
        for l in range(self.num_levels):
            o = info[1 + 8 * l: 9 + 8 * l]
            self.levels.append(dict(resX=int(o[0]), resY=int(o[1]), resZ=int(o[2]), span=int(o[3]),
                                    numElements=int(o[4]), numNodes=int(o[5])))
        self.nE = self.levels[0]["numElements"]
        self.nN = self.levels[0]["numNodes"]
        # the coarsest solve runs as one launch per block recurrence
        self.one_launch_coarse = bool(info[1 + 8 * (self.num_levels - 1) + 6])

    def close(self):
        if getattr(self, "h", None):
            self.L.t3x_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc not in (T3X_OK, T3X_WARN_MAXIT):
            raise T3XError(f"{_STATUS.get(rc, rc)}: {self.L.t3x_last_error(self.h).decode()}")
        return rc

    # ------------------------------------------------------------- steps
    def set_modulus(self, E):
        self._check(self.L.t3x_set_modulus(self.h, np.ascontiguousarray(E, dtype=np.float64)))

    def assemble(self):
        self._check(self.L.t3x_assemble(self.h))

    def apply_K(self, u):
        y = np.zeros(3 * self.nN)
        self._check(self.L.t3x_apply_K(self.h, np.ascontiguousarray(u, dtype=np.float64), y))
        return y

    def vcycle(self, r):
        x = np.zeros(3 * self.nN)
        self._check(self.L.t3x_vcycle(self.h, np.ascontiguousarray(r, dtype=np.float64), x))
        return x

    def pcg(self, b, x0=None, tol=1e-3, maxit=800):
        x = np.zeros(3 * self.nN) if x0 is None else np.array(x0, dtype=np.float64)
        its = C.c_int(0)
        res = C.c_double(0.0)
        self._check(self.L.t3x_pcg(self.h, np.ascontiguousarray(b, dtype=np.float64), x, tol, maxit,
                                   C.byref(its), C.byref(res)))
        return x, its.value, res.value

    def compliance(self, U):
        ce = np.zeros(self.nE)
        self._check(self.L.t3x_compliance(self.h, np.ascontiguousarray(U, dtype=np.float64), ce))
        return ce

    def load_vector(self, case=0):
        """F_ of a load case (t3x_get_nodal), compact interleaved nodal order."""
        out = np.zeros(3 * self.nN)
        self._check(self.L.t3x_get_nodal(self.h, 0, int(case), out))
        return out

    def displacement(self, case=0):
        """U_ of a load case after the last solve (t3x_get_nodal)."""
        out = np.zeros(3 * self.nN)
        self._check(self.L.t3x_get_nodal(self.h, 1, int(case), out))
        return out

    def xphys(self):
        """The current physical densities (t3x_opt_result), compact cell order."""
        x = np.zeros(self.nE)
        R = Result(x.ctypes.data, None, None, None, None, None, None, None, 0, None, 0.0, 0, 0)
        self._check(self.L.t3x_opt_result(self.h, C.byref(R)))
        return x

    def filter_setup(self, slot, r):
        self._check(self.L.t3x_filter_setup(self.h, slot, float(r)))

    def filter(self, slot, src):
        dst = np.zeros(self.nE)
        its = C.c_int(0)
        self._check(self.L.t3x_filter(self.h, slot, np.ascontiguousarray(src, dtype=np.float64), dst,
                                      C.byref(its)))
        return dst, its.value

    def oc(self, x, xPhys, dc, V0):
        xx = np.array(x, dtype=np.float64)
        ch = C.c_double(0.0)
        nb = C.c_int(0)
        self._check(self.L.t3x_oc(self.h, xx, np.ascontiguousarray(xPhys, dtype=np.float64),
                                  np.ascontiguousarray(dc, dtype=np.float64), float(V0), C.byref(ch), C.byref(nb)))
        return xx, ch.value, nb.value

    def mma(self, iter_, xval, xmin, xmax, xold1, xold2, dfdx, gx, dgdx, low, upp):
        arrs = [np.array(a, dtype=np.float64) for a in (xval, xmin, xmax, xold1, xold2, dfdx, dgdx, low, upp)]
        xv, xmn, xmx, xo1, xo2, df, dg, lo, up = arrs
        nn = C.c_int(0)
        self._check(self.L.t3x_mma(self.h, int(iter_), xv, xmn, xmx, xo1, xo2, df, float(gx), dg, lo, up,
                                   C.byref(nn)))
        return dict(x=xv, xold1=xo1, xold2=xo2, low=lo, upp=up, newton=nn.value)

    # ---------------------------------------------- iteration-level driver
    def opt_begin(self, mode, V0, nLoop, rMin, rHat=0.0):
        m = {"GLOBAL": 0, "LOCAL": 1}[mode] if isinstance(mode, str) else int(mode)
        self._check(self.L.t3x_opt_begin(self.h, m, float(V0), int(nLoop), float(rMin), float(rHat)))

    def opt_step(self):
        it = Iter()
        fin = C.c_int(0)
        self._check(self.L.t3x_opt_step(self.h, C.byref(it), C.byref(fin)))
        rec = dict(loop=it.loop, pcgIts=it.pcgIts, filterIts=it.filterIts, c=it.c, vol=it.vol, cons=it.cons,
                   sharp=it.sharp, change=it.change, res=it.res, t=list(it.t))
        return rec, bool(fin.value)

    def kernel_time(self, group):
        """(launches, total ms) of a timed kernel group (needs timing=1)."""
        n = C.c_longlong(0)
        ms = C.c_double(0.0)
        self._check(self.L.t3x_kernel_time(self.h, int(group), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_passive(self, voxel_ids):
        """Passive elements: 0-based voxel ids of the unpadded grid (t3x_set_passive)."""
        ids = np.ascontiguousarray(voxel_ids, dtype=np.int32)
        self._keep.append(ids)
        self._check(self.L.t3x_set_passive(self.h, len(ids), ids.ctypes.data if len(ids) else None))

    def set_passive_layers(self, n_boundary, n_loads, n_fixations):
        """specifyPassiveRegions_ (37): TopOpti_SetPassiveElements (t3x_set_passive_layers)."""
        self._check(self.L.t3x_set_passive_layers(self.h, int(round(n_boundary)), int(round(n_loads)),
                                                  int(round(n_fixations))))

    def passive(self):
        """Passive voxel ids, 0-based on the unpadded grid (t3x_get_passive)."""
        n = C.c_int(0)
        self._check(self.L.t3x_get_passive(self.h, 0, None, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        self._check(self.L.t3x_get_passive(self.h, n.value, out.ctypes.data, C.byref(n)))
        return out[:n.value]

    def memory(self):
        """Device bytes the problem holds (t3x_memory)."""
        out = C.c_longlong(0)
        self._check(self.L.t3x_memory(self.h, C.byref(out)))
        return out.value

    def kernel_time_reset(self):
        self._check(self.L.t3x_kernel_time_reset(self.h))

    def run(self, mode, V0, nLoop, rMin, rHat=0.0, nSnap=0):
        nE = self.nE
        n = max(nLoop, 1)
        xPhys = np.zeros(nE)
        nc = self.num_cases
        cH, vH, conH, sH = (np.zeros(n) for _ in range(4))
        its = np.zeros(n * nc, dtype=np.int32)
        tH = np.zeros(n * 6)
        snap = np.zeros(max(nSnap, 1) * nE)
        resH = np.zeros(n * nc)
        R = Result(xPhys.ctypes.data, cH.ctypes.data, vH.ctypes.data, conH.ctypes.data, sH.ctypes.data,
                   its.ctypes.data, tH.ctypes.data, snap.ctypes.data, int(nSnap), resH.ctypes.data, 0.0, 0, 0)
        m = {"GLOBAL": 0, "LOCAL": 1}[mode] if isinstance(mode, str) else int(mode)
        self._check(self.L.t3x_run(self.h, m, float(V0), int(nLoop), float(rMin), float(rHat), C.byref(R)))
        k = R.loops
        return dict(xPhys=xPhys, cHist=cH[:k], volHist=vH[:k], consHist=conH[:k], sharpHist=sH[:k],
                    lssIts=its[:k * nc], resHist=resH[:k * nc], tHist=tH[:6 * k].reshape(k, 6),
                    xSnap=snap[:min(nSnap, k) * nE].reshape(-1, nE),
                    cSolid=R.cSolid, loops=k, warnings=R.warnings)


def TOP3D_XL(inputModel, consType, V0, nLoop, rMin, *varargin, **params):
    """Drop-in for TOP3D_XL.m:4-19.  inputModel is a logical (nely,nelx,nelz)
    array (the built-in cuboid BCs apply, 529-560) or a (voxels, fix, load)
    or (voxels, fix, load, extra_loads) tuple for explicit boundary
    conditions and further load cases.  out_dir=path also writes the
    reference's output files there (*_Target.dat, DesignVolume.nii/.stl)."""
    extra = ()
    if isinstance(inputModel, tuple):
        vox, fix, load = inputModel[:3]
        if len(inputModel) > 3:
            extra = inputModel[3]
    else:
        vox, fix, load = inputModel, None, None
    if consType == "GLOBAL":
        rHat = 0.0
        mode = 0
    elif consType == "LOCAL":
        if len(varargin) != 1:
            raise T3XError("Wrong Input!")  # 16
        rHat = float(varargin[0])
        mode = 1
    else:
        raise T3XError("Wrong Input!")
    out_dir = params.pop("out_dir", None)
    P = Problem(vox, fix, load, extra_loads=extra, **params)
    try:
        r = P.run(mode, V0, nLoop, rMin, rHat)
        if out_dir is not None:  # outPath (186-221 / 402-440)
            from .results import write_design_outputs
            lv = P.levels[0]
            write_design_outputs(out_dir, r, vox, (lv["resY"], lv["resX"], lv["resZ"]))
        return r
    finally:
        P.close()


def footprint(nely, nelx, nelz, voxels=None, mode="GLOBAL", **params):
    """Device bytes one problem allocates over a run (upper bound, t3x_footprint)."""
    L = load_library()
    P = default_params(**params)
    out = C.c_longlong(0)
    buf = None
    if voxels is not None:
        buf = np.ascontiguousarray(np.asarray(voxels, dtype=np.uint8).reshape(-1, order="F"))
    m = {"GLOBAL": 0, "LOCAL": 1}[mode] if isinstance(mode, str) else int(mode)
    rc = L.t3x_footprint(None if buf is None else buf.ctypes.data, nely, nelx, nelz, C.byref(P), m, C.byref(out))
    if rc != T3X_OK:
        raise T3XError(f"t3x_footprint failed ({_STATUS.get(rc, rc)})")
    return out.value


def stream_copy_ms(nbytes, reps=20):
    """Milliseconds per device-to-device copy of nbytes (t3x_stream_copy)."""
    L = load_library()
    ms = C.c_double(0.0)
    rc = L.t3x_stream_copy(int(nbytes), int(reps), C.byref(ms))
    if rc != T3X_OK:
        raise T3XError(f"t3x_stream_copy failed ({_STATUS.get(rc, rc)})")
    return ms.value


def device_synchronize():
    """hipDeviceSynchronize of the HIP runtime libt3x is bound to (a caller
    that also loads another copy of the runtime, e.g. through torch, would
    otherwise wait on the wrong one)."""
    L = load_library()
    L.hipDeviceSynchronize.restype = C.c_int
    rc = L.hipDeviceSynchronize()
    if rc != 0:
        raise T3XError(f"hipDeviceSynchronize failed ({rc})")


def oc_noalloc(n_cells, x_ptr, xphys_ptr, dc_ptr, passive_ptr, V0, stream=0):
    """t3x_oc_noalloc on device arrays given by address (e.g. torch tensors'
    data_ptr()): x updated in place on the device; returns (change, nBisect)."""
    L = load_library()
    ch, nb = C.c_double(0.0), C.c_int(0)
    err = C.create_string_buffer(256)
    rc = L.t3x_oc_noalloc(int(n_cells), x_ptr, xphys_ptr, dc_ptr, passive_ptr or None, float(V0), stream or None,
                          C.byref(ch), C.byref(nb), err, len(err))
    if rc != T3X_OK:
        raise T3XError(f"t3x_oc_noalloc failed ({_STATUS.get(rc, rc)}): {err.value.decode()}")
    return ch.value, nb.value


class Server:
    """Continuous-admission serving engine (include/t3x_serve.h)."""

    def __init__(self, max_active=4, mem_budget=0, **params):
        L = load_library()
        cfg = ServerConfig(mem_budget, max_active, default_params(**params))
        h = C.c_void_p()
        rc = L.t3x_server_create(C.byref(cfg), C.byref(h))
        if rc != T3X_OK:
            raise T3XError(f"t3x_server_create failed ({_STATUS.get(rc, rc)})")
        self.L, self.h = L, h

    def submit(self, tag, nely, nelx, nelz, mode="GLOBAL", V0=0.12, nLoop=30, rMin=3 ** 0.5, rHat=0.0,
               load_dir=(0.0, 0.0, -1.0), load_mag=1.0):
        m = {"GLOBAL": 0, "LOCAL": 1}[mode] if isinstance(mode, str) else int(mode)
        r = Request(tag, nely, nelx, nelz, m, V0, nLoop, rMin, rHat, (C.c_double * 3)(*load_dir), load_mag)
        rc = self.L.t3x_server_submit(self.h, C.byref(r))
        if rc != T3X_OK:
            raise T3XError(f"t3x_server_submit failed ({_STATUS.get(rc, rc)})")

    def poll(self, max_n=64, wait_ms=100):
        """Finished requests as dicts (results copied, server memory released)."""
        buf = (Completion * max_n)()
        n = C.c_int(0)
        self.L.t3x_server_poll(self.h, buf, max_n, wait_ms, C.byref(n))
        out = []
        for c in buf[:n.value]:
            d = dict(tag=c.tag, status=c.status, loops=c.loops, pcgIts=c.pcgIts, nE=c.nE, tSubmit=c.tSubmit,
                     tStart=c.tStart, tEnd=c.tEnd, cFinal=c.cFinal, cSolid=c.cSolid, footprint=c.footprint,
                     xPhys=np.ctypeslib.as_array(c.xPhys, (c.nE,)).copy() if c.nE else np.zeros(0),
                     cHist=np.ctypeslib.as_array(c.cHist, (c.loops,)).copy() if c.loops else np.zeros(0))
            self.L.t3x_server_release(self.h, c.tag)
            out.append(d)
        return out

    def stats(self):
        q, r, rd, m = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
        self.L.t3x_server_stats(self.h, C.byref(q), C.byref(r), C.byref(rd), C.byref(m))
        return dict(queued=q.value, running=r.value, ready=rd.value, mem_in_use=m.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.t3x_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
