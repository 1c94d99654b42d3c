"""ctypes bindings for the CPU oracle (liboracle.so) and the reference build
(_ref/libtop3d_ref*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = {}

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class OModel(C.Structure):
    _fields_ = [
        ("nelx", C.c_int), ("nely", C.c_int), ("nelz", C.c_int),
        ("resX", C.c_int), ("resY", C.c_int), ("resZ", C.c_int), ("baseL", C.c_int),
        ("nnx", C.c_int), ("nny", C.c_int), ("nnz", C.c_int),
        ("ne", C.c_int), ("nN", C.c_int), ("nDOF", C.c_int), ("nF", C.c_int),
        ("eleMapBack", C.POINTER(C.c_int32)),
        ("eNodMat", C.POINTER(C.c_int32)),
        ("color", C.POINTER(C.c_int8)),
        ("bucket", C.POINTER(C.c_int32)),
        ("bucketStart", C.c_int32 * 9),
        ("Ke", C.c_double * 576),
        ("F", C.POINTER(C.c_double)),
        ("isFree", C.POINTER(C.c_uint8)),
        ("freeDofIndex", C.POINTER(C.c_int32)),
        ("eDofFree", C.POINTER(C.c_int32)),
        ("cellToCompact", C.POINTER(C.c_int32)),
        ("compactToCell", C.POINTER(C.c_int32)),
    ]


O_MAXLEV = 8


class OMg(C.Structure):
    _fields_ = [
        ("nlev", C.c_int),
        ("resX", C.c_int * O_MAXLEV), ("resY", C.c_int * O_MAXLEV), ("resZ", C.c_int * O_MAXLEV),
        ("span", C.c_int * O_MAXLEV), ("nN", C.c_int * O_MAXLEV),
        ("W", C.POINTER(C.c_double) * O_MAXLEV),
        ("mask", C.POINTER(C.c_uint8) * O_MAXLEV),
        ("diag", C.POINTER(C.c_double) * O_MAXLEV),
        ("Lc", C.POINTER(C.c_double)), ("Nc", C.c_int),
        ("weight", C.c_double), ("Emin", C.c_double),
        ("NlimitDofs", C.c_int), ("cholOk", C.c_int),
        ("Kc", C.POINTER(C.c_double)),
        ("r", C.POINTER(C.c_double) * O_MAXLEV),
        ("x", C.POINTER(C.c_double) * O_MAXLEV),
        ("tmp", C.POINTER(C.c_double)),
    ]


class OFilter(C.Structure):
    _fields_ = [
        ("kernel", C.c_double * 64),
        ("diagPrecond", C.POINTER(C.c_double)),
        ("lastX", C.POINTER(C.c_double)),
        ("lastRhs", C.POINTER(C.c_double)),
        ("lastIters", C.c_int), ("lastWarm", C.c_int),
    ]


class OParams(C.Structure):
    _fields_ = [
        ("E0", C.c_float), ("Emin", C.c_float), ("nu", C.c_float), ("penal", C.c_float), ("cgTol", C.c_float),
        ("cgMaxIt", C.c_int), ("V0", C.c_float), ("nLoop", C.c_int), ("rMin", C.c_float),
        ("mgMaxLevels", C.c_int), ("mgWeight", C.c_float), ("nonDyadic", C.c_int), ("NlimitDofs", C.c_int),
        ("rebuildEvery", C.c_int),
    ]


class OLoopRec(C.Structure):
    _fields_ = [
        ("loop", C.c_int), ("pcgIters", C.c_int), ("filterIters", C.c_int), ("bisectSteps", C.c_int),
        ("C", C.c_double), ("Cdisp", C.c_double),
        ("volFrac", C.c_float), ("sharpness", C.c_float), ("change", C.c_float), ("fval", C.c_float),
        ("xCheck", C.c_uint64),
    ]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.o_model_new.restype = C.POINTER(OModel)
        L.o_model_new.argtypes = [C.c_int, C.c_int, C.c_int]
        L.o_model_new_voxels.restype = C.POINTER(OModel)
        L.o_model_new_voxels.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.o_model_new_bc.restype = C.POINTER(OModel)
        L.o_model_new_bc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_void_p]
        L.o_top3d_run_bc.restype = C.c_int
        L.o_top3d_run_bc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_int, C.c_void_p, C.c_void_p, C.POINTER(OParams), C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.o_model_free.argtypes = [C.POINTER(OModel)]
        L.o_canon_sum.restype = C.c_double
        L.o_canon_sum.argtypes = [f64p, C.c_size_t]
        L.o_canon_dot.restype = C.c_double
        L.o_canon_dot.argtypes = [f64p, f64p, C.c_size_t]
        L.o_voxel_ke.argtypes = [C.c_double, C.c_double, f64p]
        L.o_padding.argtypes = [C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4
        for fn in ("o_free_to_full", "o_full_to_free"):
            getattr(L, fn).argtypes = [C.POINTER(OModel), f64p, f64p]
        L.o_kxu_free.argtypes = [C.POINTER(OModel), f64p, f64p, f64p]
        L.o_kxu_full.argtypes = [C.POINTER(OModel), f64p, f64p, f64p]
        L.o_kxu_full_f32.argtypes = [C.POINTER(OModel), f32p, f32p, f32p]
        L.o_compliance.restype = C.c_double
        L.o_compliance.argtypes = [C.POINTER(OModel), f64p, f64p, f64p]
        L.o_mg_new.restype = C.POINTER(OMg)
        L.o_mg_new.argtypes = [C.POINTER(OModel), C.c_int, C.c_int, C.c_float, C.c_float, C.c_int]
        L.o_mg_free.argtypes = [C.POINTER(OMg)]
        L.o_mg_setup.argtypes = [C.POINTER(OModel), C.POINTER(OMg), f64p]
        L.o_mg_apply_full.argtypes = [C.POINTER(OModel), C.POINTER(OMg), f64p, f64p]
        for fn in ("o_restrict_vec3", "o_prolong_vec3", "o_restrict_scalar_div"):
            getattr(L, fn).argtypes = [C.POINTER(OMg), C.c_int, f64p, f64p]
        L.o_chol_inplace.restype = C.c_int
        L.o_chol_inplace.argtypes = [f64p, C.c_int]
        L.o_chol_solve.argtypes = [f64p, f64p, f64p, C.c_int]
        L.o_pcg_mg.restype = C.c_int
        L.o_pcg_mg.argtypes = [C.POINTER(OModel), C.POINTER(OMg), f64p, f64p, f64p, C.c_double, C.c_int]
        L.o_filter_new.restype = C.POINTER(OFilter)
        L.o_filter_new.argtypes = [C.POINTER(OModel), C.c_float]
        L.o_filter_free.argtypes = [C.POINTER(OFilter)]
        L.o_filter_apply.argtypes = [C.POINTER(OModel), C.POINTER(OFilter), f32p, f32p]
        L.o_pde_matvec.argtypes = [C.POINTER(OModel), C.POINTER(OFilter), f64p, f64p]
        L.o_cr_cube.restype = C.c_float
        L.o_cr_cube.argtypes = [C.c_float]
        L.o_cr_square.restype = C.c_float
        L.o_cr_square.argtypes = [C.c_float]
        L.o_seq_sum.restype = C.c_float
        L.o_seq_sum.argtypes = [f32p, C.c_int]
        L.o_seq_sum_x1mx.restype = C.c_float
        L.o_seq_sum_x1mx.argtypes = [f32p, C.c_int]
        L.o_x_check.restype = C.c_uint64
        L.o_x_check.argtypes = [f32p, C.c_int64]
        L.o_simp.argtypes = [C.c_int, f32p, C.c_float, C.c_float, f64p]
        L.o_sensitivity.argtypes = [C.c_int, f32p, f64p, C.c_double, C.c_float, C.c_float, C.c_float, f32p]
        L.o_oc_update.restype = C.c_int
        L.o_oc_update.argtypes = [C.c_int, f32p, f32p, C.c_float, C.c_float, f32p, C.POINTER(C.c_float)]
        L.o_params_default.argtypes = [C.POINTER(OParams)]
        L.o_oc_update_passive.restype = C.c_int
        L.o_oc_update_passive.argtypes = [C.c_int, f32p, f32p, C.c_void_p, C.c_float, C.c_float, f32p,
                                          C.POINTER(C.c_float)]
        L.o_top3d_run_ex.restype = C.c_int
        L.o_top3d_run_ex.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(OParams),
                                     C.POINTER(OLoopRec), f32p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.o_stl_image.restype = C.c_uint64
        L.o_stl_image.argtypes = [C.c_int, C.c_int, C.c_int, f32p, C.c_float, C.c_void_p, C.c_uint64]
        L.o_top3d_run.restype = C.c_int
        L.o_top3d_run.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(OParams), C.POINTER(OLoopRec), f32p,
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Model:
    """Oracle model (fea.cpp:181-447): arrays are copied out as numpy."""

    def __init__(self, nely, nelx, nelz, solid=None, fix=None, loads=None):
        """solid: optional uint8 array indexed [ez, ex, ey] (flat (ez*nelx + ex)*nely + ey),
        nonzero = solid voxel; None = the full cuboid.  fix / loads: boundary
        conditions as lists (o_model_new_bc): fix int32 [n, 4] = x y z mask, loads a
        pair (nodes int32 [m, 3], forces [m, 3]); None for both: the cantilever."""
        L = lib()
        self.nely, self.nelx, self.nelz = nely, nelx, nelz
        if fix is not None or loads is not None:
            sm = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
            fx, ln, lf = bc_arrays(fix, loads)
            self._keep = (sm, fx, ln, lf)
            self.ptr = L.o_model_new_bc(nely, nelx, nelz, None if sm is None else sm.ctypes.data, len(fx),
                                        fx.ctypes.data, len(ln), ln.ctypes.data, lf.ctypes.data)
            if not self.ptr:
                raise ValueError("boundary condition node outside the box")
        elif solid is None:
            self.ptr = L.o_model_new(nely, nelx, nelz)
        else:
            sm = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
            assert sm.size == nely * nelx * nelz
            self.ptr = L.o_model_new_voxels(nely, nelx, nelz, sm.ctypes.data)
        m = self.ptr.contents
        self.resX, self.resY, self.resZ, self.baseL = m.resX, m.resY, m.resZ, m.baseL
        self.ne, self.nN, self.nDOF, self.nF = m.ne, m.nN, m.nDOF, m.nF
        self.eleMapBack = _arr(m.eleMapBack, m.ne, np.int32)
        self.eNodMat = _arr(m.eNodMat, 8 * m.ne, np.int32)
        self.color = _arr(m.color, m.ne, np.int32)
        self.Ke = np.array(m.Ke[:], dtype=np.float64)
        self.F = _arr(m.F, m.nDOF, np.float64)
        self.isFree = _arr(m.isFree, m.nDOF, np.uint8)
        self.freeDofIndex = _arr(m.freeDofIndex, m.nF, np.int32)
        self.eDofFree = _arr(m.eDofFree, 24 * m.ne, np.int32)
        self.compactToCell = _arr(m.compactToCell, m.ne, np.int32)
        self.b = np.where(self.isFree.astype(bool), self.F, 0.0)

    def __del__(self):
        try:
            lib().o_model_free(self.ptr)
        except Exception:
            pass

    def free_to_full(self, v):
        out = np.zeros(self.nDOF)
        lib().o_free_to_full(self.ptr, np.ascontiguousarray(v, np.float64), out)
        return out

    def full_to_free(self, v):
        out = np.zeros(self.nF)
        lib().o_full_to_free(self.ptr, np.ascontiguousarray(v, np.float64), out)
        return out

    def kxu_free(self, E, u):
        y = np.zeros(self.nF)
        lib().o_kxu_free(self.ptr, np.ascontiguousarray(E, np.float64), np.ascontiguousarray(u, np.float64), y)
        return y

    def kxu_full(self, E, u):
        y = np.zeros(self.nDOF)
        lib().o_kxu_full(self.ptr, np.ascontiguousarray(E, np.float64), np.ascontiguousarray(u, np.float64), y)
        return y

    def kxu_full_f32(self, E, u):
        y = np.zeros(self.nDOF, np.float32)
        lib().o_kxu_full_f32(self.ptr, np.ascontiguousarray(E, np.float32), np.ascontiguousarray(u, np.float32), y)
        return y

    def compliance(self, E, U):
        ce = np.zeros(self.ne)
        C = lib().o_compliance(self.ptr, np.ascontiguousarray(E, np.float64), np.ascontiguousarray(U, np.float64), ce)
        return C, ce


class MG:
    """Oracle multigrid (multigrid/*.cpp)."""

    def __init__(self, model, nonDyadic=1, maxLevels=5, weight=0.6, Emin=1e-6, NlimitDofs=1000):
        self.model = model
        self.ptr = lib().o_mg_new(model.ptr, nonDyadic, maxLevels, weight, Emin, NlimitDofs)
        g = self.ptr.contents
        self.nlev = g.nlev
        self.dims = [(g.resX[l], g.resY[l], g.resZ[l], g.span[l]) for l in range(g.nlev)]
        self.nN = [g.nN[l] for l in range(g.nlev)]
        self.masks = [_arr(g.mask[l], 3 * g.nN[l], np.uint8) for l in range(g.nlev)]

    def __del__(self):
        try:
            lib().o_mg_free(self.ptr)
        except Exception:
            pass

    def setup(self, E):
        lib().o_mg_setup(self.model.ptr, self.ptr, np.ascontiguousarray(E, np.float64))
        g = self.ptr.contents
        self.diag = [_arr(g.diag[l], 3 * g.nN[l], np.float64) for l in range(g.nlev)]
        self.Nc = g.Nc
        self.cholOk = g.cholOk
        N = 3 * g.nN[g.nlev - 1]
        self.Kc = _arr(g.Kc, N * N, np.float64).reshape(N, N) if bool(g.Kc) else None
        self.Lc = _arr(g.Lc, g.Nc * g.Nc, np.float64).reshape(g.Nc, g.Nc) if g.Nc else None

    def apply(self, r):
        z = np.zeros(self.model.nDOF)
        lib().o_mg_apply_full(self.model.ptr, self.ptr, np.ascontiguousarray(r, np.float64), z)
        return z

    def restrict_vec3(self, lc, rf):
        rc = np.zeros(3 * self.nN[lc])
        lib().o_restrict_vec3(self.ptr, lc, np.ascontiguousarray(rf, np.float64), rc)
        return rc

    def prolong_vec3(self, lc, xc):
        xf = np.zeros(3 * self.nN[lc - 1])
        lib().o_prolong_vec3(self.ptr, lc, np.ascontiguousarray(xc, np.float64), xf)
        return xf

    def pcg(self, E, b, x, tol, maxIt, precond=True):
        x = np.ascontiguousarray(x, np.float64).copy()
        it = lib().o_pcg_mg(self.model.ptr, self.ptr if precond else None, np.ascontiguousarray(E, np.float64),
                            np.ascontiguousarray(b, np.float64), x, tol, maxIt)
        return it, x


class Filter:
    """Oracle PDE filter (filter.cpp:9-275)."""

    def __init__(self, model, rMin):
        self.model = model
        self.ptr = lib().o_filter_new(model.ptr, rMin)
        f = self.ptr.contents
        self.kernel = np.array(f.kernel[:])
        self.diag = _arr(f.diagPrecond, model.nN, np.float64)

    def __del__(self):
        try:
            lib().o_filter_free(self.ptr)
        except Exception:
            pass

    def apply(self, src):
        dst = np.zeros(self.model.ne, np.float32)
        lib().o_filter_apply(self.model.ptr, self.ptr, np.ascontiguousarray(src, np.float32), dst)
        f = self.ptr.contents
        return dst, f.lastIters, f.lastWarm

    def matvec(self, x):
        y = np.zeros(self.model.nN)
        lib().o_pde_matvec(self.model.ptr, self.ptr, np.ascontiguousarray(x, np.float64), y)
        return y


def canon_sum(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().o_canon_sum(v, v.size)


def canon_dot(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().o_canon_dot(a, b, a.size)


def default_params(**kw):
    p = OParams()
    lib().o_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def bc_arrays(fix, loads):
    fx = np.ascontiguousarray(np.zeros((0, 4)) if fix is None else fix, np.int32).reshape(-1, 4)
    ln, lf = (np.zeros((0, 3)), np.zeros((0, 3))) if loads is None else loads
    return fx, np.ascontiguousarray(ln, np.int32).reshape(-1, 3), np.ascontiguousarray(lf, np.float64).reshape(-1, 3)


def run(nely, nelx, nelz, solid=None, passive=None, fix=None, loads=None, **kw):
    """Solid solve + OC loops (driver.cpp:23-218), on a voxel domain when solid is
    given, with passive voxels held at 1 (matlab/TOP3D_XL.m:154) when passive is
    given (uint8 masks indexed [ez, ex, ey]). Returns (records, x, Csolid, solidIters)."""
    p = default_params(**kw)
    rec = (OLoopRec * max(1, p.nLoop))()
    m = Model(nely, nelx, nelz, solid=solid)
    x = np.zeros(m.ne, np.float32)
    del m
    Cs = C.c_double()
    si = C.c_int()
    sm = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
    pm = None if passive is None else np.ascontiguousarray(passive, np.uint8).reshape(-1)
    if fix is not None or loads is not None:
        fx, ln, lf = bc_arrays(fix, loads)
        n = lib().o_top3d_run_bc(nely, nelx, nelz, None if sm is None else sm.ctypes.data,
                                 None if pm is None else pm.ctypes.data, len(fx), fx.ctypes.data, len(ln),
                                 ln.ctypes.data, lf.ctypes.data, C.byref(p), C.cast(rec, C.c_void_p),
                                 x.ctypes.data, C.byref(Cs), C.byref(si))
        if n < 0:
            raise ValueError("boundary condition node outside the box")
    else:
        n = lib().o_top3d_run_ex(nely, nelx, nelz, None if sm is None else sm.ctypes.data,
                                 None if pm is None else pm.ctypes.data, C.byref(p), rec, x, C.byref(Cs),
                                 C.byref(si))
    recs = [{f: getattr(r, f) for f, _ in OLoopRec._fields_} for r in rec[:n]]
    return recs, x, Cs.value, si.value


def stl_volume(model, xPhys):
    """The padded volume of export_surface_stl (io.cpp:81-83): vol[eleMapBack[e]] =
    xPhys[e], 0 on cells without an element; vol[y + ny x + ny nx z]."""
    vol = np.zeros(model.resX * model.resY * model.resZ, np.float32)
    vol[model.eleMapBack] = np.asarray(xPhys, np.float32)
    return vol


def stl_image(ny, nx, nz, vol, iso):
    """Bytes of the binary STL of extract_faces + write_stl_binary
    (voxel_surface.hpp:12-93, export_stl.hpp:14-33) for vol[ny*nx*nz]."""
    vol = np.ascontiguousarray(vol, np.float32).reshape(-1)
    assert vol.size == ny * nx * nz
    n = lib().o_stl_image(ny, nx, nz, vol, float(iso), None, 0)
    out = np.zeros(n, np.uint8)
    lib().o_stl_image(ny, nx, nz, vol, float(iso), out.ctypes.data, n)
    return out.tobytes()


# ---------------------------------------------------------------------------
# reference build (oracle/_ref), present only where it was built
# ---------------------------------------------------------------------------
def ref_path(fast=False):
    return os.path.join(HERE, "_ref", "libtop3d_ref_fast.so" if fast else "libtop3d_ref.so")


def have_ref(fast=False):
    return os.path.exists(ref_path(fast))


def ref(fast=False):
    key = bool(fast)
    if key in _ref:
        return _ref[key]
    R = C.CDLL(ref_path(fast))
    ii = [C.c_int, C.c_int, C.c_int]
    R.ref_set_hierarchy.argtypes = [C.c_int, C.c_int]
    R.ref_sizes.argtypes = ii + [i32p]
    R.ref_model.argtypes = ii + [i32p, i32p, i32p, f64p, f64p, u8p, i32p, i32p]
    R.ref_bfree.argtypes = ii + [f64p]
    R.ref_kxu.argtypes = ii + [f64p, f64p, f64p]
    R.ref_compliance.restype = C.c_double
    R.ref_compliance.argtypes = ii + [f64p, f64p, f64p]
    R.ref_mg_levels.restype = C.c_int
    R.ref_mg_levels.argtypes = ii + [i32p]
    R.ref_mg_mask.argtypes = ii + [C.c_int, u8p]
    R.ref_mg_diag.argtypes = ii + [f64p, C.c_int, f64p]
    R.ref_mg_coarsest.restype = C.c_int
    R.ref_mg_coarsest.argtypes = ii + [f64p, C.c_float, f64p, f64p, C.POINTER(C.c_int)]
    R.ref_chol_solve.argtypes = [f64p, f64p, f64p, C.c_int]
    R.ref_chol.restype = C.c_int
    R.ref_chol.argtypes = [f64p, C.c_int]
    R.ref_restrict_vec3.argtypes = ii + [C.c_int, f64p, f64p]
    R.ref_prolong_vec3.argtypes = ii + [C.c_int, f64p, f64p]
    R.ref_mg_apply.argtypes = ii + [f64p, C.c_float, f64p, f64p]
    R.ref_pcg.restype = C.c_int
    R.ref_pcg.argtypes = ii + [f64p, C.c_float, f64p, f64p, C.c_double, C.c_int, C.c_int]
    R.ref_filter_setup.argtypes = ii + [C.c_float, f64p, f64p]
    R.ref_filter_apply.argtypes = ii + [f32p, f32p]
    R.ref_run.restype = C.c_int
    R.ref_run.argtypes = ii + [C.c_float, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, f64p, f32p,
                               C.POINTER(C.c_double)]
    R.ref_time_kxu.restype = C.c_double
    R.ref_time_kxu.argtypes = ii + [f64p, f64p, C.c_int]
    R.ref_time_pcg.restype = C.c_double
    R.ref_time_pcg.argtypes = ii + [f64p, C.c_float, C.c_double, C.c_int, C.POINTER(C.c_int)]
    R.ref_export_stl.restype = C.c_int
    R.ref_export_stl.argtypes = ii + [f32p, C.c_char_p, C.c_float]
    R.ref_stl_volume.restype = C.c_int
    R.ref_stl_volume.argtypes = ii + [f32p, C.c_float, C.c_char_p]
    _ref[key] = R
    return R


class RefModel:
    """The reference's own model arrays (via oracle/_ref); the multigrid
    functions use the hierarchy (nonDyadic, maxLevels) of MGPrecondConfig
    (multigrid.hpp:9-13), by default the reference driver's (1, 5)."""

    def __init__(self, nely, nelx, nelz, fast=False, nonDyadic=1, maxLevels=5):
        R = ref(fast)
        self.R = R
        self.dims = (nely, nelx, nelz)
        self.hier = (int(nonDyadic), int(maxLevels))
        self._sel()
        s = np.zeros(7, np.int32)
        R.ref_sizes(nely, nelx, nelz, s)
        self.resX, self.resY, self.resZ, self.ne, self.nN, self.nF, self.nlev = [int(v) for v in s]
        ne, nN, nF = self.ne, self.nN, self.nF
        self.eleMapBack = np.zeros(ne, np.int32)
        self.eNodMat = np.zeros(8 * ne, np.int32)
        self.color = np.zeros(ne, np.int32)
        self.Ke = np.zeros(576)
        self.F = np.zeros(3 * nN)
        self.isFree = np.zeros(3 * nN, np.uint8)
        self.freeDofIndex = np.zeros(nF, np.int32)
        self.eDofFree = np.zeros(24 * ne, np.int32)
        R.ref_model(nely, nelx, nelz, self.eleMapBack, self.eNodMat, self.color, self.Ke, self.F, self.isFree,
                    self.freeDofIndex, self.eDofFree)
        self.bFree = np.zeros(nF)
        R.ref_bfree(nely, nelx, nelz, self.bFree)

    def _sel(self):
        self.R.ref_set_hierarchy(*self.hier)

    def kxu(self, E, u):
        y = np.zeros(self.nF)
        self.R.ref_kxu(*self.dims, np.ascontiguousarray(E, np.float64), np.ascontiguousarray(u, np.float64), y)
        return y

    def compliance(self, E, uFree):
        ce = np.zeros(self.ne)
        C_ = self.R.ref_compliance(*self.dims, np.ascontiguousarray(E, np.float64),
                                   np.ascontiguousarray(uFree, np.float64), ce)
        return C_, ce

    def levels(self):
        self._sel()
        out = np.zeros(4 * 8, np.int32)
        n = self.R.ref_mg_levels(*self.dims, out)
        return [tuple(int(v) for v in out[4 * l:4 * l + 4]) for l in range(n)]

    def mask(self, level, n_nodes):
        self._sel()
        out = np.zeros(3 * n_nodes, np.uint8)
        self.R.ref_mg_mask(*self.dims, level, out)
        return out

    def diag(self, E, level, n_nodes):
        self._sel()
        out = np.zeros(3 * n_nodes)
        self.R.ref_mg_diag(*self.dims, np.ascontiguousarray(E, np.float64), level, out)
        return out

    def coarsest(self, E, Emin, N):
        self._sel()
        Kc = np.zeros(N * N)
        Lc = np.zeros(N * N)
        ok = C.c_int()
        self.R.ref_mg_coarsest(*self.dims, np.ascontiguousarray(E, np.float64), Emin, Kc, Lc, C.byref(ok))
        return Kc.reshape(N, N), Lc.reshape(N, N), ok.value

    def restrict_vec3(self, lc, rf, nc):
        self._sel()
        rc = np.zeros(3 * nc)
        self.R.ref_restrict_vec3(*self.dims, lc, np.ascontiguousarray(rf, np.float64), rc)
        return rc

    def prolong_vec3(self, lc, xc, nf):
        self._sel()
        xf = np.zeros(3 * nf)
        self.R.ref_prolong_vec3(*self.dims, lc, np.ascontiguousarray(xc, np.float64), xf)
        return xf

    def mg_apply(self, E, Emin, rFree):
        self._sel()
        z = np.zeros(self.nF)
        self.R.ref_mg_apply(*self.dims, np.ascontiguousarray(E, np.float64), Emin,
                            np.ascontiguousarray(rFree, np.float64), z)
        return z

    def pcg(self, E, Emin, bFree, x0, tol, maxIt, useMG=True):
        self._sel()
        x = np.ascontiguousarray(x0, np.float64).copy()
        it = self.R.ref_pcg(*self.dims, np.ascontiguousarray(E, np.float64), Emin,
                            np.ascontiguousarray(bFree, np.float64), x, tol, maxIt, 1 if useMG else 0)
        return it, x

    def filter_setup(self, rMin):
        k = np.zeros(64)
        d = np.zeros(self.nN)
        self.R.ref_filter_setup(*self.dims, rMin, k, d)
        return k, d

    def filter_apply(self, src):
        dst = np.zeros(self.ne, np.float32)
        self.R.ref_filter_apply(*self.dims, np.ascontiguousarray(src, np.float32), dst)
        return dst


def ref_run(nely, nelx, nelz, V0=0.12, nLoop=3, rMin=3 ** 0.5, Emin=1e-6, cgTol=1e-3, cgMaxIt=800):
    R = ref(False)
    R.ref_set_hierarchy(1, 5)
    rec = np.zeros(6 * max(1, nLoop))
    nel = nely * nelx * nelz
    x = np.zeros(nel, np.float32)
    Cs = C.c_double()
    n = R.ref_run(nely, nelx, nelz, V0, nLoop, rMin, Emin, cgTol, cgMaxIt, rec, x, C.byref(Cs))
    recs = [dict(Cdisp=rec[6 * k], C=rec[6 * k + 1], volFrac=np.float32(rec[6 * k + 2]),
                 sharpness=np.float32(rec[6 * k + 3]), change=np.float32(rec[6 * k + 4]), pcgIters=int(rec[6 * k + 5]))
            for k in range(n)]
    return recs, x, Cs.value
