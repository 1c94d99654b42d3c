"""MI355X-native TOP3D_XL finite-element inner loop.

Python host binding of the C ABI in include/t3d_capi.h (ctypes; no torch).
The compute path is the HIP library libt3d.so built for gfx950 by
csrc/Makefile; importing this module fails loudly when it is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libt3d.so")
REPO = os.path.dirname(HERE)

_lib = None

T3D_OK = 0


class T3DError(RuntimeError):
    pass


def build(verbose=False):
    """Compile libt3d.so for gfx950 (hipcc cross-compiles without a GPU)."""
    import subprocess
    cmd = ["make", "-C", os.path.join(HERE, "csrc"), "-j8"]
    r = subprocess.run(cmd, capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise T3DError("libt3d build failed:\n" + (r.stdout or "") + (r.stderr or ""))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise T3DError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "t3d_last_error": (C.c_char_p, []),
        "t3d_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
        "t3d_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
        "t3d_free": (C.c_int, [vp]),
        "t3d_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
        "t3d_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
        "t3d_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
        "t3d_memset": (C.c_int, [vp, C.c_int, C.c_size_t]),
        "t3d_synchronize": (C.c_int, []),
        "t3d_problem_create_cuboid": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
        "t3d_problem_create_voxels": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.POINTER(vp)]),
        "t3d_problem_create_bc": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp,
                                            C.POINTER(vp)]),
        "t3d_problem_create_model": (C.c_int, [vp, C.POINTER(vp)]),
        "t3d_topvoxel_read": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "t3d_voxel_model_destroy": (C.c_int, [vp]),
        "t3d_problem_destroy": (C.c_int, [vp]),
        "t3d_problem_sizes": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "t3d_problem_model": (C.c_int, [vp] + [vp] * 8),
        "t3d_free_to_full": (C.c_int, [vp, vp, vp]),
        "t3d_full_to_free": (C.c_int, [vp, vp, vp]),
        "t3d_load_vector": (C.c_int, [vp, vp]),
        "t3d_set_modulus": (C.c_int, [vp, vp]),
        "t3d_kxu_f64": (C.c_int, [vp, vp, vp]),
        "t3d_kxu_free_host": (C.c_int, [vp, vp, vp, vp]),
        "t3d_dot_f64": (C.c_int, [vp, vp, vp, C.c_int64, C.POINTER(C.c_double)]),
        "t3d_time_op": (C.c_int, [vp, C.c_int, vp, vp, C.c_int, C.POINTER(C.c_double)]),
        "t3d_mg_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.POINTER(vp)]),
        "t3d_mg_destroy": (C.c_int, [vp]),
        "t3d_mg_setup": (C.c_int, [vp, vp]),
        "t3d_mg_apply": (C.c_int, [vp, vp, vp]),
        "t3d_mg_info": (C.c_int, [vp, C.POINTER(C.c_int), vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "t3d_mg_mask": (C.c_int, [vp, C.c_int, vp]),
        "t3d_mg_diag": (C.c_int, [vp, C.c_int, vp]),
        "t3d_mg_coarsest": (C.c_int, [vp, vp, vp]),
        "t3d_mg_set_mode": (C.c_int, [vp, C.c_int]),
        "t3d_mg_set_vcycle": (C.c_int, [vp, C.c_int]),
        "t3d_pcg": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, C.POINTER(C.c_int)]),
        "t3d_pcg_history": (C.c_int, [vp, C.c_int, C.c_int, vp]),
        "t3d_compliance": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double)]),
        "t3d_voxel_ke": (C.c_int, [C.c_double, C.c_double, vp]),
        "t3d_compliance_noalloc": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double)]),
        "t3d_filter_create": (C.c_int, [vp, C.c_float, C.POINTER(vp)]),
        "t3d_filter_destroy": (C.c_int, [vp]),
        "t3d_filter_info": (C.c_int, [vp, vp, vp]),
        "t3d_filter_apply": (C.c_int, [vp, vp, vp, C.POINTER(C.c_int)]),
        "t3d_simp": (C.c_int, [vp, vp, C.c_float, C.c_float, vp]),
        "t3d_oc_update": (C.c_int, [vp, vp, vp, C.c_float, C.c_float, vp, vp]),
        "t3d_oc_update_passive": (C.c_int, [vp, vp, vp, vp, C.c_float, C.c_float, vp, vp]),
        "t3d_seq_sum_f32": (C.c_int, [vp, vp, C.c_int64, C.POINTER(C.c_float)]),
        "t3d_params_default": (None, [vp]),
        "t3d_mixed_create": (C.c_int, [vp, vp, C.POINTER(vp)]),
        "t3d_mixed_destroy": (C.c_int, [vp]),
        "t3d_mixed_setup": (C.c_int, [vp]),
        "t3d_mixed_coarse_solve": (C.c_int, [vp, vp, vp]),
        "t3d_mixed_vcycle": (C.c_int, [vp, vp, vp]),
        "t3d_mixed_true_residual": (C.c_int, [vp, vp, vp, vp, C.POINTER(C.c_double)]),
        "t3d_kxu_f32": (C.c_int, [vp, C.c_int, vp, vp]),
        "t3d_batch_create": (C.c_int, [C.c_int, vp, vp, vp, C.c_float, C.POINTER(C.c_void_p)]),
        "t3d_batch_create_voxels": (C.c_int, [C.c_int, vp, vp, vp, vp, C.c_float, C.POINTER(C.c_void_p)]),
        "t3d_batch_destroy": (C.c_int, [vp]),
        "t3d_batch_problem": (C.c_int, [vp, C.c_int, C.POINTER(C.c_void_p)]),
        "t3d_batch_mg": (C.c_int, [vp, C.c_int, C.POINTER(C.c_void_p)]),
        "t3d_batch_setup": (C.c_int, [vp, vp]),
        "t3d_batch_set_tiles": (C.c_int, [vp, C.c_int, C.c_int]),
        "t3d_batch_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "t3d_batch_set_timing": (C.c_int, [vp, C.c_int]),
        "t3d_batch_get_timing": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                           C.POINTER(C.c_longlong)]),
        "t3d_batch_solve": (C.c_int, [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "t3d_mixed_set_operator": (C.c_int, [vp, C.c_int]),
        "t3d_mixed_time_op": (C.c_int, [vp, C.c_int, vp, vp, C.c_int, C.POINTER(C.c_float)]),
        "t3d_mixed_set_timing": (C.c_int, [vp, C.c_int]),
        "t3d_mixed_get_timing": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "t3d_pcg_mixed": (C.c_int, [vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
        "t3d_fp_probe": (C.c_int, [C.c_int, C.c_int64, vp, vp, vp]),
        "t3d_top3d_run": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, vp, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "t3d_summary_format": (C.c_int, [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "t3d_stl_size": (C.c_int, [vp, vp, C.c_float, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "t3d_stl_image": (C.c_int, [vp, vp, C.c_float, vp, C.c_uint64]),
        "t3d_export_surface_stl": (C.c_int, [vp, vp, C.c_char_p, C.c_float]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    if rc != T3D_OK:
        msg = lib().t3d_last_error().decode(errors="replace")
        raise T3DError(f"{what} failed ({rc}): {msg}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A device allocation owned by Python (t3d_malloc/t3d_free)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().t3d_malloc(C.byref(p), max(1, self.nbytes)), "t3d_malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().t3d_memcpy_h2d(self.ptr, _ptr(a), self.nbytes), "h2d")

    def download(self):
        out = np.empty(self.n, self.dtype)
        check(lib().t3d_memcpy_d2h(_ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def zero(self):
        check(lib().t3d_memset(self.ptr, 0, self.nbytes), "memset")

    def __del__(self):
        try:
            if self.ptr:
                lib().t3d_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


class Problem:
    """Device problem: CreateVoxelFEAmodel_Cuboid + ApplyBoundaryConditions
    (cpp/src/fea.cpp:181-447) with device mirrors."""

    def __init__(self, nely, nelx, nelz, solid=None, fix=None, loads=None, model=None):
        """solid: optional voxel mask, uint8 indexed [ez, ex, ey] (flat (ez*nelx + ex)*nely + ey),
        nonzero = solid; void voxels carry no element.  None: the full cuboid.
        fix / loads: boundary conditions as lists instead of the cantilever's
        (t3d_problem_create_bc): fix int32 [n, 4] = x y z mask, loads a pair
        (nodes int32 [m, 3] = x y z, forces float64 [m, 3]); nodes on the original
        box, y from its bottom.  model: a VoxelModel (domain, fixations, loads)."""
        p = C.c_void_p()
        if model is not None:
            check(lib().t3d_problem_create_model(model.h, C.byref(p)), "t3d_problem_create_model")
        elif fix is not None or loads is not None:
            sm = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
            fx = np.ascontiguousarray(np.zeros((0, 4)) if fix is None else fix, np.int32).reshape(-1, 4)
            ln, lf = (np.zeros((0, 3)), np.zeros((0, 3))) if loads is None else loads
            ln = np.ascontiguousarray(ln, np.int32).reshape(-1, 3)
            lf = np.ascontiguousarray(lf, np.float64).reshape(-1, 3)
            check(lib().t3d_problem_create_bc(nely, nelx, nelz, None if sm is None else _ptr(sm), len(fx),
                                              _ptr(fx), len(ln), _ptr(ln), _ptr(lf), C.byref(p)),
                  "t3d_problem_create_bc")
        elif solid is None:
            check(lib().t3d_problem_create_cuboid(nely, nelx, nelz, C.byref(p)), "t3d_problem_create_cuboid")
        else:
            sm = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
            if sm.size != nely * nelx * nelz:
                raise ValueError("solid mask needs nely*nelx*nelz entries")
            check(lib().t3d_problem_create_voxels(nely, nelx, nelz, _ptr(sm), C.byref(p)),
                  "t3d_problem_create_voxels")
        self.h = p
        s = (C.c_int64 * 10)()
        check(lib().t3d_problem_sizes(self.h, s), "t3d_problem_sizes")
        (self.resX, self.resY, self.resZ, self.ne, self.nN, self.nDOF, self.nF,
         self.nnx, self.nny, self.nnz) = [int(v) for v in s]

    @classmethod
    def from_handle(cls, h):
        """A view of a problem owned elsewhere (e.g. by a Batch); not destroyed by Python."""
        obj = cls.__new__(cls)
        obj.h = None
        obj._view = h
        s = (C.c_int64 * 10)()
        check(lib().t3d_problem_sizes(h, s), "t3d_problem_sizes")
        (obj.resX, obj.resY, obj.resZ, obj.ne, obj.nN, obj.nDOF, obj.nF,
         obj.nnx, obj.nny, obj.nnz) = [int(v) for v in s]
        obj.h = h
        obj._owned = False
        return obj

    def __del__(self):
        try:
            if self.h and getattr(self, "_owned", True):
                lib().t3d_problem_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def model(self):
        ne, nN, nF = self.ne, self.nN, self.nF
        out = dict(eleMapBack=np.zeros(ne, np.int32), eNodMat=np.zeros(8 * ne, np.int32),
                   color=np.zeros(ne, np.int32), Ke=np.zeros(576), F=np.zeros(3 * nN),
                   isFree=np.zeros(3 * nN, np.uint8), freeDofIndex=np.zeros(nF, np.int32),
                   eDofFree=np.zeros(24 * ne, np.int32))
        check(lib().t3d_problem_model(self.h, *[_ptr(out[k]) for k in
                                                ("eleMapBack", "eNodMat", "color", "Ke", "F", "isFree",
                                                 "freeDofIndex", "eDofFree")]), "t3d_problem_model")
        return out

    def ele_map_back(self):
        """eleMapBack[ne] alone (the full index of each compact element, fea.cpp:206-215)."""
        out = np.zeros(self.ne, np.int32)
        check(lib().t3d_problem_model(self.h, _ptr(out), None, None, None, None, None, None, None),
              "t3d_problem_model")
        return out

    def free_to_full(self, dfree, dfull):
        check(lib().t3d_free_to_full(self.h, dfree.ptr, dfull.ptr), "t3d_free_to_full")

    def full_to_free(self, dfull, dfree):
        check(lib().t3d_full_to_free(self.h, dfull.ptr, dfree.ptr), "t3d_full_to_free")

    def load_vector(self, dfull):
        check(lib().t3d_load_vector(self.h, dfull.ptr), "t3d_load_vector")

    def set_modulus(self, dE):
        check(lib().t3d_set_modulus(self.h, dE.ptr), "t3d_set_modulus")

    def kxu_f64(self, du, dy):
        check(lib().t3d_kxu_f64(self.h, du.ptr, dy.ptr), "t3d_kxu_f64")

    def kxu_free_host(self, E, uFree):
        E = np.ascontiguousarray(E, np.float64)
        u = np.ascontiguousarray(uFree, np.float64)
        y = np.zeros(self.nF)
        check(lib().t3d_kxu_free_host(self.h, _ptr(E), _ptr(u), _ptr(y)), "t3d_kxu_free_host")
        return y

    def dot(self, da, db, n=None):
        out = C.c_double()
        check(lib().t3d_dot_f64(self.h, da.ptr, db.ptr, n if n is not None else da.n, C.byref(out)), "t3d_dot_f64")
        return out.value

    def time_op(self, op, din, dout, reps):
        ms = C.c_double()
        check(lib().t3d_time_op(self.h, op, din.ptr, dout.ptr, reps, C.byref(ms)), "t3d_time_op")
        return ms.value


class MG:
    """build_static_once + make_diagonal_preconditioner_from_static
    (cpp/src/multigrid/preconditioner.cpp:19-266) on the device."""

    def __init__(self, prob, nonDyadic=1, maxLevels=5, weight=0.6, Emin=1e-6, NlimitDofs=1000, galerkin=False,
                 standard=False):
        """galerkin=True: the Galerkin hierarchy of the MATLAB original
        (matlab/TOP3D_XL.m:772-1090) on the same levels, t3d_mg_set_mode(1)."""
        self.prob = prob
        h = C.c_void_p()
        check(lib().t3d_mg_create(prob.h, nonDyadic, maxLevels, weight, Emin, NlimitDofs, C.byref(h)),
              "t3d_mg_create")
        self.h = h
        self.galerkin = bool(galerkin)
        if galerkin:
            check(lib().t3d_mg_set_mode(self.h, 1), "t3d_mg_set_mode")
            if standard:
                check(lib().t3d_mg_set_vcycle(self.h, 1), "t3d_mg_set_vcycle")
        nl = C.c_int()
        dims = (C.c_int * 32)()
        nc = C.c_int()
        ok = C.c_int()
        check(lib().t3d_mg_info(self.h, C.byref(nl), dims, C.byref(nc), C.byref(ok)), "t3d_mg_info")
        self.nlev = nl.value
        self.dims = [tuple(dims[4 * l:4 * l + 4]) for l in range(self.nlev)]
        self.nN = [(d[0] + 1) * (d[1] + 1) * (d[2] + 1) for d in self.dims]

    def __del__(self):
        try:
            if self.h:
                lib().t3d_mg_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        nl = C.c_int()
        nc = C.c_int()
        ok = C.c_int()
        check(lib().t3d_mg_info(self.h, C.byref(nl), None, C.byref(nc), C.byref(ok)), "t3d_mg_info")
        return nc.value, ok.value

    def setup(self, dE):
        check(lib().t3d_mg_setup(self.h, dE.ptr), "t3d_mg_setup")

    def apply(self, dr, dz):
        check(lib().t3d_mg_apply(self.h, dr.ptr, dz.ptr), "t3d_mg_apply")

    def mask(self, level):
        out = np.zeros(3 * self.nN[level], np.uint8)
        check(lib().t3d_mg_mask(self.h, level, _ptr(out)), "t3d_mg_mask")
        return out

    def diag(self, level):
        out = np.zeros(3 * self.nN[level])
        check(lib().t3d_mg_diag(self.h, level, _ptr(out)), "t3d_mg_diag")
        return out

    def coarsest(self):
        N = 3 * self.nN[-1]
        Kc = np.zeros(N * N)
        L = np.zeros(N * N)
        check(lib().t3d_mg_coarsest(self.h, _ptr(Kc), _ptr(L)), "t3d_mg_coarsest")
        return Kc.reshape(N, N), L.reshape(N, N)


class Filter:
    """SetupPDEFilter / ApplyPDEFilter (cpp/src/filter.cpp:9-275) on the device."""

    def __init__(self, prob, rMin):
        self.prob = prob
        h = C.c_void_p()
        check(lib().t3d_filter_create(prob.h, rMin, C.byref(h)), "t3d_filter_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().t3d_filter_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        k = np.zeros(64)
        d = np.zeros(self.prob.nN)
        check(lib().t3d_filter_info(self.h, _ptr(k), _ptr(d)), "t3d_filter_info")
        return k, d

    def apply(self, dsrc, ddst):
        it = C.c_int()
        check(lib().t3d_filter_apply(self.h, dsrc.ptr, ddst.ptr, C.byref(it)), "t3d_filter_apply")
        return it.value


def compliance(prob, dE, dU, dce):
    c = C.c_double()
    check(lib().t3d_compliance(prob.h, dE.ptr, dU.ptr, dce.ptr, C.byref(c)), "t3d_compliance")
    return c.value


def voxel_ke(nu=0.3, cellSize=1.0):
    """ComputeVoxelKe (cpp/src/fea.cpp:117-179): the 24x24 element matrix (host)."""
    Ke = np.zeros(576)
    check(lib().t3d_voxel_ke(nu, cellSize, Ke.ctypes.data), "t3d_voxel_ke")
    return Ke.reshape(24, 24)


def compliance_noalloc(prob, dE, dU, dce):
    """t3d_compliance_noalloc: the same result with no device allocation (static scratch)."""
    c = C.c_double()
    check(lib().t3d_compliance_noalloc(prob.h, dE.ptr, dU.ptr, dce.ptr, C.byref(c)), "t3d_compliance_noalloc")
    return c.value


def simp(prob, dx, E0, Emin, dE):
    check(lib().t3d_simp(prob.h, dx.ptr, E0, Emin, dE.ptr), "t3d_simp")


def oc_update(prob, dx, ddc, V0, move, dxnew, passive=None):
    """OC update (driver.cpp:169-189); passive: optional device uint8 mask in compact
    order, elements held at density 1 (matlab/TOP3D_XL.m:154)."""
    out = np.zeros(4, np.float32)
    check(lib().t3d_oc_update_passive(prob.h, dx.ptr, ddc.ptr, passive.ptr if passive is not None else None, V0,
                                      move, dxnew.ptr, _ptr(out)), "t3d_oc_update_passive")
    return dict(change=np.float32(out[0]), sharpness=np.float32(out[1]), volFrac=np.float32(out[2]),
                steps=int(out[3]))


def fp_probe(op, a, b=None):
    """Device sqrtf / float division / double sqrt of host arrays (diagnostics)."""
    da = DeviceArray.from_host(a)
    db = DeviceArray.from_host(b) if b is not None else None
    out = DeviceArray(a.size, a.dtype)
    check(lib().t3d_fp_probe(op, a.size, da.ptr, db.ptr if db is not None else None, out.ptr), "t3d_fp_probe")
    return out.download()


def seq_sum(prob, dv, n=None):
    out = C.c_float()
    check(lib().t3d_seq_sum_f32(prob.h, dv.ptr, dv.n if n is None else n, C.byref(out)), "t3d_seq_sum_f32")
    return np.float32(out.value)


class Params(C.Structure):
    _fields_ = [
        ("E0", C.c_float), ("Emin", C.c_float), ("nu", C.c_float), ("penal", C.c_float), ("cgTol", C.c_float),
        ("cgMaxIt", C.c_int), ("V0", C.c_float), ("nLoop", C.c_int), ("rMin", C.c_float),
        ("mgMaxLevels", C.c_int), ("mgWeight", C.c_float), ("nonDyadic", C.c_int), ("NlimitDofs", C.c_int),
        ("rebuildEvery", C.c_int), ("precision", C.c_int), ("mgMode", C.c_int),
        ("stlPath", C.c_char_p), ("stlIso", C.c_float), ("summaryPath", C.c_char_p),
        ("solid", C.c_void_p), ("passive", C.c_void_p), ("model", C.c_void_p),
    ]


class Summary(C.Structure):
    """t3d_summary: the values of the reference's run summary (driver.cpp:225-258)."""
    _fields_ = [
        ("origX", C.c_int), ("origY", C.c_int), ("origZ", C.c_int),
        ("resX", C.c_int), ("resY", C.c_int), ("resZ", C.c_int),
        ("V0", C.c_float), ("iterations", C.c_int), ("cgMaxIt", C.c_int), ("Csolid", C.c_double),
        ("objFirst", C.c_float), ("objLast", C.c_float),
        ("sumCG", C.c_float), ("sumOpt", C.c_float), ("sumFilter", C.c_float), ("sumIter", C.c_float),
        ("threads", C.c_int),
    ]


def summary_format(**kw):
    """The reference's summary text (driver.cpp:225-258) for the given t3d_summary fields."""
    S = Summary(**kw)
    buf = C.create_string_buffer(4096)
    n = C.c_size_t()
    check(lib().t3d_summary_format(C.byref(S), buf, 4096, C.byref(n)), "t3d_summary_format")
    return buf.raw[:n.value].decode()


class LoopRecord(C.Structure):
    _fields_ = [
        ("loop", C.c_int), ("pcgIters", C.c_int), ("filterIters", C.c_int), ("bisectSteps", C.c_int),
        ("C", C.c_double), ("Cdisp", C.c_double),
        ("volFrac", C.c_float), ("sharpness", C.c_float), ("change", C.c_float), ("fval", C.c_float),
        ("tSolve", C.c_double), ("tOpti", C.c_double), ("tFilter", C.c_double), ("tTotal", C.c_double),
        ("xCheck", C.c_uint64),
    ]


def default_params(**kw):
    p = Params()
    lib().t3d_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class VoxelModel:
    """A TopVoxel file (matlab/TOP3D_XL.m:458-552) parsed by t3d_topvoxel_read:
    nelx, nely, nelz, solid / passive masks and density (indexed [ez, ex, ey]),
    fix (int32 [n, 4] = x y z mask), loadNode (int32 [m, 3]), loadF ([m, 3]),
    extraLoadCases."""

    class _S(C.Structure):
        _fields_ = [("nelx", C.c_int), ("nely", C.c_int), ("nelz", C.c_int),
                    ("solid", C.POINTER(C.c_uint8)), ("passive", C.POINTER(C.c_uint8)),
                    ("density", C.POINTER(C.c_float)), ("nFix", C.c_int), ("fix", C.POINTER(C.c_int32)),
                    ("nLoad", C.c_int), ("loadNode", C.POINTER(C.c_int32)), ("loadF", C.POINTER(C.c_double)),
                    ("extraLoadCases", C.c_int)]

    def __init__(self, path):
        h = C.c_void_p()
        check(lib().t3d_topvoxel_read(os.fsencode(path), C.byref(h)), "t3d_topvoxel_read")
        self.h = h
        S = C.cast(h, C.POINTER(VoxelModel._S)).contents
        self.nelx, self.nely, self.nelz = S.nelx, S.nely, S.nelz
        nv = S.nelx * S.nely * S.nelz
        arr = lambda ptr, n, dt: np.ctypeslib.as_array(ptr, (n,)).astype(dt).copy() if n else np.zeros(0, dt)
        self.solid = arr(S.solid, nv, np.uint8)
        self.passive = arr(S.passive, nv, np.uint8)
        self.density = arr(S.density, nv, np.float32) if bool(S.density) else None
        self.fix = arr(S.fix, 4 * S.nFix, np.int32).reshape(-1, 4)
        self.loadNode = arr(S.loadNode, 3 * S.nLoad, np.int32).reshape(-1, 3)
        self.loadF = arr(S.loadF, 3 * S.nLoad, np.float64).reshape(-1, 3)
        self.extraLoadCases = S.extraLoadCases

    def __del__(self):
        try:
            if self.h:
                lib().t3d_voxel_model_destroy(self.h)
                self.h = None
        except Exception:
            pass


def top3d_run(nely, nelx, nelz, solid=None, passive=None, model=None, **kw):
    """TOP3D_XL_GLOBAL (cpp/src/driver.cpp:23-218) for one simulation, on the device.
    solid / passive: optional uint8 voxel masks indexed [ez, ex, ey] (a voxel domain;
    voxels held at density 1); model: a VoxelModel (its domain, boundary conditions,
    load and passive elements).  Returns (records, x, Csolid, solidIters)."""
    P = default_params(**kw)
    if model is not None:
        P.model = model.h
        if solid is None:
            solid = model.solid
    keep = []
    for name, mask in (("solid", solid), ("passive", passive)):
        if mask is not None:
            a = np.ascontiguousarray(mask, np.uint8).reshape(-1)
            if a.size != nely * nelx * nelz:
                raise ValueError("%s mask needs nely*nelx*nelz entries" % name)
            keep.append(a)
            setattr(P, name, a.ctypes.data)
    rec = (LoopRecord * max(1, P.nLoop))()
    # the design has numElements entries: every voxel of the cuboid, or the solid ones
    ne = nelx * nely * nelz if solid is None else int(np.count_nonzero(keep[0]))
    x = np.zeros(ne, np.float32)
    Cs = C.c_double()
    si = C.c_int()
    nl = C.c_int()
    check(lib().t3d_top3d_run(nely, nelx, nelz, C.byref(P), rec, P.nLoop, _ptr(x), C.byref(Cs), C.byref(si),
                              C.byref(nl)), "t3d_top3d_run")
    recs = [{f: getattr(r, f) for f, _ in LoopRecord._fields_} for r in rec[:nl.value]]
    return recs, x, Cs.value, si.value


def stl_size(prob, xPhys, iso=0.5):
    """(bytes, triangles) of the voxel-surface STL of the device density xPhys
    (compact order, float32) at threshold iso (io.cpp:80-87)."""
    nb, nt = C.c_uint64(), C.c_uint64()
    check(lib().t3d_stl_size(prob.h, xPhys.ptr, float(iso), C.byref(nb), C.byref(nt)), "t3d_stl_size")
    return nb.value, nt.value


def stl_image(prob, xPhys, iso=0.5):
    """The binary STL file of export_surface_stl (io.cpp:80-87), built on the
    device, as bytes."""
    nbytes, _ = stl_size(prob, xPhys, iso)
    img = DeviceArray((nbytes + 3) // 4, np.uint32)
    check(lib().t3d_stl_image(prob.h, xPhys.ptr, float(iso), img.ptr, nbytes), "t3d_stl_image")
    return img.download().view(np.uint8)[:nbytes].tobytes()


def export_surface_stl(prob, xPhys, path, iso=0.5):
    """export_surface_stl (io.cpp:80-87): the device-built image written to path."""
    check(lib().t3d_export_surface_stl(prob.h, xPhys.ptr, os.fsencode(path), float(iso)),
          "t3d_export_surface_stl")


class Mixed:
    """Mixed-precision MG-PCG state: FP32 K*u and V-cycle, FP64 outer loop."""

    def __init__(self, prob, M):
        self.prob, self.M = prob, M
        h = C.c_void_p()
        check(lib().t3d_mixed_create(prob.h, M.h, C.byref(h)), "t3d_mixed_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().t3d_mixed_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def setup(self):
        check(lib().t3d_mixed_setup(self.h), "t3d_mixed_setup")

    def vcycle(self, dr, dz):
        """z = M32(r): one FP32 V-cycle of the multigrid object's cycle type (float32 device arrays)."""
        check(lib().t3d_mixed_vcycle(self.h, dr.ptr, dz.ptr), "t3d_mixed_vcycle")

    def true_residual(self, db, dx, dr):
        """r = b - K x with the FP64 operator of the residual replacement, into the
        float32 device array dr; returns the FP64 r.r."""
        rn2 = C.c_double()
        check(lib().t3d_mixed_true_residual(self.h, db.ptr, dx.ptr, dr.ptr, C.byref(rn2)), "t3d_mixed_true_residual")
        return rn2.value

    def coarse_solve(self, db, dx):
        """Coarsest-level solve of the mixed V-cycle (inverse Cholesky factor), FP64 device arrays."""
        check(lib().t3d_mixed_coarse_solve(self.h, db.ptr, dx.ptr), "t3d_mixed_coarse_solve")

    def set_operator(self, form):
        """Operator of the solver: 0 Walsh form, 2 MFMA bf16 x3, 3 MFMA bf16 x6."""
        check(lib().t3d_mixed_set_operator(self.h, form), "t3d_mixed_set_operator")

    def kxu_f32(self, du, dy, form=0):
        """y = K u in FP32 (device float arrays, full layout); form 0 Walsh, 1 node gather,
        2 MFMA bf16 x3, 3 MFMA bf16 x6."""
        check(lib().t3d_kxu_f32(self.h, form, du.ptr, dy.ptr), "t3d_kxu_f32")

    def time_kxu(self, du, dy, reps, form=0):
        ms = C.c_float()
        check(lib().t3d_mixed_time_op(self.h, form, du.ptr, dy.ptr, reps, C.byref(ms)), "t3d_mixed_time_op")
        return ms.value

    def set_timing(self, on=True):
        """HIP-event timing of the fused operator kernel inside pcg() (totals zeroed)."""
        check(lib().t3d_mixed_set_timing(self.h, int(on)), "t3d_mixed_set_timing")

    def timing(self):
        """(milliseconds, launches) of the fused operator kernel since set_timing."""
        ms, n = C.c_double(), C.c_longlong()
        check(lib().t3d_mixed_get_timing(self.h, C.byref(ms), C.byref(n)), "t3d_mixed_get_timing")
        return ms.value, n.value

    def pcg(self, db, dx, tol, maxIt, trueEvery=50, flexible=True):
        it = C.c_int()
        check(lib().t3d_pcg_mixed(self.h, db.ptr, dx.ptr, tol, maxIt, trueEvery, int(flexible), C.byref(it)),
              "t3d_pcg_mixed")
        return it.value


class Batch:
    """Many problems solved together by the batched mixed-precision MG-PCG
    (t3d_batch_*): dims = [(nely, nelx, nelz), ...]."""

    def __init__(self, dims, Emin=1e-9, solids=None, galerkin=False):
        """solids: optional list of voxel masks (as Problem's solid; None entries = cuboid).
        galerkin=True: every problem's multigrid in the Galerkin form (MG(galerkin=True))."""
        n = len(dims)
        ny = np.array([d[0] for d in dims], np.int32)
        nx = np.array([d[1] for d in dims], np.int32)
        nz = np.array([d[2] for d in dims], np.int32)
        h = C.c_void_p()
        if solids is None:
            check(lib().t3d_batch_create(n, _ptr(ny), _ptr(nx), _ptr(nz), Emin, C.byref(h)), "t3d_batch_create")
        else:
            keep = [None if m is None else np.ascontiguousarray(m, dtype=np.uint8).reshape(-1) for m in solids]
            for d, m in zip(dims, keep):
                if m is not None and m.size != d[0] * d[1] * d[2]:
                    raise ValueError("solid mask needs nely*nelx*nelz entries")
            ptrs = C.cast((C.c_void_p * n)(*[None if m is None else m.ctypes.data for m in keep]), C.c_void_p)
            check(lib().t3d_batch_create_voxels(n, _ptr(ny), _ptr(nx), _ptr(nz), ptrs, Emin, C.byref(h)),
                  "t3d_batch_create_voxels")
        self.h = h
        self.n = n
        self.problems = []
        for i in range(n):
            ph = C.c_void_p()
            check(lib().t3d_batch_problem(self.h, i, C.byref(ph)), "t3d_batch_problem")
            self.problems.append(Problem.from_handle(ph))
            if galerkin:
                gh = C.c_void_p()
                check(lib().t3d_batch_mg(self.h, i, C.byref(gh)), "t3d_batch_mg")
                check(lib().t3d_mg_set_mode(gh, 1), "t3d_mg_set_mode")
        self.galerkin = bool(galerkin)

    def __del__(self):
        try:
            if self.h:
                lib().t3d_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def setup(self, dEs):
        ptrs = (C.c_void_p * self.n)(*[d.ptr.value for d in dEs])
        check(lib().t3d_batch_setup(self.h, ptrs), "t3d_batch_setup")

    def set_tiles(self, kwNZ, rowsPerBlock):
        check(lib().t3d_batch_set_tiles(self.h, kwNZ, rowsPerBlock), "t3d_batch_set_tiles")

    def info(self):
        a, b = C.c_int(), C.c_int()
        check(lib().t3d_batch_info(self.h, C.byref(a), C.byref(b)), "t3d_batch_info")
        return a.value, b.value

    def set_timing(self, on=True):
        """HIP-event timing of the next solves (totals zeroed): per kernel of the
        directly launched iterations (useGraph=False), per graph replay (useGraph=True)."""
        check(lib().t3d_batch_set_timing(self.h, int(on)), "t3d_batch_set_timing")

    def timing(self):
        """{name: (total ms, launches)} of the kernels timed, plus "graph" for the replays."""
        buf = C.create_string_buffer(64)
        ms, n = C.c_double(), C.c_longlong()
        nk = lib().t3d_batch_get_timing(self.h, -1, buf, 64, C.byref(ms), C.byref(n))
        if nk < 0:
            check(nk, "t3d_batch_get_timing")
        out = {"graph": (ms.value, n.value)}
        for k in range(nk):
            check(0 if lib().t3d_batch_get_timing(self.h, k, buf, 64, C.byref(ms), C.byref(n)) >= 0 else -1,
                  "t3d_batch_get_timing")
            out[buf.value.decode()] = (ms.value, n.value)
        return out

    def solve(self, dbs, dxs, tol, maxIt, trueEvery=50, flexible=True, useGraph=True):
        bp = (C.c_void_p * self.n)(*[d.ptr.value for d in dbs])
        xp = (C.c_void_p * self.n)(*[d.ptr.value for d in dxs])
        it = np.zeros(self.n, np.int32)
        check(lib().t3d_batch_solve(self.h, bp, xp, tol, maxIt, trueEvery, int(flexible), int(useGraph), _ptr(it)),
              "t3d_batch_solve")
        return it


def pcg(prob, M, db, dx, tol, maxIt):
    """PCG_free (cpp/src/solver.cpp:87-217) on device vectors (full layout); returns iterations."""
    it = C.c_int()
    check(lib().t3d_pcg(prob.h, M.h if M is not None else None, db.ptr, dx.ptr, tol, maxIt, C.byref(it)), "t3d_pcg")
    return it.value


def device_info():
    name = C.create_string_buffer(64)
    cus = C.c_int()
    mem = C.c_size_t()
    check(lib().t3d_device_info(name, 64, C.byref(cus), C.byref(mem)), "t3d_device_info")
    return name.value.decode(), cus.value, mem.value
