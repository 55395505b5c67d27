"""ctypes bindings of the reference's C entry points.

`MeshLib` wraps any shared library that exports the reference's prototypes
(src/meshify.h:9 meshify, src/quadric.h:7 laplacian_smoothHC, src/bwlabel.h:10
bwlabel, src/oldcubes.h:9 marchingCubes, src/isolevel.h:10 setThreshold), with
an optional symbol prefix.  The product library (`nii2mesh_amd.load()`) is one
such library; the parity tests bind other implementations of the same
prototypes through the same class, so every caller passes identical arguments.

Arrays cross the boundary as numpy arrays: volumes float32 x-fastest
(shape (nz, ny, nx) in C order), triangles int32 (nt, 3), vertices float64
(np, 3) -- the vec3i / vec3d layouts of src/meshtypes.h.
"""
import ctypes as C

import numpy as np

_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags='C_CONTIGUOUS')
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')


def _take(ptr, n, dtype, cols=3):
    """Copy n rows out of a malloc'd C array and free it."""
    if n <= 0 or not ptr:
        if ptr:
            _libc.free(ptr)
        return np.zeros((0, cols), dtype=dtype)
    ct = C.c_double if dtype == np.float64 else C.c_int32
    buf = (ct * (n * cols)).from_address(ptr)
    out = np.frombuffer(buf, dtype=dtype).reshape(n, cols).copy()
    _libc.free(ptr)
    return out


def _dims(img):
    if img.ndim != 3:
        raise ValueError('volume must be 3-D (nz, ny, nx)')
    nz, ny, nx = img.shape
    return nx, ny, nz


class MeshLib:
    """The reference's entry points bound from `lib` (a ctypes.CDLL)."""

    def __init__(self, lib, prefix='', cubes_arg=False):
        self.lib = lib
        self.cubes_arg = cubes_arg
        p = prefix
        f = getattr(lib, p + 'meshify')
        args = [_f32p, C.POINTER(C.c_short)]
        if cubes_arg:
            args.append(C.c_int)
        args += [C.c_int, C.c_float, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_bool, C.c_bool, C.c_bool]
        if not cubes_arg:
            args.append(C.c_bool)
        f.argtypes = args
        f.restype = C.c_int
        self._meshify = f
        f = getattr(lib, p + 'laplacian_smoothHC')
        f.argtypes = [_f64p, _i32p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_bool]
        f.restype = None
        self._hc = f
        f = getattr(lib, p + 'bwlabel')
        f.argtypes = [_f32p, C.c_int, C.POINTER(C.c_size_t), C.c_bool, C.c_bool]
        f.restype = C.c_int
        self._bwlabel = f
        name = p + ('set_threshold' if cubes_arg else 'setThreshold')
        f = getattr(lib, name)
        f.argtypes = [_f32p, C.c_int, C.c_int]
        f.restype = C.c_float
        self._thr = f

    # -- src/meshify.h:9 ---------------------------------------------------
    def meshify(self, img, isolevel, preSmooth=True, onlyLargest=True, fillBubbles=False,
                originalMC=0, cubes=0, verbose=False):
        """Returns (rc, tris int32 (nt,3), pts float64 (np,3)); img is mutated."""
        nx, ny, nz = _dims(img)
        dim = (C.c_short * 3)(nx, ny, nz)
        t, p = C.c_void_p(), C.c_void_p()
        nt, npt = C.c_int(0), C.c_int(0)
        if self.cubes_arg:
            rc = self._meshify(img, dim, int(cubes), int(originalMC), float(isolevel), C.byref(t),
                               C.byref(p), C.byref(nt), C.byref(npt), bool(preSmooth),
                               bool(onlyLargest), bool(fillBubbles))
        else:
            rc = self._meshify(img, dim, int(originalMC), float(isolevel), C.byref(t), C.byref(p),
                               C.byref(nt), C.byref(npt), bool(preSmooth), bool(onlyLargest),
                               bool(fillBubbles), bool(verbose))
        if rc != 0:
            return rc, np.zeros((0, 3), np.int32), np.zeros((0, 3), np.float64)
        return rc, _take(t.value, nt.value, np.int32), _take(p.value, npt.value, np.float64)

    # -- src/quadric.h:7 ---------------------------------------------------
    def laplacian_smoothHC(self, pts, tris, alpha=0.1, beta=0.5, iters=100, lockEdges=True):
        """HC smoothing in place on a float64 (np,3) array."""
        assert pts.dtype == np.float64 and pts.flags.c_contiguous
        tris = np.ascontiguousarray(tris, dtype=np.int32)
        self._hc(pts, tris, len(pts), len(tris), float(alpha), float(beta), int(iters),
                 bool(lockEdges))
        return pts

    # -- src/bwlabel.h:10 --------------------------------------------------
    def bwlabel(self, img, conn=18, onlyLargest=False, fillBubbles=False):
        nx, ny, nz = _dims(img)
        dim = (C.c_size_t * 3)(nx, ny, nz)
        return self._bwlabel(img, int(conn), dim, bool(onlyLargest), bool(fillBubbles))

    # -- src/isolevel.h:10 -------------------------------------------------
    def setThreshold(self, img, mode):
        return float(self._thr(np.ascontiguousarray(img, dtype=np.float32).ravel(), img.size,
                               int(mode)))


def bind_marching_cubes(lib, name='marchingCubes', cubes_first=None):
    """Bind a marchingCubes(img, dim, lo, hi, originalMC, iso, &vs, &ts, &nv, &nt)."""
    f = getattr(lib, name)
    f.argtypes = [_f32p, C.POINTER(C.c_short), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                  C.c_float, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                  C.POINTER(C.c_int)]
    f.restype = C.c_int

    def call(img, lo, hi, isolevel, originalMC=0):
        nx, ny, nz = _dims(img)
        dim = (C.c_short * 3)(nx, ny, nz)
        clo = (C.c_int * 3)(*lo)
        chi = (C.c_int * 3)(*hi)
        vs, ts = C.c_void_p(), C.c_void_p()
        nv, nt = C.c_int(0), C.c_int(0)
        rc = f(img, dim, clo, chi, int(originalMC), float(isolevel), C.byref(vs), C.byref(ts),
               C.byref(nv), C.byref(nt))
        if rc != 0:
            return rc, np.zeros((0, 3), np.int32), np.zeros((0, 3), np.float64)
        return rc, _take(ts.value, nt.value, np.int32), _take(vs.value, nv.value, np.float64)
    return call


def bind_unify(lib, name):
    """Bind int f(vec3d **pts, vec3i *tris, int npt, int ntri)."""
    f = getattr(lib, name)
    f.argtypes = [C.POINTER(C.c_void_p), _i32p, C.c_int, C.c_int]
    f.restype = C.c_int

    def call(pts, tris):
        tris = np.ascontiguousarray(tris, dtype=np.int32).copy()
        n = len(pts)
        buf = _libc_malloc(n * 24)
        C.memmove(buf, np.ascontiguousarray(pts, dtype=np.float64).ctypes.data, n * 24)
        p = C.c_void_p(buf)
        nn = f(C.byref(p), tris, n, len(tris))
        return _take(p.value, nn, np.float64), tris
    return call


def bind_degenerate(lib, name):
    """Bind int f(vec3d *pts, vec3i **tris, int ntri)."""
    f = getattr(lib, name)
    f.argtypes = [_f64p, C.POINTER(C.c_void_p), C.c_int]
    f.restype = C.c_int

    def call(pts, tris):
        n = len(tris)
        buf = _libc_malloc(max(n, 1) * 12)
        C.memmove(buf, np.ascontiguousarray(tris, dtype=np.int32).ctypes.data, n * 12)
        t = C.c_void_p(buf)
        nn = f(np.ascontiguousarray(pts, dtype=np.float64), C.byref(t), n)
        return _take(t.value, nn, np.int32)
    return call


def bind_quadric(lib, name='quadric_simplify_mesh'):
    """Bind void quadric_simplify_mesh(vec3d **vs, vec3i **ts, int *nvert,
    int *ntri, int target_count, double agressiveness, bool verbose,
    bool finishLossless) (src/quadric.h:6): the call takes malloc'd arrays
    and replaces them; returns copies of the simplified (tris, pts)."""
    f = getattr(lib, name)
    f.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                  C.c_double, C.c_bool, C.c_bool]
    f.restype = None

    def call(pts, tris, target_count, agressiveness=7.0, verbose=False, finishLossless=False):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        tris = np.ascontiguousarray(tris, dtype=np.int32)
        pv = _libc_malloc(max(len(pts), 1) * 24)
        tv = _libc_malloc(max(len(tris), 1) * 12)
        C.memmove(pv, pts.ctypes.data, len(pts) * 24)
        C.memmove(tv, tris.ctypes.data, len(tris) * 12)
        p, t = C.c_void_p(pv), C.c_void_p(tv)
        nv, nt = C.c_int(len(pts)), C.c_int(len(tris))
        f(C.byref(p), C.byref(t), C.byref(nv), C.byref(nt), int(target_count), float(agressiveness), bool(verbose),
          bool(finishLossless))
        return _take(t.value, nt.value, np.int32), _take(p.value, nv.value, np.float64)
    return call


_libc.malloc.argtypes = [C.c_size_t]
_libc.malloc.restype = C.c_void_p


def _libc_malloc(n):
    p = _libc.malloc(max(int(n), 1))
    if not p:
        raise MemoryError
    return p
