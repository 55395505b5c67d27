"""nii2mesh_amd -- MI355X (gfx950) voxel-to-mesh hot path of nii2mesh.

The product is the C-ABI shared library ``lib/libnii2mesh_amd.so`` (HIP
kernels + host orchestration, see include/nii2mesh_amd.h).  This package is
the Python side of that boundary:

* `load()`      -- the ctypes handle; raises if the library is missing (there
                   is no CPU fallback);
* `lib()`       -- the reference's entry points (meshify, laplacian_smoothHC,
                   bwlabel, setThreshold, marchingCubes, ...) bound through
                   `abi.MeshLib`;
* `DeviceContext` -- the device-resident API (volume already in HBM), used by
                   bench.py;
* `synth`       -- the seeded synthetic workloads.
"""
import ctypes as C
import os
import threading

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'lib', 'libnii2mesh_amd.so')
_lib = None


def load():
    """ctypes handle of the HIP library (built by nii2mesh_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('nii2mesh_amd: %s is missing; run python -m nii2mesh_amd.build' % LIB_PATH)
        # torch ships its own libamdhip64.so.7 and HSA runtime.  Loaded first,
        # it also serves this library (same soname); loaded after a ROCm copy,
        # torch would bring a second runtime that cannot open the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(LIB_PATH)
        _bind_ext(_lib)
    return _lib


class _Params(C.Structure):
    _fields_ = [('cubes', C.c_int), ('originalMC', C.c_int), ('isolevel', C.c_float),
                ('preSmooth', C.c_int), ('onlyLargest', C.c_int), ('fillBubbles', C.c_int),
                ('verbose', C.c_int)]


class AtlasEntry(C.Structure):
    _fields_ = [('label', C.c_int), ('status', C.c_int), ('nt', C.c_int), ('np', C.c_int),
                ('tri_offset', C.c_int), ('vert_offset', C.c_int)]


ATLAS_OK, ATLAS_EMPTY, ATLAS_FAILED = 0, 1, 2


class _AtlasMesh(C.Structure):
    _fields_ = [('label', C.c_int), ('status', C.c_int), ('nt', C.c_int), ('np', C.c_int),
                ('t', C.c_void_p), ('p', C.c_void_p)]


def _bind_ext(L):
    L.n2m_set_cubes.argtypes = [C.c_int]
    L.n2m_set_cubes.restype = C.c_int
    L.n2m_get_cubes.restype = C.c_int
    L.n2m_last_error.restype = C.c_char_p
    L.n2m_device_ok.restype = C.c_int
    L.n2m_ctx_create.restype = C.c_void_p
    L.n2m_ctx_destroy.argtypes = [C.c_void_p]
    L.n2m_ctx_stream.argtypes = [C.c_void_p]
    L.n2m_ctx_stream.restype = C.c_void_p
    L.n2m_ctx_meshify.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(_Params)]
    L.n2m_ctx_pipeline.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(_Params), C.c_int,
                                   C.POINTER(C.c_float), C.c_double, C.c_double, C.c_int]
    L.n2m_ctx_set_threshold.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
    L.n2m_ctx_set_threshold.restype = C.c_float
    L.n2m_ctx_smooth_hc.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]
    L.n2m_ctx_mesh_count.argtypes = [C.c_void_p]
    L.n2m_ctx_mesh_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.n2m_ctx_download.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.n2m_ctx_enable_timing.argtypes = [C.c_void_p, C.c_int]
    L.n2m_ctx_reset_timing.argtypes = [C.c_void_p]
    L.n2m_ctx_timing_only.argtypes = [C.c_void_p, C.c_char_p]
    L.n2m_ctx_kernel_stat.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.n2m_atlas_tuned.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.n2m_quick_smooth.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                   C.c_int]
    L.n2m_laplacian_smooth.argtypes = [np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'),
                                       np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                       C.c_int]
    L.n2m_laplacian_smooth.restype = None
    f32 = C.POINTER(C.c_float)
    L.n2m_apply_sform.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'),
                                  np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                  f32, f32, f32]
    L.n2m_apply_sform.restype = None
    for fn in (L.n2m_ctx_atlas, L.n2m_ctx_atlas_loop):
        fn.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(_Params), C.c_int, f32]
    L.n2m_ctx_diff_forward.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_float, C.c_int, C.c_int,
                                       C.c_int]
    L.n2m_ctx_diff_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.n2m_laplacian_smooth_T.argtypes = [np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'),
                                         np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                         C.c_int]
    L.n2m_laplacian_smooth_T.restype = None
    L.n2m_decode.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_float,
                             np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS')]
    L.n2m_save_mesh.argtypes = [C.c_char_p, np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'),
                                np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'), C.c_int, C.c_int, C.c_bool]
    L.n2m_otsu.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                           C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.n2m_atlas.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.POINTER(C.c_short), C.c_int,
                            C.c_bool, C.c_bool, C.c_int, f32, C.c_bool, C.POINTER(C.POINTER(_AtlasMesh)),
                            C.POINTER(C.c_int)]
    L.n2m_atlas_free.argtypes = [C.POINTER(_AtlasMesh), C.c_int]
    L.n2m_atlas_free.restype = None
    L.n2m_ctx_atlas_count.argtypes = [C.c_void_p]
    L.n2m_ctx_atlas_entry.argtypes = [C.c_void_p, C.c_int, C.POINTER(AtlasEntry)]
    L.n2m_ctx_atlas_download.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]


def lib(cubes=None):
    """The reference entry points of the HIP library (abi.MeshLib)."""
    L = load()
    if cubes is not None:
        L.n2m_set_cubes(int(cubes))
    return abi.MeshLib(L)


def laplacian_smooth_T(g, tris, iters=1):
    """Transpose of laplacian_smooth (src/quadric.c:316-341), iters times,
    in place on a float64 (np, 3) array."""
    L = load()
    L.n2m_laplacian_smooth_T(g, np.ascontiguousarray(tris, np.int32), len(g), len(tris), int(iters))
    return g


DT_UINT8, DT_INT16, DT_FLOAT32, DT_UINT16 = 2, 4, 16, 512


def decode(raw, datatype, slope=1.0, inter=0.0):
    """Raw NIfTI voxels -> float32 on the GPU (src/nii2mesh.c:155-172)."""
    L = load()
    raw = np.ascontiguousarray(raw)
    out = np.empty(raw.size, np.float32)
    if L.n2m_decode(raw.ctypes.data, int(datatype), raw.size, float(slope), float(inter), out) != 0:
        raise RuntimeError(last_error())
    return out.reshape(raw.shape)


def save_mesh(fnm, tris, pts, gz=False):
    """save_mesh() (src/meshify.c:984-1019) for every format of the
    reference: the binary byte streams (.mz3, .gii, .ply, .stl, .vtk,
    FreeSurfer) encoded on the GPU, the text formats (.obj, .off, .json)
    formatted by the host C library; returns 0 on success."""
    L = load()
    return L.n2m_save_mesh(str(fnm).encode(), np.ascontiguousarray(tris, np.int32),
                           np.ascontiguousarray(pts, np.float64), len(tris), len(pts), bool(gz))


def atlas(img, originalMC=0, preSmooth=True, fillBubbles=False, post_smooth=0, srow=None, verbose=False):
    """Host-memory atlas entry (n2m_atlas): returns (rc, [(label, status,
    tris, pts)] for labels 1..n) with copies of the library's arrays."""
    L = load()
    img = np.ascontiguousarray(img, np.float32)
    nz, ny, nx = img.shape
    sr = None if srow is None else (C.c_float * 12)(*[float(v) for v in np.asarray(srow).ravel()])
    out = C.POINTER(_AtlasMesh)()
    n = C.c_int(0)
    rc = L.n2m_atlas(img, (C.c_short * 3)(nx, ny, nz), int(originalMC), bool(preSmooth), bool(fillBubbles),
                     int(post_smooth), sr, bool(verbose), C.byref(out), C.byref(n))
    res = []
    for k in range(n.value):
        m = out[k]
        t = p = None
        if m.status == ATLAS_OK:
            t = np.ctypeslib.as_array((C.c_int32 * (3 * m.nt)).from_address(m.t)).reshape(-1, 3).copy() \
                if m.nt else np.zeros((0, 3), np.int32)
            p = np.ctypeslib.as_array((C.c_double * (3 * m.np)).from_address(m.p)).reshape(-1, 3).copy()
        res.append((m.label, m.status, t, p))
    if n.value:
        L.n2m_atlas_free(out, n)
    return rc, res


def otsu(hist, mode):
    """nii_otsu (src/isolevel.c:141-243): (threshold, dark, mid, bright)."""
    L = load()
    d, m, b = C.c_int(), C.c_int(), C.c_int()
    t = L.n2m_otsu(np.ascontiguousarray(hist, np.int32), len(hist), int(mode), C.byref(d), C.byref(m), C.byref(b))
    return t, d.value, m.value, b.value


def last_error():
    return load().n2m_last_error().decode()


def atlas_tuned(device=0):
    """The batched atlas's tuned launch shapes on `device`: {stage: (tile,
    blocks_per_cu)}, -1 for a stage not tuned yet (tile -1: the stage has no
    tile choice; blocks per CU 0: one workgroup per work item)."""
    L = load()
    out = {}
    name, tile, bpc = C.c_char_p(), C.c_int(), C.c_int()
    for k in range(L.n2m_atlas_tuned(device, -1, None, None, None)):
        L.n2m_atlas_tuned(device, k, C.byref(name), C.byref(tile), C.byref(bpc))
        out[name.value.decode()] = (tile.value, bpc.value)
    return out


CUBES_CLASSIC = 0
CUBES_LEWINER = 1


_orphans = []  # contexts dropped by the garbage collector on another thread


def _close_orphans():
    while _orphans:
        load().n2m_ctx_destroy(_orphans.pop())


class DeviceContext:
    """Device-resident pipeline (one HIP stream, reusable HBM buffers).

    The library runs on the context's own stream (created non-blocking), so
    work another stream queued on an input array must be complete before a
    call (torch: torch.cuda.synchronize()).  Close a context explicitly
    (close() or a with-block).  A context the garbage collector drops is
    destroyed only on the thread that created it; dropped elsewhere (for
    example on torch's autograd thread), its handle is kept and destroyed by
    the next create() or close() on any thread."""

    def __init__(self):
        self.L = load()
        _close_orphans()
        self.h = self.L.n2m_ctx_create()
        self._owner = threading.get_ident()
        if not self.h:
            raise RuntimeError('n2m_ctx_create failed: ' + last_error())

    def close(self):
        _close_orphans()
        if self.h:
            self.L.n2m_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        h = getattr(self, 'h', None)
        if not h:
            return
        self.h = None
        if threading is not None and threading.get_ident() == getattr(self, '_owner', None):
            try:
                self.L.n2m_ctx_destroy(h)
            except Exception:
                pass
        else:
            _orphans.append(h)

    @staticmethod
    def params(isolevel, cubes=CUBES_CLASSIC, originalMC=0, preSmooth=True, onlyLargest=True,
               fillBubbles=False, verbose=False):
        return _Params(int(cubes), int(originalMC), float(isolevel), int(preSmooth), int(onlyLargest),
                       int(fillBubbles), int(verbose))

    @property
    def stream(self):
        return self.L.n2m_ctx_stream(self.h)

    def meshify(self, d_ptr, dims, prm):
        d = (C.c_int * 3)(*dims)
        return self.L.n2m_ctx_meshify(self.h, C.c_void_p(d_ptr), d, C.byref(prm))

    def pipeline(self, d_ptr, dims, prm, isolevels, alpha=0.1, beta=0.5, hc_iters=100):
        d = (C.c_int * 3)(*dims)
        iso = (C.c_float * len(isolevels))(*isolevels)
        return self.L.n2m_ctx_pipeline(self.h, C.c_void_p(d_ptr), d, C.byref(prm), len(isolevels), iso,
                                       float(alpha), float(beta), int(hc_iters))

    def atlas(self, d_ptr, dims, prm, post_smooth=0, srow=None, loop=False):
        """Mesh every label of a device label volume (float32, x fastest).
        loop=True runs the per-label loop instead of the batched path."""
        d = (C.c_int * 3)(*dims)
        sr = None if srow is None else (C.c_float * 12)(*[float(v) for v in np.asarray(srow).ravel()])
        fn = self.L.n2m_ctx_atlas_loop if loop else self.L.n2m_ctx_atlas
        return fn(self.h, C.c_void_p(d_ptr), d, C.byref(prm), int(post_smooth), sr)

    def atlas_entries(self):
        out = []
        for k in range(self.L.n2m_ctx_atlas_count(self.h)):
            e = AtlasEntry()
            self.L.n2m_ctx_atlas_entry(self.h, k, C.byref(e))
            out.append(e)
        return out

    def atlas_mesh(self, k):
        """(tris, pts) of atlas entry k (label k+1), label-local vertex ids."""
        e = AtlasEntry()
        if self.L.n2m_ctx_atlas_entry(self.h, k, C.byref(e)) != 0:
            raise IndexError(k)
        t = np.empty((e.nt, 3), np.int32)
        p = np.empty((e.np, 3), np.float64)
        if self.L.n2m_ctx_atlas_download(self.h, k, t.ctypes.data, p.ctypes.data) != 0:
            raise RuntimeError(last_error())
        return t, p

    def diff_forward(self, d_ptr, dims, isolevel=0.0, originalMC=0, lap_iter=20, exact=False):
        """Differentiable forward on a device volume; the mesh is mesh 0."""
        d = (C.c_int * 3)(*dims)
        return self.L.n2m_ctx_diff_forward(self.h, C.c_void_p(d_ptr), d, float(isolevel), int(originalMC),
                                           int(lap_iter), int(exact))

    def diff_backward(self, d_dLdp, d_dLdvol):
        """dL/dvolume (float64 device array) from dL/dp (float64 device np x 3)."""
        return self.L.n2m_ctx_diff_backward(self.h, C.c_void_p(d_dLdp), C.c_void_p(d_dLdvol))

    def set_threshold(self, d_ptr, nvox, mode):
        return float(self.L.n2m_ctx_set_threshold(self.h, C.c_void_p(d_ptr), int(nvox), int(mode)))

    def smooth_hc(self, mesh=0, alpha=0.1, beta=0.5, iters=100, lock_edges=True):
        return self.L.n2m_ctx_smooth_hc(self.h, mesh, alpha, beta, iters, int(lock_edges))

    def mesh(self, k=0):
        nt, npt = C.c_int(0), C.c_int(0)
        if self.L.n2m_ctx_mesh_size(self.h, k, C.byref(nt), C.byref(npt)) != 0:
            raise IndexError(k)
        t = np.empty((nt.value, 3), np.int32)
        p = np.empty((npt.value, 3), np.float64)
        if self.L.n2m_ctx_download(self.h, k, t.ctypes.data, p.ctypes.data) != 0:
            raise RuntimeError(last_error())
        return t, p

    def mesh_size(self, k=0):
        nt, npt = C.c_int(0), C.c_int(0)
        self.L.n2m_ctx_mesh_size(self.h, k, C.byref(nt), C.byref(npt))
        return nt.value, npt.value

    def timing(self, on=True):
        self.L.n2m_ctx_enable_timing(self.h, int(on))

    def timing_only(self, name=None):
        self.L.n2m_ctx_timing_only(self.h, None if name is None else name.encode())

    def reset_timing(self):
        self.L.n2m_ctx_reset_timing(self.h)

    def kernel_stats(self):
        """{name: (ms_total, launches, algorithmic_bytes_total)}"""
        out = {}
        name = C.c_char_p()
        ms, n, b = C.c_double(), C.c_int(), C.c_double()
        cnt = self.L.n2m_ctx_kernel_stat(self.h, -1, C.byref(name), C.byref(ms), C.byref(n), C.byref(b))
        for i in range(cnt):
            self.L.n2m_ctx_kernel_stat(self.h, i, C.byref(name), C.byref(ms), C.byref(n), C.byref(b))
            out[name.value.decode()] = (ms.value, n.value, b.value)
        return out
