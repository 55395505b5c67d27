"""Test helpers: bindings of the checkers (the oracle restatement and the
reference compiled from its own sources) and small seeded volumes."""
import ctypes as C
import os

import numpy as np
import pytest

from nii2mesh_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, 'oracle', 'liborc.so')
REF_SO = {k: os.path.join(ROOT, 'oracle', '_ref', 'libn2mref_%s.so' % k) for k in ('classic', 'lewiner')}

_cache = {}


class Oracle(abi.MeshLib):
    """MeshLib over the oracle plus its stage functions."""

    def __init__(self, handle):
        super().__init__(handle, prefix='orc_', cubes_arg=True)
        h = handle
        h.orc_quick_smooth.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                       C.c_int]
        h.orc_laplacian_smooth.argtypes = [np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'),
                                           np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int]
        h.orc_laplacian_smooth.restype = None
        h.orc_border_vertices.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                          np.ctypeslib.ndpointer(np.uint8, flags='C_CONTIGUOUS')]
        h.orc_border_vertices.restype = None
        h.orc_otsu.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        h.orc_lewiner_select.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                         C.POINTER(C.c_int)]
        f32 = C.POINTER(C.c_float)
        h.orc_apply_sform.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'),
                                      np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                      f32, f32, f32]
        h.orc_apply_sform.restype = None
        h.orc_laplacian_smooth_T.argtypes = [np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'),
                                             np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int]
        h.orc_laplacian_smooth_T.restype = None
        h.orc_diff_forward.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'),
                                       C.POINTER(C.c_short), C.c_float, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
        h.orc_diff_gradient.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'),
                                        C.POINTER(C.c_short), C.c_float, C.c_int, C.c_int,
                                        np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'),
                                        np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS')]
        h.orc_decode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float,
                                 np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS')]
        self.mc_classic = _bind_mc(h, 'orc_mc_classic', classic=True)
        self.mc_lewiner = _bind_mc(h, 'orc_mc_lewiner', classic=False)
        self.unify = abi.bind_unify(h, 'orc_unify_vertices')
        self.degenerate = abi.bind_degenerate(h, 'orc_remove_degenerate_triangles')

    def quick_smooth(self, img):
        nz, ny, nx = img.shape
        return self.lib.orc_quick_smooth(img, nx, ny, nz)

    def laplacian_smooth(self, pts, tris, iters=1):
        for _ in range(iters):
            self.lib.orc_laplacian_smooth(pts, np.ascontiguousarray(tris, np.int32), len(pts), len(tris))
        return pts

    def border(self, tris, nv):
        b = np.zeros(nv + 1, np.uint8)
        self.lib.orc_border_vertices(np.ascontiguousarray(tris, np.int32), nv, len(tris), b)
        return b[:nv]

    def otsu(self, hist, mode):
        d, m, b = C.c_int(), C.c_int(), C.c_int()
        t = self.lib.orc_otsu(np.ascontiguousarray(hist, np.int32), len(hist), mode, C.byref(d), C.byref(m),
                              C.byref(b))
        return t, d.value, m.value, b.value

    def select(self, cube, lut, originalMC=0):
        cs = C.c_int()
        row = self.lib.orc_lewiner_select(np.ascontiguousarray(cube, np.float32), lut, originalMC, C.byref(cs))
        return row, cs.value

    def apply_sform(self, tris, pts, sx, sy, sz):
        f = C.c_float * 4
        self.lib.orc_apply_sform(tris, pts, len(tris), len(pts), f(*sx), f(*sy), f(*sz))

    def decode(self, raw, datatype, slope, inter):
        raw = np.ascontiguousarray(raw)
        out = np.empty(raw.size, np.float32)
        rc = self.lib.orc_decode(raw.ctypes.data, int(datatype), raw.size, float(slope), float(inter), out)
        return rc, out

    def laplacian_smooth_T(self, g, tris, iters=1):
        for _ in range(iters):
            self.lib.orc_laplacian_smooth_T(g, np.ascontiguousarray(tris, np.int32), len(g), len(tris))
        return g

    def diff_forward(self, img, iso, lap_iter=20, originalMC=0, exact=False):
        """Lewiner over the whole volume + weld + degenerate + Laplacian."""
        nz, ny, nx = img.shape
        t, p = C.c_void_p(), C.c_void_p()
        nt, npt = C.c_int(0), C.c_int(0)
        rc = self.lib.orc_diff_forward(np.ascontiguousarray(img, np.float32), (C.c_short * 3)(nx, ny, nz),
                                       float(iso), int(originalMC), int(lap_iter), int(exact), C.byref(t),
                                       C.byref(p), C.byref(nt), C.byref(npt))
        if rc != 0:
            return rc, np.zeros((0, 3), np.int32), np.zeros((0, 3), np.float64)
        return rc, abi._take(t.value, nt.value, np.int32), abi._take(p.value, npt.value, np.float64)

    def diff_gradient(self, img, iso, dLdp, lap_iter=20, originalMC=0):
        """FP64 dL/dvolume (shape of img) from dL/dp of diff_forward's vertices."""
        nz, ny, nx = img.shape
        out = np.zeros(img.shape, np.float64)
        rc = self.lib.orc_diff_gradient(np.ascontiguousarray(img, np.float32), (C.c_short * 3)(nx, ny, nz),
                                        float(iso), int(originalMC), int(lap_iter),
                                        np.ascontiguousarray(dLdp, np.float64), out)
        if rc != 0:
            raise RuntimeError('orc_diff_gradient failed')
        return out


def _bind_mc(h, name, classic):
    f = getattr(h, name)
    args = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.POINTER(C.c_short), C.POINTER(C.c_int),
            C.POINTER(C.c_int)]
    if not classic:
        args.append(C.c_int)
    args += [C.c_float, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    f.argtypes = args
    f.restype = C.c_int

    def call(img, lo, hi, iso, originalMC=0):
        nz, ny, nx = img.shape
        dim = (C.c_short * 3)(nx, ny, nz)
        vs, ts = C.c_void_p(), C.c_void_p()
        nv, nt = C.c_int(0), C.c_int(0)
        a = [img, dim, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi)]
        if not classic:
            a.append(int(originalMC))
        a += [float(iso), C.byref(vs), C.byref(ts), C.byref(nv), C.byref(nt)]
        rc = f(*a)
        if rc != 0:
            return rc, np.zeros((0, 3), np.int32), np.zeros((0, 3), np.float64)
        return rc, abi._take(ts.value, nt.value, np.int32), abi._take(vs.value, nv.value, np.float64)
    return call


class Ref(abi.MeshLib):
    """MeshLib over the reference's own sources (oracle/_ref) plus statics."""

    def __init__(self, handle):
        super().__init__(handle)
        h = handle
        h.ref_quick_smooth.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                       C.c_int]
        h.ref_laplacian_smooth.argtypes = [np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'),
                                           np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int]
        h.ref_laplacian_smooth.restype = None
        h.ref_dilate.argtypes = [np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS'), C.POINTER(C.c_size_t)]
        h.ref_dilate.restype = None
        self.marchingCubes = abi.bind_marching_cubes(h)
        self.unify = abi.bind_unify(h, 'ref_unify_vertices')
        self.degenerate = abi.bind_degenerate(h, 'ref_remove_degenerate_triangles')
        f32 = C.POINTER(C.c_float)
        h.apply_sform.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'),
                                  np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                  f32, f32, f32]
        h.apply_sform.restype = None
        h.save_mesh.argtypes = [C.c_char_p, np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'),
                                np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS'), C.c_int, C.c_int, C.c_bool]
        h.base64_encode.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        h.base64_encode.restype = C.c_void_p
        h.ref_nii_otsu.argtypes = [np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'), C.c_int, C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]

    def quick_smooth(self, img):
        nz, ny, nx = img.shape
        return self.lib.ref_quick_smooth(img, nx, ny, nz)

    def laplacian_smooth(self, pts, tris, iters=1):
        for _ in range(iters):
            self.lib.ref_laplacian_smooth(pts, np.ascontiguousarray(tris, np.int32), len(pts), len(tris))
        return pts

    def dilate(self, img):
        nz, ny, nx = img.shape
        self.lib.ref_dilate(img, (C.c_size_t * 3)(nx, ny, nz))

    def save_mesh(self, fnm, tris, pts, gz):
        return self.lib.save_mesh(fnm.encode(), np.ascontiguousarray(tris, np.int32),
                                  np.ascontiguousarray(pts, np.float64), len(tris), len(pts), bool(gz))

    def base64(self, data):
        n = C.c_size_t()
        ptr = self.lib.base64_encode(data, len(data), C.byref(n))
        out = C.string_at(ptr, n.value)
        self.lib.ref_free(C.c_void_p(ptr))
        return out

    def apply_sform(self, tris, pts, sx, sy, sz):
        f = C.c_float * 4
        self.lib.apply_sform(tris, pts, len(tris), len(pts), f(*sx), f(*sy), f(*sz))

    def otsu(self, hist, mode):
        """nii_otsu (src/isolevel.c:141): (threshold, dark, mid, bright)."""
        d, m, b = C.c_int(), C.c_int(), C.c_int()
        h = np.ascontiguousarray(hist, np.int32).copy()
        t = self.lib.ref_nii_otsu(h, len(h), mode, C.byref(d), C.byref(m), C.byref(b))
        return t, d.value, m.value, b.value


def otsu_histograms():
    """256-bin histograms for nii_otsu: random, flat, empty, one spike, four
    equal spikes (many equal objective values), mirror-symmetric, bimodal,
    and sparse Poisson ones."""
    rng = np.random.default_rng(77)
    out = [rng.integers(0, 5000, 256), np.ones(256, int), np.zeros(256, int)]
    h = np.zeros(256, int)
    h[40] = 1000
    out.append(h)
    h = np.zeros(256, int)
    h[[10, 80, 150, 220]] = 500
    out.append(h)
    sym = rng.integers(0, 100, 128)
    out.append(np.concatenate([sym, sym[::-1]]))
    x = np.arange(256)
    out.append((1000 * (np.exp(-(x - 60) ** 2 / 200) + np.exp(-(x - 190) ** 2 / 300))).astype(int))
    for k in range(6):
        out.append(rng.poisson(rng.uniform(1, 300), 256) * (rng.uniform(size=256) < 0.7))
    return [np.ascontiguousarray(h, np.int32) for h in out]


def oracle():
    if 'orc' not in _cache:
        if not os.path.exists(ORACLE_SO):
            pytest.skip('oracle not built (make -C oracle oracle)')
        _cache['orc'] = Oracle(C.CDLL(ORACLE_SO))
    return _cache['orc']


def ref(kind):
    if kind not in _cache:
        if not os.path.exists(REF_SO[kind]):
            pytest.skip('reference build absent (make -C oracle ref)')
        _cache[kind] = Ref(C.CDLL(REF_SO[kind]))
    return _cache[kind]


def blobs(shape, seed, nblobs=6, noise=0.0, dtype=np.float32):
    """Sum of Gaussian blobs on a grid (a smooth field with several
    components, holes and saddles), values roughly 0..100."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    v = np.zeros(shape, np.float64)
    for _ in range(nblobs):
        c = rng.uniform(0, 1, 3) * np.array([nz, ny, nx])
        s = rng.uniform(1.5, max(2.0, 0.25 * max(shape)))
        a = rng.uniform(-60, 100)
        v += a * np.exp(-((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2) / (2 * s * s))
    if noise:
        v += rng.normal(0, noise, shape)
    return np.ascontiguousarray(v.astype(dtype))


def sphere_loss(p, centre, r0):
    """Mean squared radial distance of the vertices to a sphere and its
    gradient with respect to the vertices."""
    d = p - np.asarray(centre, np.float64)
    r = np.sqrt((d * d).sum(1))
    e = r - r0
    L = float((e * e).mean())
    g = (2.0 / len(p)) * (e / r)[:, None] * d
    return L, g


def sphere_field(n, radius, seed, nharm=4, amp=1.5, noise=0.0):
    """Small perturbed-sphere signed field on an n^3 grid (float32)."""
    rng = np.random.default_rng(seed)
    c = (n - 1) / 2.0
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing='ij')
    dx, dy, dz = x - c, y - c, z - c
    r = np.sqrt(dx * dx + dy * dy + dz * dz) + 1e-9
    pert = np.zeros_like(r)
    for _ in range(nharm):
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        pert += amp * np.cos(rng.integers(1, 4) * np.arccos(np.clip((dx * k[0] + dy * k[1] + dz * k[2]) / r, -1, 1))
                             + rng.uniform(0, 2 * np.pi))
    f = r - (radius + pert / nharm)
    if noise:
        f += rng.normal(0, noise, f.shape)
    return f.astype(np.float32)


def random_mask(shape, seed, p=0.5):
    rng = np.random.Generator(np.random.PCG64(seed))
    return (rng.uniform(size=shape) < p).astype(np.float32)


def nan_volume(shape, seed, frac=0.002):
    """blobs() with a fraction of voxels set to NaN, as in masked float maps;
    voxel 0 is NaN too (the range loops of the reference start from it)."""
    img = blobs(shape, seed, noise=2.0)
    rng = np.random.default_rng(seed)
    img[rng.uniform(size=img.shape) < frac] = np.nan
    img[0, 0, 0] = np.nan
    return img


def nan_eq(a, b):
    """Bitwise equality of float arrays, except that any NaN equals any NaN
    (the payload a NaN carries through arithmetic is not part of the
    contract)."""
    a, b = np.asarray(a), np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if a.shape != b.shape or a.dtype != b.dtype or not np.array_equal(na, nb):
        return False
    u = np.uint64 if a.dtype.itemsize == 8 else np.uint32
    return np.array_equal(np.where(na, 0, a).view(u), np.where(nb, 0, b).view(u))


def tiny_volume(shape):
    """A small test volume: blobs() when every side is at least 2, else
    seeded normal noise (blobs needs room for its shapes)."""
    if min(shape) > 1:
        return blobs(shape, 3, noise=1.0)
    return np.random.default_rng(1).normal(size=shape).astype(np.float32)


# ------------------------------------------------------------- real data
DATA = os.path.join(ROOT, 'tests', 'golden', 'data')
DT_BYTES = {2: (np.uint8, 1), 4: (np.int16, 2), 16: (np.float32, 4), 512: (np.uint16, 2)}


def read_nifti(path):
    """NIfTI-1 (.nii or .nii.gz, native little endian) as (header dict, raw
    voxels (nz, ny, nx) in the file's datatype) -- the fields the reference
    loader uses (src/nii2mesh.c:78-172): dim, datatype, vox_offset,
    scl_slope (0 reads as 1), scl_inter, srow_x/y/z."""
    import gzip
    import struct
    with (gzip.open(path, 'rb') if path.endswith('.gz') else open(path, 'rb')) as f:
        b = f.read()
    dim = struct.unpack('<8h', b[40:56])
    dt = struct.unpack('<h', b[70:72])[0]
    off = int(round(struct.unpack('<f', b[108:112])[0]))
    slope, inter = struct.unpack('<2f', b[112:120])
    srow = np.frombuffer(b[280:328], '<f4').astype(np.float32).reshape(3, 4)
    nx, ny, nz = dim[1], dim[2], dim[3]
    typ, _ = DT_BYTES[dt]
    raw = np.frombuffer(b, typ, count=nx * ny * nz, offset=off).reshape(nz, ny, nx)
    return {'dim': (nx, ny, nz), 'datatype': dt, 'slope': np.float32(slope if slope != 0 else 1.0),
            'inter': np.float32(inter), 'srow': srow}, raw


def loader_float(hdr, raw):
    """The reference loader's conversion: raw * scl_slope + scl_inter in
    float32, one rounding per operation (src/nii2mesh.c:155-172)."""
    return np.ascontiguousarray(raw.astype(np.float32) * hdr['slope'] + hdr['inter'], dtype=np.float32)


def mesh_digest(rc, tris, pts):
    """rc, counts and SHA-256 of the vec3i / vec3d bytes of a mesh."""
    import hashlib
    t = np.ascontiguousarray(tris, '<i4')
    p = np.ascontiguousarray(pts, '<f8')
    return {'rc': int(rc), 'nt': int(len(t)), 'np': int(len(p)),
            'tris': hashlib.sha256(t.tobytes()).hexdigest(), 'pts': hashlib.sha256(p.tobytes()).hexdigest()}


def volume_digest(vol):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(vol, '<f4').tobytes()).hexdigest()


# the README's flag sets for bet.nii.gz (README.md:76-124): (name, isolevel
# (1/2/3 = dark/medium/bright, or a number), preSmooth, onlyLargest,
# fillBubbles, postSmooth); the CLI defaults are medium, -p 1, -l 1, -b 0, -s 0
BET_CASES = [
    ('p0', 2, False, True, False, 0),
    ('p1', 2, True, True, False, 0),
    ('l0', 2, True, False, False, 0),
    ('l1', 2, True, True, False, 0),
    ('i122_l0_b0', 122.0, True, False, False, 0),
    ('i122_l0_b1', 122.0, True, False, True, 0),
    ('id', 1, True, True, False, 0),
    ('im', 2, True, True, False, 0),
    ('ib', 3, True, True, False, 0),
    ('i122', 122.0, True, True, False, 0),
    ('i120_p0_s0', 120.0, False, True, False, 0),
    ('i120_p0_s100', 120.0, False, True, False, 100),
]


def write_nifti(path, vol, datatype=16, srow=None):
    """Minimal NIfTI-1 (.nii, little endian, vox_offset 352) of a (nz, ny, nx)
    volume, with an sform (identity unless srow, a 3 x 4 float array)."""
    import struct
    nz, ny, nx = vol.shape
    typ, bpp = DT_BYTES[datatype]
    h = bytearray(348)
    struct.pack_into('<i', h, 0, 348)
    struct.pack_into('<8h', h, 40, 3, nx, ny, nz, 1, 1, 1, 1)
    struct.pack_into('<2h', h, 70, datatype, 8 * bpp)
    struct.pack_into('<8f', h, 76, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into('<f', h, 108, 352.0)
    struct.pack_into('<2f', h, 112, 1.0, 0.0)
    struct.pack_into('<h', h, 254, 1)  # sform_code
    s = np.eye(3, 4, dtype=np.float32) if srow is None else np.asarray(srow, np.float32)
    h[280:328] = s.astype('<f4').tobytes()
    h[344:348] = b'n+1\0'
    with open(path, 'wb') as f:
        f.write(bytes(h) + b'\0\0\0\0' + np.ascontiguousarray(vol, typ).tobytes())
