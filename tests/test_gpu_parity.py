"""GPU parity: every stage of the HIP library against the oracle, bit-exact.

Integer/index work (labels, cases, triangle lists, vertex numbering) must be
identical; floating-point outputs must be bit-identical too, because both
sides perform the reference's operations in the same order and precision
(src/meshify.c, src/bwlabel.c, src/oldcubes.c, src/MarchingCubes.c,
src/quadric.c, src/isolevel.c).
"""
import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu

SHAPES = [(7, 9, 11), (24, 31, 40), (48, 48, 48), (33, 65, 70)]


def _eq(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize('shape', SHAPES + [(5, 5, 5), (4, 9, 9)])
@pytest.mark.parametrize('seed', [1, 2])
def test_quick_smooth(gpu, oracle, shape, seed):
    """src/meshify.c:170-216"""
    img = support.blobs(shape, seed, noise=3.0)
    a, b = img.copy(), img.copy()
    ra = gpu.load().n2m_quick_smooth(a, shape[2], shape[1], shape[0])
    rb = oracle.quick_smooth(b)
    assert ra == rb
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize('shape', [(13, 17, 19), (40, 33, 70), (1, 9, 9), (3, 2, 40)])
@pytest.mark.parametrize('p', [0.3, 0.55, 0.75])
@pytest.mark.parametrize('conn', [6, 18, 26])
@pytest.mark.parametrize('largest,fill', [(False, False), (True, False), (False, True), (True, True)])
def test_bwlabel(gpu, oracle, shape, p, conn, largest, fill):
    """src/bwlabel.c:429-543: labels numbered by first voxel, largest cluster,
    bubble fill."""
    m = support.random_mask(shape, int(p * 100) + conn + 7 * largest + 3 * fill, p)
    a, b = m.copy(), m.copy()
    na = gpu.lib().bwlabel(a, conn, largest, fill)
    nb = oracle.bwlabel(b, conn, largest, fill)
    assert na == nb
    assert np.array_equal(a, b)


@pytest.mark.parametrize('shape', SHAPES)
@pytest.mark.parametrize('seed', [3, 4])
@pytest.mark.parametrize('flavour', ['classic', 'lewiner', 'lewiner_o1'])
def test_marching_cubes(gpu, oracle, shape, seed, flavour):
    """src/oldcubes.c:465-522 and src/MarchingCubes.c:1086-1143 on the whole
    box and on an interior sub-box (Lewiner reads past hi with wrap)."""
    img = support.blobs(shape, seed, noise=0.5)
    nz, ny, nx = shape
    iso = float(np.float32(np.percentile(img, 60)))
    L = gpu.load()
    import nii2mesh_amd.abi as abi
    mc = abi.bind_marching_cubes(L)
    for lo, hi in [((0, 0, 0), (nx, ny, nz)), ((1, 2, 1), (nx - 1, ny - 2, nz))]:
        if flavour == 'classic':
            L.n2m_set_cubes(0)
            got = mc(img, lo, hi, iso)
            exp = oracle.mc_classic(img, lo, hi, iso)
        else:
            L.n2m_set_cubes(1)
            o1 = 1 if flavour == 'lewiner_o1' else 0
            got = mc(img, lo, hi, iso, o1)
            exp = oracle.mc_lewiner(img, lo, hi, iso, o1)
        assert got[0] == exp[0]
        assert _eq(got[1], exp[1])
        assert _eq(got[2], exp[2])


def _raw_mesh(oracle, shape, seed, flavour):
    img = support.blobs(shape, seed, noise=0.5)
    nz, ny, nx = shape
    iso = float(np.float32(np.percentile(img, 55)))
    if flavour == 'classic':
        return oracle.mc_classic(img, (0, 0, 0), (nx, ny, nz), iso)
    return oracle.mc_lewiner(img, (0, 0, 0), (nx, ny, nz), iso, 0)


@pytest.mark.parametrize('shape', SHAPES[1:])
@pytest.mark.parametrize('flavour', ['classic', 'lewiner'])
def test_unify_and_degenerate(gpu, oracle, shape, flavour):
    """src/meshify.c:45-106 (weld) and :113-168 (degenerate filter)."""
    import nii2mesh_amd.abi as abi
    rc, tris, pts = _raw_mesh(oracle, shape, 5, flavour)
    assert rc == 0
    L = gpu.load()
    gu = abi.bind_unify(L, 'n2m_unify_vertices')
    gd = abi.bind_degenerate(L, 'n2m_remove_degenerate_triangles')
    p1, t1 = gu(pts, tris)
    p2, t2 = oracle.unify(pts, tris)
    assert _eq(p1, p2) and _eq(t1, t2)
    d1 = gd(p1, t1)
    d2 = oracle.degenerate(p2, t2)
    assert _eq(d1, d2)


@pytest.mark.parametrize('n', [2048, 2049, 128 * 2048, 128 * 2048 + 1])
def test_unify_scan_tile_boundaries(gpu, oracle, n):
    """Weld and degenerate removal on n points and n triangles at the scan's
    size boundaries: one 2048-word tile, two tiles, the largest single-pass
    scan (128 tiles) and the smallest three-level one (129 tiles).  Points
    on a 40^3 integer lattice, so most of them are welded."""
    import nii2mesh_amd.abi as abi
    rng = np.random.default_rng(n)
    pts = np.ascontiguousarray(rng.integers(0, 40, size=(n, 3)).astype(np.float64))
    tris = np.ascontiguousarray(rng.integers(0, n, size=(n, 3)), np.int32)
    L = gpu.load()
    gu = abi.bind_unify(L, 'n2m_unify_vertices')
    gd = abi.bind_degenerate(L, 'n2m_remove_degenerate_triangles')
    p1, t1 = gu(pts, tris)
    p2, t2 = oracle.unify(pts, tris)
    assert len(p1) < n and _eq(p1, p2) and _eq(t1, t2)
    d1 = gd(p1, t1)
    d2 = oracle.degenerate(p2, t2)
    assert _eq(d1, d2)


@pytest.mark.parametrize('first_nan', [False, True])
def test_unify_nan_rows(gpu, oracle, first_nan):
    """Weld with vertices at NaN positions: each is a class of its own (its
    key sorts last); with vertex 0 at NaN every key is NaN and nothing
    merges."""
    import nii2mesh_amd.abi as abi
    rc, tris, pts = _raw_mesh(oracle, (24, 31, 40), 5, 'classic')
    assert rc == 0
    rng = np.random.Generator(np.random.PCG64(3))
    n, k = len(pts), 90
    at = np.sort(rng.choice(np.arange(0 if first_nan else 1, n + k), size=k, replace=False))
    if first_nan:
        at[0] = 0
    isnan = np.zeros(n + k, bool)
    isnan[at] = True
    newidx = np.flatnonzero(~isnan)
    pts2 = np.full((n + k, 3), np.nan)
    pts2[newidx] = pts
    tris2 = np.concatenate([newidx[tris], at[:k // 3 * 3].reshape(-1, 3)]).astype(np.int32)
    gu = abi.bind_unify(gpu.load(), 'n2m_unify_vertices')
    p1, t1 = gu(pts2, tris2)
    p2, t2 = oracle.unify(pts2, tris2)
    assert support.nan_eq(p1, p2) and _eq(t1, t2)
    assert (len(p1) == n + k) if first_nan else (len(p1) < n + k)


@pytest.mark.parametrize('fillers', [0, 8, 70])
def test_unify_claimed_twice(gpu, oracle, fillers):
    """A vertex claimed by a seed and then by a later seed of the same key
    window ends in the later class (src/meshify.c:66-79); the earlier class
    keeps only its own members' positions.  Seed i = 1 claims j = 3 (8e-6
    apart); i' = 2 is 1.07e-5 from i, so it seeds its own class, and claims
    j (8.6e-6 apart).  `fillers` more vertices in the same key segment, far
    apart in 3-D, make the segment take the wave path (8: 11 vertices) or
    the one-lane path (70: 73 vertices, longer than a wave)."""
    import nii2mesh_amd.abi as abi
    pts = [(0.0, 0.0, 0.0), (10.0, 0.0, 0.0), (10.0000065, 8.5e-6, 0.0), (10.000008, 0.0, 0.0)]
    for k in range(1, fillers + 1):
        r, th = 10.0 - 5e-6 * k, 0.1 * k
        pts.append((r * np.cos(th), r * np.sin(th), 0.0))
    pts = np.array(pts, np.float64)
    tris = np.array([[0, 1, 2], [1, 2, 3]] + [[0, 1, 3 + k] for k in range(1, fillers + 1)], np.int32)
    gu = abi.bind_unify(gpu.load(), 'n2m_unify_vertices')
    p1, t1 = gu(pts, tris)
    p2, t2 = oracle.unify(pts, tris)
    assert _eq(p1, p2) and _eq(t1, t2)
    assert len(p2) == len(pts) - 1 and t2[1, 1] == t2[1, 2]  # j went to i''s class


@pytest.mark.parametrize('flavour', ['classic', 'lewiner'])
@pytest.mark.parametrize('lock', [True, False])
def test_hc_smooth(gpu, oracle, flavour, lock):
    """src/quadric.c:343-394: FP64, same summation order -> bit-identical."""
    img = support.blobs((40, 44, 48), 9, noise=2.0)
    rc, tris, pts = oracle.meshify(img.copy(), float(np.percentile(img, 50)), preSmooth=True,
                                   onlyLargest=False, fillBubbles=False, cubes=0 if flavour == 'classic' else 1)
    assert rc == 0
    a, b = pts.copy(), pts.copy()
    gpu.lib().laplacian_smoothHC(a, tris, 0.1, 0.5, 25, lock)
    oracle.laplacian_smoothHC(b, tris, 0.1, 0.5, 25, lock)
    assert np.array_equal(a, b)
    # plain Laplacian (src/quadric.c:316-341)
    c, d = pts.copy(), pts.copy()
    gpu.load().n2m_laplacian_smooth(c, np.ascontiguousarray(tris), len(c), len(tris), 7)
    oracle.laplacian_smooth(d, tris, 7)
    assert np.array_equal(c, d)


FORM_ENV = {'window': ('0', '0'), 'halo': ('0', '1'), 'ring': ('1', '1')}  # N2M_HC_RING, N2M_HC_HALO


@pytest.mark.parametrize('form', ['window', 'halo', 'ring'])
@pytest.mark.parametrize('iters', [1, 2, 12])
@pytest.mark.parametrize('lock', [True, False])
def test_hc_gather_forms(gpu, oracle, monkeypatch, form, iters, lock):
    """The three forms of the HC iteration -- two launches with an LDS window
    and global escapes, two launches with LDS block + halo and 16-bit slots,
    and one launch with block + two rings in LDS -- give the oracle's
    laplacian_smoothHC bit for bit on a Morton-relabelled mesh of several
    blocks.  The stage statistics show which form ran."""
    import torch
    ring, halo = FORM_ENV[form]
    monkeypatch.setenv('N2M_HC_RING', ring)
    monkeypatch.setenv('N2M_HC_HALO', halo)
    # 34865 vertices: Morton-relabelled (>= 32768), 35 blocks of 1024, 46 of 768
    img = support.blobs((80, 84, 88), 4, noise=2.0)
    iso = float(np.float32(np.percentile(img, 50)))
    ctx = gpu.DeviceContext()
    ctx.timing(True)
    d = torch.from_numpy(img).cuda()
    torch.cuda.synchronize()  # the library runs on its own stream
    prm = ctx.params(iso, cubes=1, preSmooth=True, onlyLargest=False, fillBubbles=False)
    assert ctx.meshify(d.data_ptr(), (88, 84, 80), prm) == 0
    assert ctx.smooth_hc(0, 0.1, 0.5, iters, lock) == 0
    t, p = ctx.mesh(0)
    stats = ctx.kernel_stats()
    ctx.close()
    rc, to, po = oracle.meshify(img.copy(), iso, True, False, False, cubes=1)
    assert rc == 0 and len(po) >= 32768
    oracle.laplacian_smoothHC(po, to, 0.1, 0.5, iters, lock)
    assert _eq(t, to) and np.array_equal(p, po)
    assert ('hc_halo' in stats) == (form == 'halo')
    assert ('hc_ring' in stats) == (form == 'ring')
    assert ('hc_a' in stats) == (form != 'ring')


@pytest.mark.parametrize('rb', ['512', '640', '768'])
@pytest.mark.parametrize('lock', [True, False])
def test_hc_ring_block_sizes(gpu, oracle, monkeypatch, rb, lock):
    """The one-launch HC iteration at each block size (N2M_HC_RB) on a noisy
    surface of 55000 vertices.  With 768-vertex blocks, 18 of the 72 blocks
    have more than 256 ring-1 neighbours (1024 threads - 768), so their
    workgroups compute ring-1 rows in a second pass.  The rings fit, the
    one-launch form runs, and the result is the oracle's bit for bit."""
    import torch
    monkeypatch.setenv('N2M_HC_RB', rb)
    img = support.blobs((80, 84, 88), 4, noise=8.0)
    iso = float(np.float32(np.percentile(img, 50)))
    ctx = gpu.DeviceContext()
    ctx.timing(True)
    d = torch.from_numpy(img).cuda()
    torch.cuda.synchronize()  # the library runs on its own stream
    prm = ctx.params(iso, cubes=1, preSmooth=True, onlyLargest=False, fillBubbles=False)
    assert ctx.meshify(d.data_ptr(), (88, 84, 80), prm) == 0
    assert ctx.smooth_hc(0, 0.1, 0.5, 9, lock) == 0
    t, p = ctx.mesh(0)
    stats = ctx.kernel_stats()
    ctx.close()
    rc, to, po = oracle.meshify(img.copy(), iso, True, False, False, cubes=1)
    assert rc == 0 and len(po) == 55000
    oracle.laplacian_smoothHC(po, to, 0.1, 0.5, 9, lock)
    assert _eq(t, to) and np.array_equal(p, po)
    assert 'hc_ring' in stats and 'hc_a' not in stats


@pytest.mark.parametrize('threads,hc_in_worker', [('1', '1'), ('1', '0'), ('0', '0')])
@pytest.mark.parametrize('cubes', [0, 1])
def test_pipeline_isolevels(gpu, oracle, monkeypatch, threads, hc_in_worker, cubes):
    """The device pipeline (one pre-smooth shared by several isolevels, each
    meshed and HC-smoothed -- by one host thread and context per isolevel,
    smoothing as soon as its mesh is ready or after all meshing, or one
    after the other) gives, per isolevel, the oracle's meshify +
    laplacian_smoothHC(lockEdges) bit for bit."""
    import torch
    monkeypatch.setenv('N2M_PIPE_THREADS', threads)
    monkeypatch.setenv('N2M_PIPE_HC', hc_in_worker)
    img = support.blobs((60, 64, 68), 8, noise=2.0)
    isos = [float(np.float32(np.percentile(img, q))) for q in (40, 60, 80)]
    ctx = gpu.DeviceContext()
    d = torch.from_numpy(img).cuda()
    torch.cuda.synchronize()  # the library runs on its own stream
    prm = ctx.params(0.0, cubes=cubes, preSmooth=True, onlyLargest=True, fillBubbles=True)
    assert ctx.pipeline(d.data_ptr(), (68, 64, 60), prm, isos, hc_iters=10) == 0
    got = [ctx.mesh(k) for k in range(len(isos))]
    ctx.close()
    for k, iso in enumerate(isos):
        rc, to, po = oracle.meshify(img.copy(), iso, True, True, True, cubes=cubes)
        assert rc == 0
        oracle.laplacian_smoothHC(po, to, 0.1, 0.5, 10, True)
        assert _eq(got[k][0], to) and np.array_equal(got[k][1], po), k


def test_hc_halo_overflow_falls_back(gpu, oracle):
    """A mesh without locality (random triangles over 40000 vertices): every
    block has far more outside neighbours than its rings or halo hold, so the
    two-ring build and the halo build overflow and the window kernels run; the
    result is still the oracle's, bit for bit."""
    import torch
    rng = np.random.default_rng(17)
    nv, nt = 40000, 30000
    pts = rng.normal(size=(nv, 3)) * 10.0
    tris = np.stack([rng.permutation(nv)[:nt], rng.permutation(nv)[:nt], rng.permutation(nv)[:nt]], 1)
    tris = tris[(tris[:, 0] != tris[:, 1]) & (tris[:, 1] != tris[:, 2]) & (tris[:, 0] != tris[:, 2])]
    tris = np.ascontiguousarray(tris, np.int32)
    a, b = pts.copy(), pts.copy()
    gpu.lib().laplacian_smoothHC(a, tris, 0.1, 0.5, 6, True)
    oracle.laplacian_smoothHC(b, tris, 0.1, 0.5, 6, True)
    assert np.array_equal(a, b)
    c, d = pts.copy(), pts.copy()
    gpu.load().n2m_laplacian_smooth(c, tris, len(c), len(tris), 4)
    oracle.laplacian_smooth(d, tris, 4)
    assert np.array_equal(c, d)


def test_hc_large_host_transfer(gpu, oracle):
    """laplacian_smoothHC on host arrays of more than 32 MiB each (a
    1205 x 1203 grid with a rough height field: 1.45 M vertices, 2.89 M
    triangles, sizes no multiple of the 4 MiB staging chunk): the vertices
    and triangles go up, and the vertices come back, through the pinned
    multi-threaded staging; the result is the oracle's, bit for bit."""
    nx, ny = 1205, 1203
    rng = np.random.default_rng(23)
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    pts = np.ascontiguousarray(np.stack([x, y, rng.normal(size=(ny, nx)) * 0.3], -1).reshape(-1, 3))
    v = np.arange(nx * ny, dtype=np.int32).reshape(ny, nx)
    a0, b0, c0, d0 = v[:-1, :-1].ravel(), v[:-1, 1:].ravel(), v[1:, :-1].ravel(), v[1:, 1:].ravel()
    tris = np.ascontiguousarray(np.concatenate([np.stack([a0, b0, d0], 1), np.stack([a0, d0, c0], 1)]), np.int32)
    assert pts.nbytes > (32 << 20) and tris.nbytes > (32 << 20) and pts.nbytes % (4 << 20) != 0
    a, b = pts.copy(), pts.copy()
    gpu.lib().laplacian_smoothHC(a, tris, 0.1, 0.5, 2, True)
    oracle.laplacian_smoothHC(b, tris, 0.1, 0.5, 2, True)
    assert np.array_equal(a, b)


@pytest.mark.parametrize('case', ['no_iterations', 'no_triangles', 'one_triangle', 'isolated_vertices',
                                  'ragged_blocks'])
def test_hc_small_and_ragged(gpu, oracle, case):
    """laplacian_smoothHC edge cases: 0 iterations, no triangle at all, a
    single triangle (all border), vertices no triangle uses, and a vertex
    count that is not a multiple of the 1024-vertex block."""
    rng = np.random.default_rng(23)
    iters = 5
    if case == 'no_iterations':
        img = support.blobs((20, 22, 24), 2, noise=1.0)
        _, tris, pts = oracle.meshify(img.copy(), float(np.percentile(img, 50)), True, False, False, cubes=1)
        iters = 0
    elif case == 'no_triangles':
        pts, tris = rng.normal(size=(50, 3)), np.zeros((0, 3), np.int32)
    elif case == 'one_triangle':
        pts, tris = rng.normal(size=(3, 3)), np.array([[0, 1, 2]], np.int32)
    elif case == 'isolated_vertices':
        pts = rng.normal(size=(12, 3))
        tris = np.array([[0, 1, 2], [2, 1, 3], [3, 1, 4], [4, 1, 0]], np.int32)
    else:
        img = support.blobs((70, 74, 78), 6, noise=2.0)
        _, tris, pts = oracle.meshify(img.copy(), float(np.percentile(img, 50)), True, False, False, cubes=0)
        assert len(pts) % 1024 != 0 and len(pts) > 32768
    for lock in (True, False):
        a, b = pts.copy(), pts.copy()
        gpu.lib().laplacian_smoothHC(a, tris, 0.1, 0.5, iters, lock)
        oracle.laplacian_smoothHC(b, tris, 0.1, 0.5, iters, lock)
        assert np.array_equal(a, b), (case, lock)


@pytest.mark.parametrize('seed', [11, 12, 13])
@pytest.mark.parametrize('mode', [1, 2, 3])
def test_set_threshold(gpu, oracle, seed, mode):
    """src/isolevel.c:245-277 (GPU histograms, host Otsu)."""
    img = support.blobs((30, 40, 50), seed, noise=4.0)
    assert gpu.lib().setThreshold(img, mode) == oracle.setThreshold(img, mode)


MESHIFY_CASES = [
    # (preSmooth, onlyLargest, fillBubbles)
    (True, True, False),
    (True, True, True),
    (True, False, True),
    (False, True, False),
    (True, False, False),
    (False, False, False),
]


@pytest.mark.parametrize('shape', [(24, 31, 40), (48, 48, 48)])
@pytest.mark.parametrize('flags', MESHIFY_CASES)
@pytest.mark.parametrize('flavour', ['classic', 'lewiner', 'lewiner_o1'])
def test_meshify_blobs(gpu, oracle, shape, flags, flavour):
    """meshify() end to end (src/meshify.c:286-389), including the mutated
    volume the reference leaves behind."""
    img = support.blobs(shape, 21, noise=3.0)
    iso = float(np.float32(np.percentile(img, 65)))
    cubes = 0 if flavour == 'classic' else 1
    o1 = 1 if flavour == 'lewiner_o1' else 0
    a, b = img.copy(), img.copy()
    g = gpu.lib(cubes)
    ra = g.meshify(a, iso, *flags, originalMC=o1)
    rb = oracle.meshify(b, iso, *flags, originalMC=o1, cubes=cubes)
    assert ra[0] == rb[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert _eq(ra[1], rb[1])
    assert _eq(ra[2], rb[2])


def _sign_bits_volume(kind, shape, seed):
    if kind == 'blobs':
        return support.blobs(shape, seed, noise=3.0), None
    if kind == 'face':  # an object cut by the x = nx - 1 face: the sub-box reaches nx (the wrapped reads)
        img = support.blobs(shape, seed, noise=1.0)
        img[:, :, -3:] += 80.0
        return img, None
    if kind == 'ties':  # integers at an integer isolevel: values equal to iso, FLT_EPSILON clamps
        rng = np.random.default_rng(seed)
        return rng.integers(0, 4, size=shape).astype(np.float32), 2.0
    img = support.nan_volume(shape, seed)  # 'nan'
    return img, float(np.float32(np.nanpercentile(img, 60)))


@pytest.mark.parametrize('shape', [(9, 11, 32), (20, 23, 64), (14, 18, 96)])
@pytest.mark.parametrize('kind', ['blobs', 'face', 'ties', 'nan'])
@pytest.mark.parametrize('flags', [(True, True, True), (False, False, False), (True, False, True)])
@pytest.mark.parametrize('o1', [0, 1])
def test_meshify_lewiner_sign_bits(gpu, oracle, shape, kind, flags, o1):
    """Lewiner meshify on volumes whose rows are whole 32-voxel words, where
    the classify pass reads the apply stage's sign bits instead of the
    values: the same meshes and mutated volume as the oracle."""
    img, iso = _sign_bits_volume(kind, shape, 30 + shape[2])
    if iso is None:
        iso = float(np.float32(np.percentile(img, 62)))
    a, b = img.copy(), img.copy()
    ra = gpu.lib(1).meshify(a, iso, *flags, originalMC=o1)
    rb = oracle.meshify(b, iso, *flags, originalMC=o1, cubes=1)
    assert ra[0] == rb[0]
    assert support.nan_eq(a, b)
    assert _eq(ra[1], rb[1]) and support.nan_eq(ra[2], rb[2])


@pytest.mark.parametrize('flags', [(True, True, True), (False, False, False), (False, True, False)])
@pytest.mark.parametrize('flavour', ['classic', 'lewiner'])
def test_meshify_nan_volume(gpu, oracle, flags, flavour):
    """Volumes with NaN voxels: a vertex at a NaN position welds alone, and a
    triangle whose corner the Lewiner tables left without a vertex (-1) is
    dropped by the degenerate filter -- the rules of the oracle, where the
    reference indexes old2new[-1] (src/meshify.c:66-79, :88-100)."""
    img = support.nan_volume((26, 28, 30), 3)
    iso = float(np.float32(np.nanpercentile(img, 60)))
    cubes = 0 if flavour == 'classic' else 1
    a, b = img.copy(), img.copy()
    ra = gpu.lib(cubes).meshify(a, iso, *flags)
    rb = oracle.meshify(b, iso, *flags, cubes=cubes)
    assert ra[0] == rb[0] == 0
    assert support.nan_eq(a, b)
    assert _eq(ra[1], rb[1]) and support.nan_eq(ra[2], rb[2])
    assert ra[1].min() >= 0 and ra[1].max() < len(ra[2])


@pytest.mark.parametrize('shape', [(1, 6, 7), (2, 2, 2), (3, 1, 4), (5, 5, 1), (1, 1, 1), (6, 5, 4)])
@pytest.mark.parametrize('cubes', [0, 1])
def test_meshify_tiny_volumes(gpu, oracle, shape, cubes):
    """Volumes with a dimension of 1 or 2: the same return codes, meshes and
    mutated volume as the oracle."""
    img = support.tiny_volume(shape)
    iso = float(np.float32(np.median(img)))
    for flags in [(True, True, True), (False, False, False)]:
        a, b = img.copy(), img.copy()
        ra = gpu.lib(cubes).meshify(a, iso, *flags)
        rb = oracle.meshify(b, iso, *flags, cubes=cubes)
        assert ra[0] == rb[0]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert _eq(ra[1], rb[1]) and _eq(ra[2], rb[2])


@pytest.mark.parametrize('cubes', [0, 1])
def test_meshify_concurrent_threads(gpu, oracle, cubes):
    """meshify() and laplacian_smoothHC() called from several host threads
    at once, as the atlas mode's OpenMP loop does (src/nii2mesh.c:546-551):
    each thread uses its own context and stream, and every result is the
    oracle's."""
    import threading
    lib = gpu.lib(cubes)
    vols = [support.blobs((36 + 4 * i, 44, 48), 30 + i, noise=2.0) for i in range(6)]
    isos = [float(np.float32(np.percentile(v, 55))) for v in vols]
    out = [None] * len(vols)

    def work(i):
        r = lib.meshify(vols[i].copy(), isos[i], True, True, True)
        p = r[2].copy()
        if r[0] == 0:
            lib.laplacian_smoothHC(p, r[1], 0.1, 0.5, 10, True)
        out[i] = (r[0], r[1], p)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(vols))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i, v in enumerate(vols):
        rc, to, po = oracle.meshify(v.copy(), isos[i], True, True, True, cubes=cubes)
        assert rc == 0
        oracle.laplacian_smoothHC(po, to, 0.1, 0.5, 10, True)
        assert out[i] is not None and out[i][0] == rc
        assert _eq(out[i][1], to) and np.array_equal(out[i][2], po), i


def test_meshify_iso_fallback_and_failures(gpu, oracle):
    """Isolevel outside the range -> midpoint (src/meshify.c:316-319);
    constant volume -> failure (:312-315)."""
    img = support.blobs((20, 20, 20), 5)
    for iso in (1e9, -1e9):
        a, b = img.copy(), img.copy()
        ra = gpu.lib(0).meshify(a, iso)
        rb = oracle.meshify(b, iso, cubes=0)
        assert ra[0] == rb[0] and _eq(ra[1], rb[1]) and _eq(ra[2], rb[2])
    c = np.full((10, 10, 10), 3.0, np.float32)
    assert gpu.lib(0).meshify(c.copy(), 1.0)[0] == 1


@pytest.mark.parametrize('flavour,flags', [('classic', (True, True, False)), ('lewiner', (True, True, True))])
def test_meshify_plain_u8_128(gpu, oracle, flavour, flags):
    """The 128^3 uint8 workload at isolevel 128."""
    from nii2mesh_amd import synth
    img = synth.plain_u8()
    cubes = 0 if flavour == 'classic' else 1
    a, b = img.copy(), img.copy()
    ra = gpu.lib(cubes).meshify(a, 128.0, *flags)
    rb = oracle.meshify(b, 128.0, *flags, cubes=cubes)
    assert ra[0] == rb[0] == 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert _eq(ra[1], rb[1]) and _eq(ra[2], rb[2])
    pa, pb = ra[2].copy(), rb[2].copy()
    gpu.lib().laplacian_smoothHC(pa, ra[1], 0.1, 0.5, 10, True)
    oracle.laplacian_smoothHC(pb, rb[1], 0.1, 0.5, 10, True)
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize('lock', [True, False])
def test_hc_large_mesh_relabelled(gpu, oracle, lock):
    """Meshes above the relabelling threshold (Morton-ordered gathers) stay
    bit-identical to the reference order of operations."""
    img = support.blobs((96, 100, 104), 31, noise=4.0)
    rc, tris, pts = oracle.meshify(img.copy(), float(np.percentile(img, 50)), True, False, False, cubes=1)
    assert rc == 0 and len(pts) > 40000
    a, b = pts.copy(), pts.copy()
    gpu.lib().laplacian_smoothHC(a, tris, 0.1, 0.5, 15, lock)
    oracle.laplacian_smoothHC(b, tris, 0.1, 0.5, 15, lock)
    assert np.array_equal(a, b)
    c, d = pts.copy(), pts.copy()
    gpu.load().n2m_laplacian_smooth(c, np.ascontiguousarray(tris), len(c), len(tris), 4)
    oracle.laplacian_smooth(d, tris, 4)
    assert np.array_equal(c, d)


@pytest.mark.parametrize('nv,nt', [(50, 80), (40000, 70000)])
def test_hc_arbitrary_triangles(gpu, oracle, nv, nt):
    """laplacian_smoothHC is a public entry (obj2mesh calls it): arbitrary
    triangle lists, including repeated vertices, orphans and open borders."""
    rng = np.random.Generator(np.random.PCG64(nv))
    pts = rng.normal(size=(nv, 3)) * 10
    tris = rng.integers(0, nv - 3, size=(nt, 3)).astype(np.int32)
    tris[::7, 1] = tris[::7, 0]  # degenerate: repeated vertex
    a, b = pts.copy(), pts.copy()
    gpu.lib().laplacian_smoothHC(a, tris, 0.1, 0.5, 6, True)
    oracle.laplacian_smoothHC(b, tris, 0.1, 0.5, 6, True)
    assert np.array_equal(a, b)


@pytest.mark.parametrize('shape', [(9, 11, 1100), (20, 24, 600), (6, 5, 1025)])
@pytest.mark.parametrize('p', [0.35, 0.6])
@pytest.mark.parametrize('conn', [6, 18, 26])
def test_bwlabel_wide_rows(gpu, oracle, shape, p, conn):
    """Rows longer than one labelling tile (512 voxels): runs that cross
    tile and word boundaries in x."""
    m = support.random_mask(shape, conn + int(10 * p), p)
    for largest, fill in [(False, False), (True, True)]:
        a, b = m.copy(), m.copy()
        assert gpu.lib().bwlabel(a, conn, largest, fill) == oracle.bwlabel(b, conn, largest, fill)
        assert np.array_equal(a, b)


def test_bwlabel_long_runs(gpu, oracle):
    """Solid slabs and lines: runs spanning whole rows, giant components."""
    m = np.zeros((17, 19, 700), np.float32)
    m[2:15, 3:16, 5:690] = 1
    m[8, :, :] = 1
    m[:, 9, 100:600] = 0
    for conn in (6, 18, 26):
        for largest, fill in [(False, False), (True, False), (True, True)]:
            a, b = m.copy(), m.copy()
            assert gpu.lib().bwlabel(a, conn, largest, fill) == oracle.bwlabel(b, conn, largest, fill)
            assert np.array_equal(a, b)


def test_golden_fixtures_on_gpu(gpu):
    """The reference's own outputs (tests/golden, written from the
    reference build) reproduced by the HIP library."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
    files = sorted(f for f in os.listdir(here) if f.endswith('.npz'))
    assert files
    for f in files:
        g = np.load(os.path.join(here, f))
        kind = str(g['kind'])
        if kind == 'meshify':
            img = g['img'].copy()
            rc, t, p = gpu.lib(int(g['cubes'])).meshify(img, float(g['iso']), *g['flags'].tolist(),
                                                       originalMC=int(g['o1']))
            assert rc == int(g['rc']) and _eq(t, g['tris']) and _eq(p, g['pts']), f
            assert np.array_equal(img, g['img_after']), f
            q = p.copy()
            gpu.lib().laplacian_smoothHC(q, t, 0.1, 0.5, int(g['hc_iters']), True)
            assert np.array_equal(q, g['pts_hc']), f
        elif kind == 'bwlabel':
            img = g['img'].copy()
            n = gpu.lib().bwlabel(img, int(g['conn']), bool(g['largest']), bool(g['fill']))
            assert n == int(g['n']) and np.array_equal(img, g['out']), f
        elif kind == 'threshold':
            for mode, val in zip(g['modes'], g['values']):
                assert np.float32(gpu.lib().setThreshold(g['img'], int(mode))) == np.float32(val), f


@pytest.mark.parametrize('flip', [False, True])
def test_apply_sform_matches_reference(gpu, oracle, ref_classic, flip):
    """n2m_apply_sform (FP64 on the GPU, winding swapped on the GPU when the
    product of the row sums is negative) against the reference's own
    apply_sform (src/meshify.c:1025-1045), bit for bit."""
    import ctypes as C
    img = support.blobs((26, 24, 22), 12)
    rc, t, p = oracle.meshify(img, float(np.percentile(img, 60)), True, True, False, cubes=0)
    assert rc == 0 and len(t) > 100
    srow = np.array([[0.9973992, -0.0386101, 0.0608605, -69.86039],
                     [0.0402, 1.0213, -0.0129, -102.5],
                     [-0.0598, 0.0155, 0.9987, -73.25]], np.float32)
    if flip:
        srow[0] *= -1.5  # negative product of the row sums
    ta, pa = t.copy(), p.copy()
    tb, pb = t.copy(), p.copy()
    f = C.c_float * 4
    gpu.load().n2m_apply_sform(ta, pa, len(ta), len(pa), f(*srow[0]), f(*srow[1]), f(*srow[2]))
    ref_classic.apply_sform(tb, pb, srow[0], srow[1], srow[2])
    assert np.array_equal(ta, tb) and np.array_equal(pa.view(np.uint64), pb.view(np.uint64))
    assert flip == (not np.array_equal(ta, t))


def fan_mesh(n_apex, seed, repeats=False):
    """A disk fanned around vertex 0 (n_apex triangles), a second apex of
    valence 20 (a short fan), a vertex of valence 14, and, with repeats,
    triangles listing a high-valence vertex twice."""
    rng = np.random.default_rng(seed)
    pts = [np.array([0.0, 0.0, 0.0])]
    tris = []
    a = np.linspace(0, 2 * np.pi, n_apex, endpoint=False)
    rim = np.stack([np.cos(a), np.sin(a), 0.05 * rng.normal(size=n_apex)], 1) * 10
    pts.extend(rim)
    for i in range(n_apex):
        tris.append((0, 1 + i, 1 + (i + 1) % n_apex))
    for m, val in ((20, 20), (14, 14)):
        c = len(pts)
        pts.append(np.array([30.0 + m, 0.0, 0.0]))
        b = np.linspace(0, 2 * np.pi, val, endpoint=False)
        base = len(pts)
        pts.extend(np.stack([30.0 + m + np.cos(b), np.sin(b), 0.1 * rng.normal(size=val)], 1))
        for i in range(val):
            tris.append((c, base + i, base + (i + 1) % val))
    if repeats:
        tris.append((0, 0, 5))
        tris.append((7, 0, 7))
    t = np.array(tris, np.int32)
    order = rng.permutation(len(t))  # high-valence references spread over the triangle list
    p = np.array(pts, np.float64) + rng.normal(scale=0.01, size=(len(pts), 3))
    return t[order], p


@pytest.mark.parametrize('apex', [2000, 12000])
@pytest.mark.parametrize('repeats', [False, True])
@pytest.mark.parametrize('lock', [True, False])
def test_hc_high_valence_fan(gpu, oracle, repeats, lock, apex):
    """A vertex in `apex` triangles (and others of valence 20 and 14): the
    valence-20 list is sorted and counted by one wave, the apex's by one
    workgroup (an LDS sort; with 12000 its 24000 or 36000 border entries
    exceed the LDS and are compared from global memory); HC and the plain
    Laplacian equal the oracle's bit for bit."""
    t, p = fan_mesh(apex, 3, repeats)
    a, b = p.copy(), p.copy()
    gpu.lib().laplacian_smoothHC(a, t, 0.1, 0.5, 5, lock)
    oracle.laplacian_smoothHC(b, t, 0.1, 0.5, 5, lock)
    assert np.array_equal(a, b)
    a, b = p.copy(), p.copy()
    gpu.load().n2m_laplacian_smooth(a, t, len(a), len(t), 3)
    oracle.laplacian_smooth(b, t, 3)
    assert np.array_equal(a, b)
