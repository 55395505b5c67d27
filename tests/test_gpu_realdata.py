"""Parity on the reference's own data and on the timed workloads at full
size, against digests the reference build wrote (tests/golden/make_digests.py):
rc, counts and SHA-256 of the vec3i / vec3d bytes of every mesh.

* bet.nii.gz with the README's flag sets (README.md:76-124): the raw uint8
  voxels are decoded on the GPU, the isolevel comes from the GPU
  setThreshold, then meshify, apply_sform and HC run on the GPU -- the steps
  of nii2() (src/nii2mesh.c:321-331);
* D99_atlas_v2.0_right.nii.gz with -a 1 (README.md:136, src/nii2mesh.c:492-573),
  every label, through the host atlas entry;
* the 128^3 uint8 volume at isolevel 128 without pre-smoothing (vertices on
  grid corners: weld ties, Lewiner's FLT_EPSILON clamp, merge-made
  degenerates) with 100 HC iterations;
* the 512^3 float32 workload exactly as bench.py times it (GPU thresholds,
  the device pipeline with its threaded meshing, 100 two-ring HC
  iterations per mesh), and the 256^3 / 370-label atlas as bench.py times
  it (batched, classic, pre-smooth), label by label.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def digests(group):
    path = os.path.join(GOLDEN, 'digests_%s.json' % group)
    if not os.path.exists(path):
        pytest.fail('missing %s (python tests/golden/make_digests.py %s)' % (path, group))
    with open(path) as f:
        return json.load(f)


def same(got, want, what):
    for k in ('rc', 'nt', 'np', 'tris', 'pts'):
        assert got[k] == want[k], '%s: %s differs (got %s, reference %s)' % (what, k, got[k], want[k])


def gpu_nii2(gpu, cubes, img, iso, preSmooth, onlyLargest, fillBubbles, post, srow, originalMC=0):
    """meshify -> apply_sform -> laplacian_smoothHC on the GPU."""
    L = gpu.load()
    a = img.copy()
    rc, t, p = gpu.lib(cubes).meshify(a, iso, preSmooth, onlyLargest, fillBubbles, originalMC=originalMC)
    if rc == 0:
        if srow is not None:
            f = C.c_float * 4
            L.n2m_apply_sform(t, p, len(t), len(p), f(*srow[0]), f(*srow[1]), f(*srow[2]))
        if post > 0:
            gpu.lib().laplacian_smoothHC(p, t, 0.1, 0.5, post, True)
    d = support.mesh_digest(rc, t, p)
    d['img_after'] = support.volume_digest(a)
    return d


@pytest.fixture(scope='module')
def bet(gpu):
    hdr, raw = support.read_nifti(os.path.join(support.DATA, 'bet.nii.gz'))
    img = gpu.decode(raw, hdr['datatype'], hdr['slope'], hdr['inter']).reshape(raw.shape)
    return hdr, img


@pytest.mark.parametrize('name', [c[0] for c in support.BET_CASES])
@pytest.mark.parametrize('kind,o1', [('classic', 0), ('lewiner', 0), ('lewiner', 1)])
def test_bet_readme_flag_sets(gpu, bet, kind, o1, name):
    want = digests('bet')
    hdr, img = bet
    assert support.volume_digest(img) == want['volume'], 'GPU decode differs from the loader'
    case = dict((c[0], c) for c in support.BET_CASES)[name]
    _, iso, ps, ol, fb, post = case
    cubes = 0 if kind == 'classic' else 1
    w = want['cases']['%s_o%d_%s' % (kind, o1, name)]
    level = float(np.float32(gpu.lib(cubes).setThreshold(img, iso))) if isinstance(iso, int) else float(iso)
    assert level == w['isolevel']
    got = gpu_nii2(gpu, cubes, img, level, ps, ol, fb, post, hdr['srow'], o1)
    same(got, w, name)
    assert got['img_after'] == w['img_after'], '%s: mutated volume differs' % name


@pytest.mark.parametrize('kind', ['classic', 'lewiner'])
def test_d99_atlas_every_label(gpu, kind):
    """nii2mesh D99_atlas_v2.0_right.nii.gz -a 1: every label's mesh (after
    apply_sform) equals the reference's."""
    want = digests('d99')
    hdr, raw = support.read_nifti(os.path.join(support.DATA, 'D99_atlas_v2.0_right.nii.gz'))
    img = support.loader_float(hdr, raw)
    assert support.volume_digest(img) == want['volume']
    L = gpu.load()
    prev = L.n2m_set_cubes(0 if kind == 'classic' else 1)
    try:
        rc, res = gpu.atlas(img, 0, True, False, 0, hdr['srow'].ravel())
    finally:
        L.n2m_set_cubes(prev)
    assert rc == 0
    ref = want['builds'][kind]
    assert len(res) == want['nlabels'] == len(ref)
    nmesh = 0
    for label, status, t, p in res:
        w = ref[str(label)]
        if w is None:
            assert status == gpu.ATLAS_EMPTY, label
            continue
        if w['rc'] != 0:
            assert status == gpu.ATLAS_FAILED, label
            continue
        assert status == gpu.ATLAS_OK, label
        same(support.mesh_digest(0, t, p), w, 'label %d' % label)
        nmesh += 1
    assert nmesh > 300


@pytest.mark.parametrize('name', ['classic_p0_l1_b0', 'lewiner_p0_l1_b1', 'classic_p0_l0_b1', 'lewiner_p0_l0_b0'])
def test_plain128_tie_path(gpu, name):
    """128^3 uint8 at isolevel 128 (an integer value the volume holds)
    without pre-smoothing, + 100 HC iterations."""
    from nii2mesh_amd import synth
    want = digests('plain128')
    img = synth.plain_u8()
    assert support.volume_digest(img) == want['volume']
    kind, p, l, b = name.split('_')
    cubes = 0 if kind == 'classic' else 1
    got = gpu_nii2(gpu, cubes, img, 128.0, p == 'p1', l == 'l1', b == 'b1', 100, None)
    w = want['cases'][name]
    same(got, w, name)
    assert got['img_after'] == w['img_after']


def test_c4_pipeline_as_benched(gpu):
    """The 512^3 step of bench.py: GPU thresholds (modes 1, 2, 3 from one
    set of histograms), then n2m_ctx_pipeline (shared pre-smooth, threaded
    meshing, Lewiner, largest + bubbles, 100 HC iterations per mesh)."""
    import torch
    from nii2mesh_amd import synth
    want = digests('c4')
    vol = synth.large_f32()
    assert support.volume_digest(vol) == want['volume']
    d = torch.from_numpy(vol).to('cuda')
    del vol
    torch.cuda.synchronize()
    with gpu.DeviceContext() as ctx:
        L = ctx.L
        L.n2m_ctx_set_thresholds.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_float)]
        isos = (C.c_float * 3)()
        assert L.n2m_ctx_set_thresholds(ctx.h, C.c_void_p(d.data_ptr()), d.numel(), 3, (C.c_int * 3)(1, 2, 3),
                                        isos) == 0
        levels = [float(v) for v in isos]
        assert levels == [w['isolevel'] for w in want['levels']]
        prm = ctx.params(0.0, cubes=gpu.CUBES_LEWINER, preSmooth=True, onlyLargest=True, fillBubbles=True)
        assert ctx.pipeline(d.data_ptr(), (512, 512, 512), prm, levels, 0.1, 0.5, 100) == 0
        for k, w in enumerate(want['levels']):
            t, p = ctx.mesh(k)
            same(support.mesh_digest(0, t, p), w, 'level %d' % k)
            del t, p


@pytest.mark.parametrize('loop', [False, True])
def test_c2_atlas_as_benched(gpu, loop):
    """The 256^3 / 370-label atlas of bench.py (binarise 0.5, pre-smooth,
    classic, onlyLargest off), batched and per-label loop, every label."""
    import torch
    from nii2mesh_amd import synth
    want = digests('c2')
    vol = synth.atlas_u16()
    assert support.volume_digest(vol) == want['volume'], 'the atlas generator differs on this machine'
    d = torch.from_numpy(vol).to('cuda')
    torch.cuda.synchronize()
    with gpu.DeviceContext() as ctx:
        prm = ctx.params(0.5, cubes=gpu.CUBES_CLASSIC, preSmooth=True, onlyLargest=False, fillBubbles=False)
        assert ctx.atlas(d.data_ptr(), tuple(vol.shape[::-1]), prm, 0, None, loop=loop) == 0
        ents = ctx.atlas_entries()
        assert len(ents) == want['nlabels'] == 370
        for k, e in enumerate(ents):
            w = want['labels'][str(k + 1)]
            assert w is not None and w['rc'] == 0 and e.status == gpu.ATLAS_OK, k + 1
            t, p = ctx.atlas_mesh(k)
            same(support.mesh_digest(0, t, p), w, 'label %d' % (k + 1))
