"""Atlas meshing (src/nii2mesh.c:492-583) on the GPU against the oracle.

Every label i of an indexed volume is binarised as (i-0.5 < v < i+0.5),
meshed by meshify() at isolevel 0.5 with onlyLargest off, mapped through the
sform and HC-smoothed; the batched path (all labels on crops in one pass)
and the per-label loop must both give, label by label, the oracle's mesh
bit for bit, and the same status (skipped / failed / meshed).
"""
import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu

OK, EMPTY, FAILED = 0, 1, 2


def small_atlas(shape, nlab, seed):
    """Voronoi cells of an ellipsoid that touches the x faces; label 2 left
    out, a one-voxel label, a label stored off-integer, a label on a corner."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    seeds = rng.uniform(0, 1, (nlab, 3)) * np.array([nz, ny, nx])
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    d = ((zz[..., None] - seeds[:, 0]) ** 2 + (yy[..., None] - seeds[:, 1]) ** 2
         + (xx[..., None] - seeds[:, 2]) ** 2)
    lab = (np.argmin(d, axis=-1) + 1).astype(np.float32)
    r = (((xx - nx / 2 + 0.5) / (nx / 2)) ** 2 + ((yy - ny / 2 + 0.5) / (0.42 * ny)) ** 2
         + ((zz - nz / 2 + 0.5) / (0.42 * nz)) ** 2)
    lab[r > 1.0] = 0
    lab[lab == 2] = 0                      # an id with no voxels
    lab[lab == 3] = 3.3                    # still label 3
    lab[nz // 2, ny // 2, nx // 2] = nlab + 1  # one voxel
    lab[0:3, 0:3, 0:3] = nlab + 2          # a corner block touching three faces
    return lab


def expect(orc, vol, cubes, originalMC, preSmooth, fillBubbles, post, srow):
    """Oracle per label: (status, tris, pts)."""
    nl = int(np.trunc(vol.max()))
    out = []
    for i in range(1, nl + 1):
        b = ((vol > np.float32(i - 0.5)) & (vol < np.float32(i + 0.5))).astype(np.float32)
        if not b.any():
            out.append((EMPTY, None, None))
            continue
        rc, t, p = orc.meshify(b, 0.5, preSmooth, False, fillBubbles, originalMC=originalMC, cubes=cubes)
        if rc != 0:
            out.append((FAILED, None, None))
            continue
        if srow is not None:
            orc.apply_sform(t, p, srow[0:4], srow[4:8], srow[8:12])
        if post:
            orc.laplacian_smoothHC(p, t, 0.1, 0.5, post, True)
        out.append((OK, t, p))
    return out


def run_gpu(gpu, vol, cubes, originalMC, preSmooth, fillBubbles, post, srow, loop):
    import torch
    ctx = gpu.DeviceContext()
    d = torch.from_numpy(np.ascontiguousarray(vol)).to('cuda')
    prm = ctx.params(0.5, cubes=cubes, originalMC=originalMC, preSmooth=preSmooth, onlyLargest=False,
                     fillBubbles=fillBubbles)
    nz, ny, nx = vol.shape
    torch.cuda.synchronize()  # the library runs on its own stream
    rc = ctx.atlas(d.data_ptr(), (nx, ny, nz), prm, post, srow, loop=loop)
    res = []
    for k, e in enumerate(ctx.atlas_entries()):
        assert e.label == k + 1
        if e.status == OK:
            t, p = ctx.atlas_mesh(k)
            res.append((OK, t, p))
        else:
            res.append((e.status, None, None))
    ctx.close()
    return rc, res


def compare(got, want):
    assert len(got) == len(want)
    for k, ((sg, tg, pg), (sw, tw, pw)) in enumerate(zip(got, want)):
        assert sg == sw, 'label %d: status %d, oracle %d' % (k + 1, sg, sw)
        if sw != OK:
            continue
        assert tg.shape == tw.shape and np.array_equal(tg, tw), 'label %d: triangles differ' % (k + 1)
        assert pg.shape == pw.shape and np.array_equal(pg, pw), 'label %d: vertices differ' % (k + 1)


SROW_FLIP = np.array([-1.5, 0.0, 0.0, 90.0, 0.0, 1.25, 0.1, -120.0, 0.0, 0.0, 1.1, -70.0], np.float32)


@pytest.mark.parametrize('loop', [False, True])
@pytest.mark.parametrize('cubes,originalMC', [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize('preSmooth', [True, False])
def test_atlas_small(gpu, oracle, cubes, originalMC, preSmooth, loop):
    vol = small_atlas((30, 34, 38), 14, 11)
    want = expect(oracle, vol, cubes, originalMC, preSmooth, False, 0, None)
    rc, got = run_gpu(gpu, vol, cubes, originalMC, preSmooth, False, 0, None, loop)
    compare(got, want)
    assert rc == (0 if any(s == OK for s, _, _ in want) else 1)
    assert sum(s == EMPTY for s, _, _ in want) >= 1


@pytest.mark.parametrize('loop', [False, True])
@pytest.mark.parametrize('cubes', [0, 1])
def test_atlas_sform_hc(gpu, oracle, cubes, loop):
    vol = small_atlas((28, 30, 33), 10, 5)
    want = expect(oracle, vol, cubes, 0, True, False, 12, SROW_FLIP)
    rc, got = run_gpu(gpu, vol, cubes, 0, True, False, 12, SROW_FLIP, loop)
    compare(got, want)
    assert rc == 0


@pytest.mark.parametrize('cubes', [0, 1])
def test_atlas_fill_bubbles(gpu, oracle, cubes):
    vol = small_atlas((26, 30, 28), 8, 7)
    # a hollow shell label around a cavity so the bubble fill matters
    vol[8:18, 10:20, 9:19] = 20
    vol[11:15, 13:17, 12:16] = 0
    want = expect(oracle, vol, cubes, 0, True, True, 3, None)
    rc, got = run_gpu(gpu, vol, cubes, 0, True, True, 3, None, False)
    compare(got, want)


@pytest.mark.parametrize('cubes', [0, 1])
def test_atlas_chunks(gpu, oracle, cubes, monkeypatch):
    """Crops split over several passes (small crop budget)."""
    monkeypatch.setenv('N2M_ATLAS_CHUNK', '3000')
    vol = small_atlas((30, 32, 36), 12, 3)
    want = expect(oracle, vol, cubes, 0, True, False, 4, None)
    rc, got = run_gpu(gpu, vol, cubes, 0, True, False, 4, None, False)
    compare(got, want)


def test_atlas_edge_cases(gpu, oracle):
    # a label that fills the volume: no variability after smoothing
    vol = np.ones((12, 13, 14), np.float32)
    rc, got = run_gpu(gpu, vol, 1, 0, True, False, 0, None, False)
    assert rc != 0 and [s for s, _, _ in got] == [FAILED]
    # no label >= 1: the CLI refuses the volume
    import torch
    ctx = gpu.DeviceContext()
    z = torch.zeros((8, 8, 8), dtype=torch.float32, device='cuda')
    torch.cuda.synchronize()
    assert ctx.atlas(z.data_ptr(), (8, 8, 8), ctx.params(0.5, onlyLargest=False)) != 0
    assert 'label' in gpu.last_error()
    ctx.close()
    # labels on a 1-voxel-thin volume, and a tiny label that takes the
    # isolevel fallback
    vol = np.zeros((1, 9, 11), np.float32)
    vol[0, 2:6, 3:8] = 1
    vol[0, 7, 9] = 2
    for cubes in (0, 1):
        want = expect(oracle, vol, cubes, 0, True, False, 0, None)
        rc, got = run_gpu(gpu, vol, cubes, 0, True, False, 0, None, False)
        compare(got, want)
    vol = np.zeros((10, 11, 12), np.float32)
    vol[5, 5, 5] = 1
    vol[1:4, 1:5, 2:6] = 3
    for cubes in (0, 1):
        want = expect(oracle, vol, cubes, 0, True, False, 2, None)
        rc, got = run_gpu(gpu, vol, cubes, 0, True, False, 2, None, False)
        compare(got, want)


def test_atlas_256_batched_vs_loop(gpu, oracle):
    """The 256^3 / 370-label workload: batched == per-label loop for every
    label; a few labels (largest, smallest, median) also against the oracle."""
    from nii2mesh_amd import synth
    vol = synth.atlas_u16()
    rc_b, got_b = run_gpu(gpu, vol, 1, 0, True, False, 10, None, False)
    rc_l, got_l = run_gpu(gpu, vol, 1, 0, True, False, 10, None, True)
    assert rc_b == rc_l == 0
    compare(got_b, got_l)
    sizes = [(len(t) if s == OK else -1) for s, t, _ in got_b]
    ok = [k for k in range(len(sizes)) if sizes[k] >= 0]
    ok.sort(key=lambda k: sizes[k])
    for k in {ok[0], ok[len(ok) // 2], ok[-1]}:
        i = k + 1
        b = ((vol > np.float32(i - 0.5)) & (vol < np.float32(i + 0.5))).astype(np.float32)
        rc, t, p = oracle.meshify(b, 0.5, True, False, False, cubes=1)
        assert rc == 0
        oracle.laplacian_smoothHC(p, t, 0.1, 0.5, 10, True)
        assert np.array_equal(got_b[k][1], t) and np.array_equal(got_b[k][2], p), 'label %d' % i


@pytest.mark.parametrize('tile', [0, 1, 2, 3])
@pytest.mark.parametrize('cubes', [0, 1])
def test_atlas_every_tile(gpu, oracle, cubes, tile, monkeypatch):
    """Each crop tile the autotuner may pick gives the oracle's meshes."""
    monkeypatch.setenv('N2M_ATLAS_TILE', str(tile))
    vol = small_atlas((33, 35, 31), 11, 8)
    want = expect(oracle, vol, cubes, 0, True, False, 0, None)
    rc, got = run_gpu(gpu, vol, cubes, 0, True, False, 0, None, False)
    compare(got, want)


_PERSIST_WANT = {}


@pytest.mark.parametrize('bpc', [0, 1, 4])
@pytest.mark.parametrize('cubes', [0, 1])
def test_atlas_persistent_grids(gpu, oracle, cubes, bpc, monkeypatch):
    """Every stage launched persistent at one blocks-per-CU choice (index
    into 0 / 1 / 2 / 4 / 8; 0 is one workgroup per item) on crops with more
    work items than a 1-per-CU grid has workgroups, so the workgroups loop:
    the oracle's meshes, HC included."""
    monkeypatch.setenv('N2M_ATLAS_BPC', str(bpc))
    vol = small_atlas((96, 96, 96), 24, 5)
    if cubes not in _PERSIST_WANT:
        _PERSIST_WANT[cubes] = expect(oracle, vol, cubes, 0, True, False, 3, None)
    rc, got = run_gpu(gpu, vol, cubes, 0, True, False, 3, None, False)
    assert rc == 0
    compare(got, _PERSIST_WANT[cubes])


def test_atlas_tuner_report(gpu):
    """After a batched run every stage it launched reports a tuned shape."""
    vol = small_atlas((40, 42, 44), 9, 3)
    rc, _ = run_gpu(gpu, vol, 1, 0, True, False, 0, None, False)
    assert rc == 0
    t = gpu.atlas_tuned(0)
    assert set(t) == {'smooth', 'apply', 'classic_count', 'classic_emit', 'lewiner_classify', 'lewiner_emit'}
    for stage in ('smooth', 'apply', 'lewiner_classify', 'lewiner_emit'):
        tile, bpc = t[stage]
        assert bpc in (0, 1, 2, 4, 8), (stage, t[stage])
        assert (tile >= 0) == (stage == 'smooth'), (stage, t[stage])


@pytest.mark.parametrize('cubes', [0, 1])
def test_atlas_host_entry(gpu, oracle, cubes):
    """n2m_atlas (host arrays, the CLI's form) gives the oracle's meshes."""
    vol = small_atlas((27, 29, 31), 9, 12)
    want = expect(oracle, vol, cubes, 0, True, False, 5, SROW_FLIP)
    L = gpu.load()
    prev = L.n2m_set_cubes(cubes)
    try:
        rc, res = gpu.atlas(vol, 0, True, False, 5, SROW_FLIP)
    finally:
        L.n2m_set_cubes(prev)
    assert rc == 0
    compare([(s, t, p) for _, s, t, p in res], want)
    assert [lab for lab, _, _, _ in res] == list(range(1, len(want) + 1))
