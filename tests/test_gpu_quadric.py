"""quadric_simplify_mesh (src/quadric.c:396-518) with its first update pass
on the GPU, against the reference's own build (oracle/_ref) bit for bit:
marching-cubes meshes before and after HC, a real-data mesh, high-valence
fans with borders, and the CLI's reductions and quality settings (lossy at
0.5 / 0.25 / 0.15 with aggressiveness 8 / 7 / 5, the lossless finish of
-q 2, and a target above the triangle count)."""
import numpy as np
import pytest

from nii2mesh_amd import abi
from tests import support

pytestmark = pytest.mark.gpu


def meshes(gpu, oracle):
    out = []
    img = support.blobs((40, 44, 48), 3, noise=2.0)
    for cubes in (0, 1):
        rc, t, p = oracle.meshify(img.copy(), float(np.float32(np.percentile(img, 60))), True, True, False,
                                  cubes=cubes)
        assert rc == 0
        out.append(('blobs_c%d' % cubes, t, p))
        q = p.copy()
        oracle.laplacian_smoothHC(q, t, 0.1, 0.5, 20, True)
        out.append(('blobs_c%d_hc' % cubes, t, q))
    from tests.test_gpu_parity import fan_mesh
    t, p = fan_mesh(300, 4)
    out.append(('fan', t, p))
    return out


@pytest.mark.parametrize('r,aggr,lossless', [(0.5, 8.0, False), (0.25, 7.0, False), (0.15, 5.0, True),
                                             (1.0, 5.0, True), (2.0, 7.0, False)])
def test_quadric_matches_reference(gpu, oracle, ref_classic, r, aggr, lossless):
    mine = abi.bind_quadric(gpu.load())
    theirs = abi.bind_quadric(ref_classic.lib)
    for name, t, p in meshes(gpu, oracle):
        target = int(round(np.float32(len(t)) * np.float32(r)))
        ta, pa = mine(p, t, target, aggr, False, lossless)
        tb, pb = theirs(p, t, target, aggr, False, lossless)
        assert len(ta) < len(t) or r >= 1.0, name
        assert np.array_equal(ta, tb) and np.array_equal(pa.view(np.uint64), pb.view(np.uint64)), name


def test_quadric_bet_readme(gpu, ref_classic):
    """The README's -r 0.15 on bet.nii.gz (classic, default flags)."""
    import os
    import json
    hdr, raw = support.read_nifti(os.path.join(support.DATA, 'bet.nii.gz'))
    img = support.loader_float(hdr, raw)
    rc, t, p = gpu.lib(0).meshify(img.copy(), float(np.float32(gpu.lib(0).setThreshold(img, 2))), True, True,
                                  False)
    assert rc == 0
    target = int(round(np.float32(len(t)) * np.float32(0.15)))
    ta, pa = abi.bind_quadric(gpu.load())(p, t, target, 7.0)
    tb, pb = abi.bind_quadric(ref_classic.lib)(p, t, target, 7.0)
    assert np.array_equal(ta, tb) and np.array_equal(pa, pb)
    assert len(ta) <= target + 2
