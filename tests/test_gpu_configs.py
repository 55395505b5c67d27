"""GPU parity at the full sizes of the workload configurations, against the
oracle, bit for bit (the CPU side of each test runs for up to ~40 s)."""
import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return a.shape == b.shape and np.array_equal(a, b)


def test_drop_in_lewiner_int16_256(gpu, oracle):
    """The 256^3 int16 drop-in Lewiner workload: isolevel 500, pre-smooth,
    largest cluster, bubble fill, then 100 HC iterations (src/meshify.c:286-389,
    src/quadric.c:343-394)."""
    from nii2mesh_amd import synth
    img = synth.lewiner_i16()
    a, b = img.copy(), img.copy()
    ra = gpu.lib(1).meshify(a, 500.0, True, True, True)
    rb = oracle.meshify(b, 500.0, True, True, True, cubes=1)
    assert ra[0] == rb[0] == 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert _eq(ra[1], rb[1]) and _eq(ra[2], rb[2])
    pa, pb = ra[2].copy(), rb[2].copy()
    gpu.lib().laplacian_smoothHC(pa, ra[1], 0.1, 0.5, 100, True)
    oracle.laplacian_smoothHC(pb, rb[1], 0.1, 0.5, 100, True)
    assert np.array_equal(pa, pb)


def test_large_f32_512_thresholds_and_mid_level(gpu, oracle):
    """The 512^3 f32 large-pipeline volume: the three multi-Otsu isolevels
    (GPU histograms, src/isolevel.c:245-277) equal the oracle's, and the mid
    isolevel's mesh (pre-smooth, largest cluster, bubble fill, Lewiner) and 3
    HC iterations on it equal the oracle's."""
    from nii2mesh_amd import synth
    img = synth.large_f32()
    thr = [gpu.lib().setThreshold(img, m) for m in (1, 2, 3)]
    assert thr == [oracle.setThreshold(img, m) for m in (1, 2, 3)]
    a, b = img.copy(), img.copy()
    ra = gpu.lib(1).meshify(a, thr[1], True, True, True)
    rb = oracle.meshify(b, thr[1], True, True, True, cubes=1)
    assert ra[0] == rb[0] == 0
    assert _eq(ra[1], rb[1]) and _eq(ra[2], rb[2])
    pa, pb = ra[2].copy(), rb[2].copy()
    gpu.lib().laplacian_smoothHC(pa, ra[1], 0.1, 0.5, 3, True)
    oracle.laplacian_smoothHC(pb, rb[1], 0.1, 0.5, 3, True)
    assert np.array_equal(pa, pb)


def test_large_f32_512_slab_all_levels(gpu, oracle):
    """All three isolevels of the 512^3 volume (including the bright one,
    whose mesh is the largest) on its 512 x 512 x 64 slab z in [224, 288):
    meshify (pre-smooth, largest cluster, bubble fill, Lewiner) and 3 HC
    iterations equal the oracle's."""
    from nii2mesh_amd import synth
    img = synth.large_f32()
    thr = [gpu.lib().setThreshold(img, m) for m in (1, 2, 3)]
    slab = np.ascontiguousarray(img[224:288])
    del img
    for iso in thr:
        a, b = slab.copy(), slab.copy()
        ra = gpu.lib(1).meshify(a, iso, True, True, True)
        rb = oracle.meshify(b, iso, True, True, True, cubes=1)
        assert ra[0] == rb[0] == 0
        assert _eq(ra[1], rb[1]) and _eq(ra[2], rb[2]), iso
        pa, pb = ra[2].copy(), rb[2].copy()
        gpu.lib().laplacian_smoothHC(pa, ra[1], 0.1, 0.5, 3, True)
        oracle.laplacian_smoothHC(pb, rb[1], 0.1, 0.5, 3, True)
        assert np.array_equal(pa, pb), iso
