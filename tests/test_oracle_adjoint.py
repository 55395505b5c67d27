"""The oracle's differentiable path, pinned on the CPU.

* Its forward (exact=False) equals the reference's own chain --
  marchingCubes (Lewiner) over the whole volume, unify_vertices,
  remove_degenerate_triangles, laplacian_smooth -- from the reference build.
* Its transposed Laplacian satisfies <A x, y> = <x, A^T y> with the
  reference's laplacian_smooth as A.
* Its FP64 gradient matches central differences of the FP64-position
  forward (exact=True) on near-surface voxels whose perturbation leaves the
  mesh topology unchanged.
"""
import numpy as np
import pytest

from tests import support


@pytest.mark.parametrize('originalMC', [0, 1])
def test_diff_forward_is_reference_chain(oracle, ref_lewiner, originalMC):
    img = support.sphere_field(34, 10.0, 4, noise=0.01)
    rc, t, p = oracle.diff_forward(img, 0.0, lap_iter=7, originalMC=originalMC)
    assert rc == 0
    nz, ny, nx = img.shape
    rc2, t2, p2 = ref_lewiner.marchingCubes(img.copy(), (0, 0, 0), (nx - 1, ny - 1, nz - 1), 0.0, originalMC)
    assert rc2 == 0
    p2, t2 = ref_lewiner.unify(p2, t2)
    t2 = ref_lewiner.degenerate(p2, t2)
    ref_lewiner.laplacian_smooth(p2, t2, 7)
    assert np.array_equal(t, t2) and np.array_equal(p, p2)


def test_laplacian_transpose_identity(oracle, ref_lewiner):
    img = support.blobs((24, 26, 22), 9)
    rc, t, p = oracle.meshify(img, float(np.percentile(img, 60)), True, False, False, cubes=1)
    assert rc == 0
    rng = np.random.default_rng(0)
    nv = len(p)
    # isolated vertices (kept by the forward) exercise the identity rows
    x = rng.normal(size=(nv + 3, 3))
    y = rng.normal(size=(nv + 3, 3))
    ax = x.copy()
    ref_lewiner.laplacian_smooth(ax, t, 1)
    aty = y.copy()
    oracle.laplacian_smooth_T(aty, t, 1)
    lhs, rhs = (ax * y).sum(), (x * aty).sum()
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    # three iterations compose
    ax3, aty3 = x.copy(), y.copy()
    ref_lewiner.laplacian_smooth(ax3, t, 3)
    oracle.laplacian_smooth_T(aty3, t, 3)
    assert abs((ax3 * y).sum() - (x * aty3).sum()) <= 1e-12 * max(abs((ax3 * y).sum()), 1.0)


def loss_at(oracle, img, centre, r0, lap):
    rc, t, p = oracle.diff_forward(img, 0.0, lap_iter=lap, exact=True)
    assert rc == 0
    return support.sphere_loss(p, centre, r0)[0], t


def test_oracle_gradient_vs_central_differences(oracle):
    n, lap, r0 = 30, 5, 8.0
    img = support.sphere_field(n, 9.0, 6, noise=0.01)
    c = ((n - 1) / 2.0,) * 3
    rc, t, p = oracle.diff_forward(img, 0.0, lap_iter=lap, exact=True)
    assert rc == 0
    L0, g = support.sphere_loss(p, c, r0)
    grad = oracle.diff_gradient(img, 0.0, g, lap_iter=lap)
    rng = np.random.default_rng(1)
    near = np.argwhere(np.abs(img) < 1.5)
    checked, worst = 0, 0.0
    eps = 1e-3
    for k in rng.permutation(len(near)):
        if checked == 24:
            break
        idx = tuple(near[k])
        a, b = img.copy(), img.copy()
        a[idx] += eps
        b[idx] -= eps
        La, ta = loss_at(oracle, a, c, r0, lap)
        Lb, tb = loss_at(oracle, b, c, r0, lap)
        if not (np.array_equal(ta, t) and np.array_equal(tb, t)):
            continue  # the perturbation changes a case
        fd = (La - Lb) / (2 * eps)
        if abs(fd) < 1e-9:
            continue
        worst = max(worst, abs(grad[idx] - fd) / abs(fd))
        checked += 1
    assert checked >= 16
    assert worst < 1e-3, worst
