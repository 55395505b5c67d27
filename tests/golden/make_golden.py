#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference build.

The reference ships no expected outputs, so these vectors are produced by
the reference compiled from its own sources (oracle/_ref, built by
`make -C oracle ref`) on small seeded volumes; the inputs are regenerated
here, the outputs are the reference's.  Run from the repository root:

    python tests/golden/make_golden.py [--force]

Existing files are kept unless --force is given.  The tie cases mesh
integer-valued volumes at an integer isolevel without pre-smoothing, so
vertices land exactly on grid corners: they drive the weld's merges across
edges (src/meshify.c:61-79), Lewiner's FLT_EPSILON clamp
(src/MarchingCubes.c:252-255) and the degenerate triangles those merges make
(src/meshify.c:128,138); they run 100 HC iterations.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests import support  # noqa: E402


def tie_volume(kind, shape, seed):
    """Integer-valued float32 volumes: uniform 0..3 noise, or a sphere of
    rounded doubled radii (2r rounded: shells of equal value)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind == 'noise':
        return rng.integers(0, 4, shape).astype(np.float32)
    nz, ny, nx = shape
    c = rng.uniform(0.4, 0.6, 3) * np.array([nz, ny, nx])
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    r = np.sqrt((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2)
    return np.rint(2.0 * r).astype(np.float32)


TIE_CASES = [
    # name, volume kind, shape, seed, iso, flags (preSmooth, onlyLargest, fillBubbles), cubes, o1
    ('meshify_tie_noise_classic', 'noise', (18, 20, 22), 31, 2.0, (False, False, False), 0, 0),
    ('meshify_tie_noise_lewiner', 'noise', (18, 20, 22), 32, 2.0, (False, True, True), 1, 0),
    ('meshify_tie_noise_lewiner_o1', 'noise', (17, 19, 21), 33, 2.0, (False, False, True), 1, 1),
    ('meshify_tie_sphere_classic', 'sphere', (24, 26, 22), 34, 16.0, (False, True, True), 0, 0),
    ('meshify_tie_sphere_lewiner', 'sphere', (24, 26, 22), 35, 16.0, (False, False, False), 1, 0),
    ('meshify_tie_sphere_lewiner_o1', 'sphere', (23, 21, 25), 36, 15.0, (False, True, False), 1, 1),
]


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument('--force', action='store_true')
    force = ap.parse_args().force
    rc_ = support.ref('classic')
    rl_ = support.ref('lewiner')
    n = 0

    def fresh(name):
        return force or not os.path.exists(os.path.join(HERE, name + '.npz'))

    for name, kind, shape, seed, iso, flags, cubes, o1 in TIE_CASES:
        if not fresh(name):
            continue
        img = tie_volume(kind, shape, seed)
        ref = rc_ if cubes == 0 else rl_
        a = img.copy()
        rc, t, p = ref.meshify(a, iso, *flags, originalMC=o1)
        q = p.copy()
        if rc == 0:
            ref.laplacian_smoothHC(q, t, 0.1, 0.5, 100, True)
        np.savez_compressed(os.path.join(HERE, name + '.npz'), kind='meshify', img=img, iso=iso,
                            flags=np.array(flags), cubes=cubes, o1=o1, rc=rc, tris=t, pts=p, img_after=a,
                            pts_hc=q, hc_iters=100)
        print(name, rc, len(t), len(p))
        n += 1
    cases = [
        # name, shape, seed, iso percentile, flags (preSmooth, onlyLargest, fillBubbles), cubes, o1
        ('meshify_classic_a', (18, 20, 22), 1, 60, (True, True, False), 0, 0),
        ('meshify_classic_b', (16, 17, 21), 2, 45, (False, False, True), 0, 0),
        ('meshify_lewiner_a', (18, 20, 22), 3, 55, (True, True, True), 1, 0),
        ('meshify_lewiner_o1', (17, 19, 16), 4, 65, (True, False, False), 1, 1),
    ]
    for name, shape, seed, pct, flags, cubes, o1 in cases:
        if not fresh(name):
            continue
        img = support.blobs(shape, seed, noise=2.0)
        iso = float(np.float32(np.percentile(img, pct)))
        ref = rc_ if cubes == 0 else rl_
        a = img.copy()
        rc, t, p = ref.meshify(a, iso, *flags, originalMC=o1)
        q = p.copy()
        if rc == 0:
            ref.laplacian_smoothHC(q, t, 0.1, 0.5, 10, True)
        np.savez_compressed(os.path.join(HERE, name + '.npz'), kind='meshify', img=img, iso=iso,
                            flags=np.array(flags), cubes=cubes, o1=o1, rc=rc, tris=t, pts=p, img_after=a,
                            pts_hc=q, hc_iters=10)
        n += 1
    for conn in (6, 18, 26):
        m = support.random_mask((9, 10, 11), 200 + conn, 0.5)
        for largest, fill in [(False, False), (True, True)]:
            if not fresh('bwlabel_c%d_l%d_f%d' % (conn, largest, fill)):
                continue
            a = m.copy()
            nl = rc_.bwlabel(a, conn, largest, fill)
            np.savez_compressed(os.path.join(HERE, 'bwlabel_c%d_l%d_f%d.npz' % (conn, largest, fill)),
                                kind='bwlabel', img=m, conn=conn, largest=largest, fill=fill, n=nl, out=a)
            n += 1
    if not fresh('threshold'):
        print('wrote %d fixtures' % n)
        return
    img = support.blobs((20, 21, 22), 9, noise=5.0)
    vals = [rc_.setThreshold(img, m) for m in (1, 2, 3)]
    np.savez_compressed(os.path.join(HERE, 'threshold.npz'), kind='threshold', img=img, modes=np.array([1, 2, 3]),
                        values=np.array(vals, np.float32))
    n += 1
    print('wrote %d fixtures' % n)


if __name__ == '__main__':
    main()
