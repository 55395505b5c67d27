#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference build.

The reference ships no expected outputs, so these vectors are produced by
the reference compiled from its own sources (oracle/_ref, built by
`make -C oracle ref`) on small seeded volumes; the inputs are regenerated
here, the outputs are the reference's.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests import support  # noqa: E402


def main():
    rc_ = support.ref('classic')
    rl_ = support.ref('lewiner')
    n = 0
    cases = [
        # name, shape, seed, iso percentile, flags (preSmooth, onlyLargest, fillBubbles), cubes, o1
        ('meshify_classic_a', (18, 20, 22), 1, 60, (True, True, False), 0, 0),
        ('meshify_classic_b', (16, 17, 21), 2, 45, (False, False, True), 0, 0),
        ('meshify_lewiner_a', (18, 20, 22), 3, 55, (True, True, True), 1, 0),
        ('meshify_lewiner_o1', (17, 19, 16), 4, 65, (True, False, False), 1, 1),
    ]
    for name, shape, seed, pct, flags, cubes, o1 in cases:
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
            a = m.copy()
            nl = rc_.bwlabel(a, conn, largest, fill)
            np.savez_compressed(os.path.join(HERE, 'bwlabel_c%d_l%d_f%d.npz' % (conn, largest, fill)),
                                kind='bwlabel', img=m, conn=conn, largest=largest, fill=fill, n=nl, out=a)
            n += 1
    img = support.blobs((20, 21, 22), 9, noise=5.0)
    vals = [rc_.setThreshold(img, m) for m in (1, 2, 3)]
    np.savez_compressed(os.path.join(HERE, 'threshold.npz'), kind='threshold', img=img, modes=np.array([1, 2, 3]),
                        values=np.array(vals, np.float32))
    n += 1
    print('wrote %d fixtures' % n)


if __name__ == '__main__':
    main()
