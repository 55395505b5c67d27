"""Seeded synthetic volumes for the five workloads (numpy PCG64, seed 1000+k).

All volumes are returned as C-ordered arrays of shape (nz, ny, nx), i.e. the
x-fastest layout the reference uses (src/meshify.c:186).

* `head(scale)`       128^3*scale: ellipsoid (semi-axes 50/40/45 * scale) at
  160, six Gaussian blobs (sigma 4..10, +-40) inside it, two mirrored
  ventricles at 40, a hollow sphere (r 8, shell 2) and a solid sphere (r 6)
  at 200 outside it; Gaussian noise, clipped to 0..255.
* `plain_u8()`        128^3 uint8, noise 8, rounded          (plain CPU path)
* `lewiner_i16()`     256^3 int16, 0..1000, noise 16          (drop-in Lewiner)
* `atlas_u16()`       256^3 uint16 weighted-Voronoi labels (370,
                      log-normal sizes: median ~2k, max ~3-4e5)  (batched atlas)
* `sphere_f32()`      192^3 float32 perturbed-sphere distance  (differentiable)
* `large_f32()`       512^3 float32, noise 8                   (large pipeline)
"""
import numpy as np

SEED0 = 1000


def _rng(k):
    return np.random.Generator(np.random.PCG64(SEED0 + k))


def _grid_box(shape, centre, radius):
    """Index box (slices) and coordinate grids of a cube around centre."""
    sl, coords = [], []
    for d in range(3):
        lo = max(int(np.floor(centre[d] - radius)), 0)
        hi = min(int(np.ceil(centre[d] + radius)) + 1, shape[d])
        sl.append(slice(lo, hi))
        coords.append(np.arange(lo, hi, dtype=np.float32))
    z, y, x = np.meshgrid(*coords, indexing='ij')
    return tuple(sl), z, y, x


def head(scale, noise, rng, dtype=np.float32):
    n = 128 * scale
    shape = (n, n, n)
    c = 64.0 * scale
    vol = np.zeros(shape, np.float32)
    ax = np.array([50.0, 40.0, 45.0]) * scale  # semi-axes x, y, z
    # ellipsoid, processed slab by slab to bound memory
    zz = np.arange(n, dtype=np.float32)
    yy = zz[:, None]
    xx = zz[None, :]
    inside = np.zeros(shape, bool)
    for z in range(n):
        r = ((xx - c) / ax[0]) ** 2 + ((yy - c) / ax[1]) ** 2 + ((z - c) / ax[2]) ** 2
        inside[z] = r <= 1.0
    vol[inside] = 160.0
    # six Gaussian blobs centred inside the ellipsoid
    blobs = 0
    while blobs < 6:
        p = rng.uniform(-1, 1, 3)
        if (p ** 2).sum() > 1.0:
            continue
        centre = (c + p[2] * ax[2], c + p[1] * ax[1], c + p[0] * ax[0])  # z, y, x
        sigma = rng.uniform(4.0, 10.0) * scale
        amp = 40.0 if rng.uniform() < 0.5 else -40.0
        sl, z, y, x = _grid_box(shape, centre, 4 * sigma)
        d2 = (z - centre[0]) ** 2 + (y - centre[1]) ** 2 + (x - centre[2]) ** 2
        vol[sl] += np.where(inside[sl], amp * np.exp(-d2 / (2 * sigma * sigma)), 0).astype(np.float32)
        blobs += 1
    # two ventricles mirrored about the mid-sagittal plane x = c
    va = np.array([6.0, 12.0, 8.0]) * scale
    for sx in (-1, 1):
        centre = (c + 2.0 * scale, c - 4.0 * scale, c + sx * 12.0 * scale)
        sl, z, y, x = _grid_box(shape, centre, va.max() + 1)
        r = ((x - centre[2]) / va[0]) ** 2 + ((y - centre[1]) / va[1]) ** 2 + ((z - centre[0]) / va[2]) ** 2
        sub = vol[sl]
        sub[r <= 1.0] = 40.0
    # hollow sphere (r 8, shell 2) and solid sphere (r 6), outside the ellipsoid
    centre = (c, c + 52.0 * scale, c)
    sl, z, y, x = _grid_box(shape, centre, 9 * scale)
    r = np.sqrt((x - centre[2]) ** 2 + (y - centre[1]) ** 2 + (z - centre[0]) ** 2)
    sub = vol[sl]
    sub[(r <= 8 * scale) & (r > 6 * scale)] = 200.0
    centre = (110.0 * scale, 18.0 * scale, 18.0 * scale)
    sl, z, y, x = _grid_box(shape, centre, 7 * scale)
    r = np.sqrt((x - centre[2]) ** 2 + (y - centre[1]) ** 2 + (z - centre[0]) ** 2)
    sub = vol[sl]
    sub[r <= 6 * scale] = 200.0
    # additive noise, slab by slab
    for z in range(n):
        vol[z] += rng.standard_normal((n, n), dtype=np.float32) * np.float32(noise)
    np.clip(vol, 0, 255, out=vol)
    return vol


def plain_u8():
    """128^3 uint8 head, noise 8 (returned as the float32 the loader makes)."""
    v = head(1, 8.0, _rng(0))
    return np.rint(v).astype(np.uint8).astype(np.float32)


def lewiner_i16():
    """256^3 int16 head, intensities 0..1000, noise 16."""
    rng = _rng(1)
    v = head(2, 0.0, rng)
    v *= np.float32(1000.0 / 255.0)
    for z in range(v.shape[0]):
        v[z] += rng.standard_normal(v.shape[1:], dtype=np.float32) * np.float32(16.0)
    np.clip(v, 0, 1000, out=v)
    return np.rint(v).astype(np.int16).astype(np.float32)


def atlas_u16(n=256, nlabels=370, passes=4):
    """256^3 uint16 labels: the head ellipsoid (semi-axes 50/40/45 * n/128)
    split into `nlabels` cells of a multiplicatively weighted Voronoi
    diagram, a voxel going to the seed with the smallest |x - s_i| / r_i.

    Target cell sizes are drawn log-normal (median 2000 voxels, sigma 1.68
    in log space, rescaled to fill the ellipsoid: about 3.0 M voxels, the
    largest cells about 3-4e5 voxels); r_i starts at the radius of a ball of
    the target size and is corrected `passes - 1` times by the cube root of
    target / measured size.  Seeds sit on integer coordinates, so every
    squared distance is an integer below 2^24 and exact in float32 whatever
    the summation order; the only rounding is the one product with 1/r_i^2,
    so the labels are the same on every machine."""
    import torch
    rng = _rng(2)
    s = n / 128.0
    c = 64.0 * s
    ax = np.array([50.0, 40.0, 45.0]) * s  # x, y, z
    g = np.arange(n, dtype=np.float64)
    inside = (((g[None, None, :] - c) / ax[0]) ** 2 + ((g[None, :, None] - c) / ax[1]) ** 2
              + ((g[:, None, None] - c) / ax[2]) ** 2) <= 1.0
    idx = np.flatnonzero(inside)
    P = np.stack([idx % n, (idx // n) % n, idx // (n * n)], 1).astype(np.float32)
    nin = len(P)
    tgt = np.exp(rng.normal(np.log(2000.0), 1.68, nlabels))
    tgt *= nin / tgt.sum()
    rad = np.cbrt(3.0 * tgt / (4.0 * np.pi))
    # seeds, largest cell first, kept apart by 0.9 x the sum of the radii
    seeds = np.zeros((nlabels, 3))
    placed = []
    for i in np.argsort(-tgt, kind='stable'):
        q = None
        for attempt in range(200):
            cand = rng.uniform(-1, 1, (64, 3))
            cand = np.rint(c + cand[(cand ** 2).sum(1) <= 1.0] * ax)
            if not placed:
                q = cand[0]
                break
            d = np.sqrt(((cand[:, None, :] - seeds[placed][None]) ** 2).sum(-1))
            ok = np.flatnonzero(np.all(d >= 0.9 * (rad[i] + rad[placed])[None], axis=1))
            if len(ok) or attempt == 199:
                q = cand[ok[0] if len(ok) else 0]
                break
        seeds[i] = q
        placed.append(i)
    Pt = torch.from_numpy(P)
    St = torch.from_numpy(seeds.astype(np.float32))
    P2 = (Pt * Pt).sum(1)
    S2 = (St * St).sum(1)
    lab = torch.empty(nin, dtype=torch.int64)
    w = rad.copy()
    B = 1 << 17
    for k in range(passes):
        W2 = torch.from_numpy((1.0 / (w * w)).astype(np.float32))
        for b in range(0, nin, B):
            d = torch.addmm(S2[None], Pt[b:b + B], St.T, alpha=-2.0)  # exact integers
            d.add_(P2[b:b + B, None]).mul_(W2[None])
            lab[b:b + B] = d.argmin(1)
        if k + 1 < passes:
            sz = np.bincount(lab.numpy(), minlength=nlabels)
            w = w * np.cbrt(np.clip(tgt / np.maximum(sz, 1), 0.25, 4.0))
    vol = np.zeros(n * n * n, np.float32)
    vol[idx] = (lab.numpy() + 1).astype(np.float32)
    return vol.reshape(n, n, n)


def _real_sph_harm(l, m, theta, phi):
    from scipy.special import sph_harm
    y = sph_harm(abs(m), l, phi, theta)
    if m > 0:
        return np.sqrt(2) * (-1) ** m * y.real
    if m < 0:
        return np.sqrt(2) * (-1) ** m * y.imag
    return y.real


def sphere_f32(n=192, radius=60.0, nharm=8, amp=3.0, noise=0.02):
    """Signed distance-like field |x-c| - (r + sum_k amp*Y_k(dir)) + noise."""
    rng = _rng(3)
    c = (n - 1) / 2.0
    harm = []
    for _ in range(nharm):
        l = int(rng.integers(1, 5))
        m = int(rng.integers(-l, l + 1))
        harm.append((l, m))
    # normalise each harmonic to unit peak on a fine sphere grid
    th = np.linspace(0, np.pi, 181)[:, None]
    ph = np.linspace(0, 2 * np.pi, 361)[None, :]
    scales = [1.0 / np.abs(_real_sph_harm(l, m, th, ph)).max() for l, m in harm]
    vol = np.empty((n, n, n), np.float32)
    g = np.arange(n, dtype=np.float64) - c
    yy, xx = np.meshgrid(g, g, indexing='ij')
    for z in range(n):
        zz = g[z]
        r = np.sqrt(xx * xx + yy * yy + zz * zz)
        theta = np.arccos(np.clip(zz / np.maximum(r, 1e-9), -1, 1))
        phi = np.arctan2(yy, xx)
        pert = np.zeros_like(r)
        for (l, m), s in zip(harm, scales):
            pert += amp * s * _real_sph_harm(l, m, theta, phi)
        vol[z] = (r - (radius + pert)).astype(np.float32)
        vol[z] += rng.standard_normal((n, n), dtype=np.float32) * np.float32(noise)
    return vol


def large_f32():
    """512^3 float32 head, noise 8 (not rounded)."""
    return head(4, 8.0, _rng(4))


WORKLOADS = {
    'plain_u8_128': plain_u8,
    'lewiner_i16_256': lewiner_i16,
    'atlas_u16_256': atlas_u16,
    'sphere_f32_192': sphere_f32,
    'large_f32_512': large_f32,
}
