"""nii_otsu (src/isolevel.c:141-243) through the library: the three-threshold
search of modes 1 and 5 runs on the GPU, the others on the host.  Every
mode must return the oracle's thresholds, including histograms whose
objective has ties (the reference keeps the first maximum in (l, m, h)
order) and degenerate ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def histograms():
    rng = np.random.default_rng(77)
    out = [rng.integers(0, 5000, 256), np.ones(256, int), np.zeros(256, int)]
    h = np.zeros(256, int)
    h[40] = 1000
    out.append(h)  # a single spike
    h = np.zeros(256, int)
    h[[10, 80, 150, 220]] = 500
    out.append(h)  # four equal spikes: many equal objective values
    sym = rng.integers(0, 100, 128)
    out.append(np.concatenate([sym, sym[::-1]]))  # mirror-symmetric
    x = np.arange(256)
    out.append((1000 * (np.exp(-(x - 60) ** 2 / 200) + np.exp(-(x - 190) ** 2 / 300))).astype(int))
    for k in range(6):
        out.append(rng.poisson(rng.uniform(1, 300), 256) * (rng.uniform(size=256) < 0.7))
    return [np.ascontiguousarray(h, np.int32) for h in out]


@pytest.mark.parametrize('mode', [1, 2, 3, 4, 5])
def test_otsu_matches_oracle(gpu, oracle, mode):
    for k, h in enumerate(histograms()):
        assert gpu.otsu(h, mode) == oracle.otsu(h, mode), 'histogram %d' % k
