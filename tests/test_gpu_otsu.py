"""nii_otsu (src/isolevel.c:141-243) through the library: the three-threshold
search of modes 1 and 5 runs on the GPU, the others on the host.  Every
mode must return the oracle's thresholds, including histograms whose
objective has ties (the reference keeps the first maximum in (l, m, h)
order) and degenerate ones.  The oracle is pinned on the same histograms
against the reference's own nii_otsu (tests/test_oracle.py)."""
import pytest

from tests import support

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('mode', [1, 2, 3, 4, 5])
def test_otsu_matches_oracle(gpu, oracle, mode):
    for k, h in enumerate(support.otsu_histograms()):
        assert gpu.otsu(h, mode) == oracle.otsu(h, mode), 'histogram %d' % k
