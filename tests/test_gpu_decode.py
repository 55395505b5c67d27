"""Device-side loader conversion (src/nii2mesh.c:155-172) against the oracle,
bit for bit, for every datatype the CLI reads."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('dtype,code', [(np.uint8, 2), (np.int16, 4), (np.uint16, 512), (np.float32, 16)])
@pytest.mark.parametrize('slope,inter', [(1.0, 0.0), (0.0, 0.0), (0.003921569, -0.5), (-2.75, 1000.125)])
def test_decode_matches_oracle(gpu, oracle, dtype, code, slope, inter):
    rng = np.random.default_rng(code)
    if dtype == np.float32:
        raw = rng.normal(0, 1000, (17, 19, 23)).astype(np.float32)
    else:
        info = np.iinfo(dtype)
        raw = rng.integers(info.min, info.max, (17, 19, 23), endpoint=True).astype(dtype)
    got = gpu.decode(raw, code, slope, inter)
    rc, want = oracle.decode(raw, code, slope, inter)
    assert rc == 0 and np.array_equal(got.ravel(), want)


def test_decode_rejects_other_types(gpu):
    with pytest.raises(RuntimeError):
        gpu.decode(np.zeros(8, np.int32), 8)
