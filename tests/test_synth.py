"""The synthetic workloads meet their stated shapes and statistics."""
import numpy as np


def test_atlas_label_sizes_log_normal():
    """256^3 atlas: 370 labels, all present, sizes log-normal with median
    about 2k voxels and a largest label of a few 1e5 voxels."""
    from nii2mesh_amd import synth
    vol = synth.atlas_u16()
    assert vol.shape == (256, 256, 256) and vol.dtype == np.float32
    lab = vol.astype(np.int64).ravel()
    assert np.array_equal(vol.ravel(), lab.astype(np.float32))  # integer labels
    sz = np.bincount(lab, minlength=371)[1:]
    assert len(sz) == 370 and sz.min() > 0
    assert 1500 <= np.median(sz) <= 2500, np.median(sz)
    assert 250_000 <= sz.max() <= 500_000, sz.max()
    ls = np.log(sz)
    # log sizes: spread of the drawn log-normal (sigma 1.68), little skew
    assert 1.4 <= ls.std() <= 2.0, ls.std()
    assert abs(np.mean(ls) - np.median(ls)) < 0.3
