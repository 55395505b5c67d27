"""n2m_set_thresholds_noalloc (no allocation; one fixed device array per
device) against n2m_ctx_set_thresholds, through the C ABI: two calls in a
row, the second on a smaller volume, must give the same isolevels bit for
bit; sizes beyond the stated limits are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu


def _bind(L):
    L.n2m_set_thresholds_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_float), C.c_void_p]
    L.n2m_set_thresholds_noalloc.restype = C.c_int
    L.n2m_ctx_set_thresholds.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_float)]
    return L


def _both(gpu, torch, vol, modes, stream):
    L = _bind(gpu.load())
    d = torch.from_numpy(vol).cuda()
    torch.cuda.synchronize()
    n = len(modes)
    cm = (C.c_int * n)(*modes)
    a, b = (C.c_float * n)(), (C.c_float * n)()
    assert L.n2m_set_thresholds_noalloc(C.c_void_p(d.data_ptr()), d.numel(), n, cm, a, stream) == 0, \
        gpu.last_error()
    with gpu.DeviceContext() as ctx:
        assert L.n2m_ctx_set_thresholds(ctx.h, C.c_void_p(d.data_ptr()), d.numel(), n, cm, b) == 0
    return np.array(a[:], np.float32), np.array(b[:], np.float32)


@pytest.mark.parametrize('own_stream', [False, True])
def test_noalloc_thresholds_match_context_entry(gpu, own_stream):
    """On the default stream, and on a stream of the caller's own."""
    import torch
    side = torch.cuda.Stream() if own_stream else None  # default priority
    stream = C.c_void_p(side.cuda_stream) if own_stream else None
    big = support.blobs((72, 80, 88), 21, noise=3.0)
    small = support.blobs((20, 24, 28), 22, noise=1.0)
    for vol, modes in ((big, [1, 2, 3]), (small, [3, 1, 2, 2])):
        a, b = _both(gpu, torch, vol, modes, stream)
        print('noalloc', a, 'context', b, flush=True)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_edge_volumes(gpu):
    """Constant, tiny (fewer than 100 voxels), NaN-holding and all-NaN
    volumes take the range fallbacks of src/isolevel.c:35-139."""
    import torch
    cases = [np.full((6, 7, 8), 3.0, np.float32), support.blobs((3, 4, 5), 2),
             support.nan_volume((20, 22, 24), 5), np.full((4, 4, 4), np.nan, np.float32)]
    for k, vol in enumerate(cases):
        a, b = _both(gpu, torch, vol, [1, 2, 3], None)
        assert support.nan_eq(a, b), k


def test_noalloc_refuses_beyond_limits(gpu):
    """Refused on the host before any GPU work (the volume is never read)."""
    import torch
    L = _bind(gpu.load())
    d = torch.zeros(128, dtype=torch.float32, device='cuda')
    torch.cuda.synchronize()
    p = C.c_void_p(d.data_ptr())
    out = (C.c_float * 4)()
    one = (C.c_int * 1)(2)
    assert L.n2m_set_thresholds_noalloc(p, 0, 1, one, out, None) == 1
    assert 'voxels' in gpu.last_error()
    assert L.n2m_set_thresholds_noalloc(p, 1 << 32, 1, one, out, None) == 1
    assert 'voxels' in gpu.last_error()
    assert L.n2m_set_thresholds_noalloc(p, 100, 17, (C.c_int * 17)(*[2] * 17), out, None) == 1
    assert 'modes' in gpu.last_error()
    assert L.n2m_set_thresholds_noalloc(p, 100, 1, (C.c_int * 1)(4), out, None) == 1
    assert 'mode 4' in gpu.last_error()
