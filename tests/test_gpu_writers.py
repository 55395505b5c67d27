"""Binary mesh files written with GPU-encoded byte streams must equal the
reference's save_mesh() output (src/meshify.c:602-930,984-1019) byte for
byte: .mz3 and .gii with and without compression, .ply, and the default
(unknown extension -> .mz3 appended)."""
import os

import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu


def mesh(oracle, seed):
    img = support.blobs((30, 32, 28), seed)
    rc, t, p = oracle.meshify(img, float(np.percentile(img, 60)), True, True, False, cubes=1)
    assert rc == 0
    return t, p


@pytest.mark.parametrize('name,gz', [('m.ply', False), ('m.gii', False), ('m.gii', True), ('m.mz3', False),
                                     ('m.mz3', True), ('m.surf', False), ('m', True)])
def test_writer_bytes_match_reference(gpu, oracle, ref_classic, tmp_path, name, gz):
    t, p = mesh(oracle, 5)
    a, b = str(tmp_path / ('gpu_' + name)), str(tmp_path / ('ref_' + name))
    assert gpu.save_mesh(a, t, p, gz) == 0
    assert ref_classic.save_mesh(b, t, p, gz) == 0
    if not os.path.exists(a):  # no known extension: ".mz3" appended
        a, b = a + '.mz3', b + '.mz3'
    assert open(a, 'rb').read() == open(b, 'rb').read()


def test_base64_matches_reference_and_rfc(gpu, oracle, ref_classic, tmp_path):
    """RFC 4648 test vectors through the GIfTI path's encoder: a 1..3-byte
    tail must pad like the reference's base64_encode."""
    for raw, enc in [(b'f', b'Zg=='), (b'fo', b'Zm8='), (b'foo', b'Zm9v'), (b'foobar', b'Zm9vYmFy')]:
        assert ref_classic.base64(raw) == enc
    # triangle counts giving 12*nt % 3 == 0 and float arrays of any length
    for nt in (1, 2, 7):
        t = np.arange(3 * nt, dtype=np.int32).reshape(nt, 3)
        p = np.linspace(-3, 5, 3 * (nt + 1)).reshape(-1, 3)
        a, b = str(tmp_path / 'a.gii'), str(tmp_path / 'b.gii')
        assert gpu.save_mesh(a, t, p, False) == 0 and ref_classic.save_mesh(b, t, p, False) == 0
        assert open(a, 'rb').read() == open(b, 'rb').read()


def test_text_formats_refused(gpu, tmp_path):
    t = np.zeros((1, 3), np.int32)
    p = np.zeros((3, 3))
    assert gpu.save_mesh(str(tmp_path / 'x.obj'), t, p) != 0
    assert 'save_mesh' in gpu.last_error()
