"""Mesh files written by the library must equal the reference's save_mesh()
output (src/meshify.c:418-1019) byte for byte: .mz3 and .gii with and
without compression, .ply, .stl, .vtk, the FreeSurfer surfaces (.pial,
.inflated: all but the time stamp line), the text formats .obj, .off and
.json, and the default (unknown extension -> .mz3 appended)."""
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
                                     ('m.mz3', True), ('m.surf', False), ('m', True), ('m.stl', False),
                                     ('m.vtk', False), ('m.obj', False), ('m.off', False), ('m.json', False),
                                     ('m.obj', True)])
def test_writer_bytes_match_reference(gpu, oracle, ref_classic, tmp_path, name, gz):
    t, p = mesh(oracle, 5)
    a, b = str(tmp_path / ('gpu_' + name)), str(tmp_path / ('ref_' + name))
    assert gpu.save_mesh(a, t, p, gz) == 0
    rc = ref_classic.save_mesh(b, t, p, gz)
    if not name.endswith('.stl'):  # the reference's save_stl returns no value
        assert rc == 0
    if not os.path.exists(a):  # no known extension: ".mz3" appended
        a, b = a + '.mz3', b + '.mz3'
    assert open(a, 'rb').read() == open(b, 'rb').read()


@pytest.mark.parametrize('name', ['m.pial', 'm.inflated'])
def test_freesurfer_matches_reference(gpu, oracle, ref_classic, tmp_path, name):
    """FreeSurfer binary surfaces (src/meshify.c:418-456): magic, a time
    stamp line ending in two newlines, big-endian counts, float32 vertices
    and triangles.  All bytes but the stamp's text must match."""
    t, p = mesh(oracle, 6)
    a, b = str(tmp_path / ('gpu_' + name)), str(tmp_path / ('ref_' + name))
    assert gpu.save_mesh(a, t, p) == 0 and ref_classic.save_mesh(b, t, p, False) == 0
    ra, rb = open(a, 'rb').read(), open(b, 'rb').read()
    ea, eb = ra.index(b'\n\n') + 2, rb.index(b'\n\n') + 2
    assert ra[:3] == rb[:3] == b'\xff\xff\xfe'
    assert ra[3:ea].startswith(b'created by niimath on ') and rb[3:eb].startswith(b'created by niimath on ')
    assert ra[ea:] == rb[eb:]
    assert len(ra) - ea == 8 + 12 * len(p) + 12 * len(t)


def test_text_writers_large_mesh(gpu, oracle, ref_classic, tmp_path):
    """More rows than one formatting chunk (the rows are formatted by host
    threads in parallel and written in order), with negative and large
    coordinates."""
    rng = np.random.default_rng(3)
    p = rng.normal(0, 300, (70000, 3)) * rng.choice([1e-6, 1.0, 1e5], (70000, 3))
    t = rng.integers(0, len(p), (90000, 3), dtype=np.int32)
    for ext in ('.obj', '.off', '.json'):
        a, b = str(tmp_path / ('gpu' + ext)), str(tmp_path / ('ref' + ext))
        assert gpu.save_mesh(a, t, p) == 0 and ref_classic.save_mesh(b, t, p, False) == 0
        assert open(a, 'rb').read() == open(b, 'rb').read(), ext


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
