"""The reference CLI linked unchanged against the library (the drop-in
recipe; oracle/Makefile builds oracle/_ref/nii2mesh_gpu from the
reference's own nii2mesh.c): run as a child process with the README's flag
sets, its mesh files must equal, byte for byte, those of the reference CLI
built from its own sources (oracle/_ref/nii2mesh_{classic,lewiner}) --
load_nii, setThreshold, meshify, apply_sform, HC, the quadric reduction and
the writers (src/nii2mesh.c:321-356, src/meshify.c:984-1019).
"""
import os
import subprocess

import numpy as np
import pytest

from tests import support

pytestmark = pytest.mark.gpu

REF = os.path.join(support.ROOT, 'oracle', '_ref')


def run(exe, args, cwd, cubes=None):
    env = dict(os.environ)
    if cubes is not None:
        env['N2M_CUBES'] = cubes
    print('run', os.path.basename(exe), ' '.join(args), flush=True)
    r = subprocess.run([os.path.join(REF, exe)] + args, cwd=cwd, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (exe, args, r.stdout[-2000:], r.stderr[-2000:])
    return r


def outputs(d):
    return {f: open(os.path.join(d, f), 'rb').read() for f in sorted(os.listdir(d)) if not f.endswith('.nii')}


def compare_cli(tmp_path, inp, args, ref_exe, cubes, ext):
    a, b = tmp_path / 'ref', tmp_path / 'gpu'
    a.mkdir()
    b.mkdir()
    run(ref_exe, [inp] + args + ['out' + ext], str(a))
    run('nii2mesh_gpu', [inp] + args + ['out' + ext], str(b), cubes)
    ra, rb = outputs(str(a)), outputs(str(b))
    assert ra and sorted(ra) == sorted(rb)
    for f in ra:
        assert ra[f] == rb[f], '%s differs (%d vs %d bytes)' % (f, len(ra[f]), len(rb[f]))
    return ra


@pytest.fixture(scope='module')
def cli_built():
    for exe in ('nii2mesh_gpu', 'nii2mesh_classic', 'nii2mesh_lewiner'):
        if not os.path.exists(os.path.join(REF, exe)):
            pytest.fail('%s is not built (make -C oracle ref)' % exe)


BET_ARGS = [
    ['-p', '1', '-r', '1'],
    ['-i', '122', '-l', '0', '-b', '1', '-r', '1'],
    ['-i', '120', '-r', '1', '-p', '0', '-s', '100'],
    ['-i', 'b', '-r', '0.15'],
]


@pytest.mark.parametrize('k,ext', [(0, '.ply'), (0, '.gii'), (0, '.mz3'), (1, '.gii'), (2, '.ply'), (3, '.mz3')])
@pytest.mark.parametrize('build', ['classic', 'lewiner', 'lewiner_o1'])
def test_cli_bet(cli_built, tmp_path, build, k, ext):
    inp = os.path.join(support.DATA, 'bet.nii.gz')
    args = list(BET_ARGS[k])
    if build == 'lewiner_o1':
        args += ['-o', '1']
    ref_exe = 'nii2mesh_classic' if build == 'classic' else 'nii2mesh_lewiner'
    compare_cli(tmp_path, inp, args, ref_exe, 'classic' if build == 'classic' else 'lewiner', ext)


@pytest.mark.parametrize('build', ['classic', 'lewiner'])
def test_cli_atlas(cli_built, tmp_path, build):
    """-a 1 on a small float32 atlas: one file per label, each identical."""
    rng = np.random.default_rng(5)
    nz, ny, nx = 30, 34, 38
    seeds = rng.uniform(0, 1, (9, 3)) * np.array([nz, ny, nx])
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    d = ((z[..., None] - seeds[:, 0]) ** 2 + (y[..., None] - seeds[:, 1]) ** 2 + (x[..., None] - seeds[:, 2]) ** 2)
    lab = (np.argmin(d, -1) + 1).astype(np.float32)
    r = ((x - nx / 2) / (0.45 * nx)) ** 2 + ((y - ny / 2) / (0.45 * ny)) ** 2 + ((z - nz / 2) / (0.45 * nz)) ** 2
    lab[r > 1] = 0
    lab[lab == 4] = 0  # a label id without voxels
    inp = str(tmp_path / 'atlas.nii')
    srow = [[-1.5, 0.0, 0.0, 90.0], [0.0, 1.25, 0.1, -120.0], [0.0, 0.0, 1.1, -70.0]]
    support.write_nifti(inp, lab, 16, srow)
    ref_exe = 'nii2mesh_classic' if build == 'classic' else 'nii2mesh_lewiner'
    out = compare_cli(tmp_path, inp, ['-a', '1', '-r', '1', '-s', '5'], ref_exe, build, '.gii')
    assert len(out) == 8  # labels 1..9 less the empty one
