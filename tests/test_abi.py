"""The C-ABI boundary (CPU): the HIP library loads, exports exactly the
entry points include/nii2mesh_amd.h declares, and without a device reports
failure instead of computing anything on the CPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import nii2mesh_amd
from nii2mesh_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'nii2mesh_amd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    names = []
    for decl in re.findall(r'^[A-Za-z_][^;#{}]*?\(.*?\)\s*;', text, flags=re.M | re.S):
        if decl.lstrip().startswith('typedef'):
            continue
        m = re.search(r'(\w+)\s*\(', decl)
        if m:
            names.append(m.group(1))
    return names


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ('meshify', 'laplacian_smoothHC', 'bwlabel', 'marchingCubes', 'setThreshold'):
        assert ref in names
    assert len(names) > 25


def test_library_exports_every_declared_symbol():
    L = nii2mesh_amd.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_only_the_abi_is_exported():
    out = subprocess.run(['nm', '-D', '--defined-only', nii2mesh_amd.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if ' T ' in l}
    assert exported == set(declared_functions())


def _has_device():
    try:
        return nii2mesh_amd.load().n2m_device_ok() == 1
    except Exception:
        return False


@pytest.mark.skipif(_has_device(), reason='a GPU is present: the no-device path cannot be exercised')
def test_no_device_is_a_reported_failure():
    L = nii2mesh_amd.load()
    assert L.n2m_device_ok() == 0
    assert nii2mesh_amd.last_error()
    img = np.zeros((8, 8, 8), np.float32)
    img[2:6, 2:6, 2:6] = 10
    rc, t, p = nii2mesh_amd.lib().meshify(img, 5.0)
    assert rc == 1 and len(t) == 0
    assert nii2mesh_amd.load().n2m_ctx_create() is None


def test_cubes_switch():
    L = nii2mesh_amd.load()
    prev = L.n2m_set_cubes(1)
    assert L.n2m_get_cubes() == 1
    L.n2m_set_cubes(0)
    assert L.n2m_get_cubes() == 0
    L.n2m_set_cubes(prev)
