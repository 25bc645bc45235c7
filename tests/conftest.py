import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')
MESH_600 = os.path.join(GOLDEN, 'meshes', '4testing-600_120.mail')
MESH_940 = os.path.join(GOLDEN, 'meshes', '4testing-940_266_sq.mail')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')


@pytest.fixture(scope='session')
def mats():
    with open(os.path.join(GOLDEN, 'materials.json')) as f:
        d = json.load(f)

    def get(name, wl):
        v = d[name][repr(float(wl))]
        return complex(v[0], v[1])
    return get


def k_parallel_norm(wl, theta, phi, period, n_inc=1.0):
    """Light(theta, phi).k_pll * period with the 1e-5 degree nudge
    (reference objects.py:1736-1743)."""
    if abs(theta) < 1e-5:
        theta += 1e-5
    if abs(phi) < 1e-5:
        phi += 1e-5
    k = 2 * np.pi * n_inc / wl
    th, ph = np.deg2rad(theta), np.deg2rad(phi)
    return k * np.sin(th) * np.array([np.cos(ph), np.sin(ph)]) * period


def load_txt(case, name):
    return np.loadtxt(os.path.join(GOLDEN, case, name))


@pytest.fixture(scope='session')
def gpu_ctx():
    from emustack_amd import _lib
    return _lib.Context(0)
