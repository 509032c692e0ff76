import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X); run with -m gpu')
    config.addinivalue_line('markers', 'slow: long CPU test')


def pytest_collection_modifyitems(config, items):
    # GPU tests cannot run without a device; skip them loudly instead of failing in setup
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason='no HIP device visible')
    for it in items:
        if 'gpu' in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope='session')
def golden():
    import json
    with open(os.path.join(ROOT, 'tests', 'golden', 'ref_tables.json')) as f:
        return json.load(f)


def random_modulus(n, seed=2, Emin=1e-6, penal=3.0):
    """x ~ U(0,1) (seeded); E = Emin + x^p (1 - Emin) (the K*u micro-benchmark input)."""
    rng = np.random.default_rng(seed)
    x = rng.random(n)
    return Emin + x ** penal * (1 - Emin)


def holey_mask(nely, nelx, nelz, seed=0, frac=0.8):
    """Random non-cuboid mask whose clamped (x = 0) face and loaded (x = end) face stay solid."""
    rng = np.random.default_rng(seed)
    vox = rng.random((nely, nelx, nelz)) < frac
    vox[:, 0, :] = True
    vox[:, -1, :] = True
    return vox


def smooth_modulus(level1, per=1, Emin=1e-6, penal=3.0):
    """A smooth density field x = 0.12 + 0.05 sin(x/7) cos(y/5) sin(z/6 + 1) on the solid elements
    (the kind of field the optimisation produces; MG-PCG converges in ~10-20 iterations)."""
    e = level1.eleMapBack.astype(np.int64) - 1
    ey, ex, ez = e % level1.resY, (e // level1.resY) % level1.resX, e // (level1.resY * level1.resX)
    x = 0.12 + 0.05 * np.sin(ex / 7.0) * np.cos(ey / 5.0) * np.sin(ez / 6.0 + 1)
    E = Emin + x ** penal * (1 - Emin)
    if per > 1:
        E = np.repeat(E[None, :], per, axis=0)
    return E
