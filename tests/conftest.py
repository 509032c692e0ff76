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
