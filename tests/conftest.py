"""Shared test setup.

`gpu` marks tests that need an MI355X; the CPU suite runs with -m "not gpu".
The product package lives in the directory laplace-fmm_amd/ (not a valid
Python identifier) and is imported here as `laplace_fmm_amd`; the oracle is
the CPU checker under oracle/.
"""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_product():
    name = "laplace_fmm_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "laplace-fmm_amd", "__init__.py"),
                                                  submodule_search_locations=[os.path.join(ROOT, "laplace-fmm_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def lfmm():
    return load_product()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def uniform_cube(n, seed, dim=3):
    """Seeded synthetic input: positions U[0,1)^dim drawn before charges U[-1,1]."""
    rng = np.random.default_rng(seed)
    x = rng.random((n, dim))
    q = rng.uniform(-1.0, 1.0, n)
    return x, q


def plummer(n, seed, a=0.05, rmax=1.0):
    """Seeded Plummer sphere: scale radius a, truncated at radius rmax (radii
    redrawn beyond it), centred at (1/2, 1/2, 1/2); charges U[-1,1].  Radii
    from the inverse of the Plummer mass profile, r = a / sqrt(u^(-2/3) - 1),
    with isotropic directions.  About 1% of points fall outside the unit cube;
    they are kept (the tree's bounding box grows to hold them)."""
    rng = np.random.default_rng(seed)
    r = np.empty(0)
    while len(r) < n:
        u = rng.random(2 * n)
        rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = np.concatenate([r, rr[rr <= rmax]])
    r = r[:n]
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    x = 0.5 + r[:, None] * d
    q = rng.uniform(-1.0, 1.0, n)
    return x, q


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
