import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def pytest_collection_modifyitems(config, items):
    # tests/test_zz_* run last (pytest -x stops at the first failure)
    items.sort(key=lambda it: 1 if os.path.basename(str(it.fspath)).startswith("test_zz_") else 0)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ladybug_path(tmp_path_factory):
    import gzip
    import shutil
    out = tmp_path_factory.mktemp("bal") / "ladybug49.txt"
    with gzip.open(os.path.join(GOLDEN, "ladybug49.txt.gz"), "rb") as src, open(out, "wb") as dst:
        shutil.copyfileobj(src, dst)
    return str(out)


def theta_tolerance(cam, base):
    """Normwise Jacobian tolerance: `base`, widened inside (sqrt(eps), 1e-4) where the reference's
    forward-mode graph itself loses ~1e-2 * eps / theta through the 1 - cos(theta) cancellation."""
    th = float(np.linalg.norm(np.asarray(cam)[:3]))
    if th * th > 2.0 ** -52:
        return base + 0.1 * np.finfo(np.float64).eps / th
    return base


def tiny_problem(seed=0, nc=3, npts=12, noise=0.5, loss_pert=True):
    """Small random problem (all cameras see all points) for brute-force checks."""
    rng = np.random.default_rng(seed)
    cams = []
    for c in range(nc):
        aa = rng.normal(0, 0.3, 3)
        t = np.array([rng.normal(0, 0.5), rng.normal(0, 0.5), -10.0 + rng.normal(0, 0.5)])
        cams.append(np.concatenate([aa, t, [rng.normal(500, 20), rng.normal(0, 1e-3), rng.normal(0, 1e-3)]]))
    cams = np.array(cams)
    pts = rng.normal(0, 1.5, (npts, 3))
    ci = np.repeat(np.arange(nc), npts).astype(np.int32)
    pi = np.tile(np.arange(npts), nc).astype(np.int32)
    return cams, pts, ci, pi, rng
