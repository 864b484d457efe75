import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _have_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


# ---- lensing fixtures: the reference's mock cosmology and five shells
# (tests/fixtures/domain.py:36-143), restated


class MockCosmology:
    """Omega_m0 = 0.3, Hubble distance 4400 Mpc, E(z) of flat LCDM and the
    dimensionless transverse comoving distance x_M(z) = 1000 z (or 1000 (z2 - z))."""

    Omega_m0 = 0.3
    hubble_distance = 4.4e3

    def H_over_H0(self, z):  # noqa: N802
        return (self.Omega_m0 * (1 + z) ** 3 + 1 - self.Omega_m0) ** 0.5

    def xm(self, z, z2=None):
        if z2 is None:
            return np.asarray(z) * 1_000
        return (np.asarray(z2) - np.asarray(z)) * 1_000

    def transverse_comoving_distance(self, z, z2=None):
        return self.hubble_distance * self.xm(z, z2)


class Window:
    def __init__(self, za, wa, zeff):
        self.za, self.wa, self.zeff = np.asarray(za), np.asarray(wa), zeff


@pytest.fixture
def cosmo():
    return MockCosmology()


@pytest.fixture
def shells():
    return [Window([i, i + 1.0, i + 2.0], [0.0, 1.0, 0.0], i + 1.0) for i in range(5)]
