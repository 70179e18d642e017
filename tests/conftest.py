import faulthandler
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# a crash in native code prints the Python stack of every thread, naming the test
faulthandler.enable(all_threads=True)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


# species of the committed fixtures (tests/golden/make_golden.py)
FIXTURE_SPECIES = {
    0: dict(),
    1: dict(dirichlet=1),
    2: dict(neumann=1),
    3: dict(dirichlet=1),
    10: dict(epsilon=1e6, removable=1, dirichlet=1, phi=5000.0, mass=55.845, evap_charge=2, evap_field=3.0e10,
             evap_energy=1.0),
    11: dict(epsilon=1e6, removable=1, dirichlet=1, phi=5000.0, mass=58.69, evap_charge=2, evap_field=2.5e10,
             evap_energy=1.0),
}


def coordination_species():
    """FIXTURE_SPECIES plus ids 12..127 as matrix atoms whose evaporation field
    grows with the id (the species table of the update_coordination fixture;
    values as the fixture's configuration text gives them)"""
    sp = {k: dict(v) for k, v in FIXTURE_SPECIES.items()}
    for k in range(12, 128):
        field = float(np.float32(float("%.8e" % (3.0e10 * (1.0 + 0.002 * (k - 10))))))
        sp[k] = dict(epsilon=1e6, removable=1, dirichlet=1, phi=5000.0, mass=55.845, evap_charge=2, evap_field=field,
                     evap_energy=1.0)
    return sp


@pytest.fixture(scope="session")
def tip_small():
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "tip_small.npz")))


@pytest.fixture(scope="session")
def tetgen_random():
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "tetgen_random.npz")))


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False
