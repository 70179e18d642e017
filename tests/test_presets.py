"""Initial potentials (Grid_3d::Table::randomReset, preset_sphericalPotential)
in the oracle, pinned by their invariants: Dirichlet nodes keep their
preset; random values lie in the configured range and repeat with the seed;
the ramp equals pmax at its centre and falls monotonically with distance."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import FIXTURE_SPECIES


def make(fx, mode):
    o = O.Oracle(mode, fx["xyz"], fx["ids"], fx["tets"], fx["ref_neigh"], O.species_table(FIXTURE_SPECIES))
    o.sync()
    return o


@pytest.mark.parametrize("mode", [O.F32, O.F64])
def test_random_reset(tip_small, mode):
    o = make(tip_small, mode)
    flags = o.flags()
    ids = tip_small["ids"]
    presets = {k: v.get("phi", 0.0) for k, v in FIXTURE_SPECIES.items()}
    lo, hi = min(presets.values()), max(presets.values())
    o.srand(7)
    o.random_reset()
    a = o.phi(0)
    assert np.array_equal(a, o.phi(1))
    d = (flags & 2) != 0
    assert np.array_equal(a[d], np.array([presets[i] for i in ids[d]], dtype=a.dtype))
    assert np.all(a[~d] >= lo) and np.all(a[~d] <= hi) and np.std(a[~d]) > 0.1 * (hi - lo)
    o.srand(7)
    o.random_reset()
    assert np.array_equal(o.phi(0), a)


@pytest.mark.parametrize("mode", [O.F32, O.F64])
def test_spherical_preset(tip_small, mode):
    o = make(tip_small, mode)
    flags = o.flags()
    xyz = tip_small["xyz"]
    free = np.nonzero((flags & 2) == 0)[0]
    c = xyz[free[len(free) // 2]].copy()
    o.spherical_preset(c, 100.0, 4000.0)
    a = o.phi(0)
    assert a[free[len(free) // 2]] == 4000.0
    dist = np.linalg.norm(xyz[free].astype(np.float64) - c, axis=1)
    order = np.argsort(dist, kind="stable")
    assert np.all(np.diff(a[free][order]) <= 1e-9 * 4000.0)   # non-increasing with distance
    corner = np.linalg.norm(xyz.min(0).astype(np.float64) - c)
    slope = (4000.0 - 100.0) / corner
    # float mode: the ramp is formed in float (a few ulp of 4000 V absolute)
    atol = 4000.0 * 5e-7 if mode == O.F32 else 4000.0 * 1e-12
    np.testing.assert_allclose(a[free], 4000.0 - slope * dist, rtol=1e-5, atol=atol)
