"""Surface normal (Surface_3d::Table::normal, surface_3d.cpp:123-169, with
Geometry_3d::find_best_plane, geometry_3d.cpp:1806-1870): the oracle's
restatement against closed forms of the eigenproblem.  Its bits are pinned by
the reference's own event normals in test_oracle_golden."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import FIXTURE_SPECIES


def test_plane_z0_gives_z_axis():
    rng = np.random.default_rng(1)
    off = np.zeros((40, 3), np.float32)
    off[:, :2] = rng.standard_normal((40, 2)).astype(np.float32) * 1e-10
    n = O.best_plane(off)
    assert abs(abs(n[2]) - 1.0) <= 1e-6 and abs(n[0]) <= 1e-6 and abs(n[1]) <= 1e-6


def test_no_offsets_gives_first_axis():
    # a zero scatter matrix: tred2 leaves the identity, every eigenvalue is 0,
    # and the first minimum picks column 0
    assert np.array_equal(O.best_plane(np.zeros((0, 3), np.float32)), np.array([1, 0, 0], np.float32))


@pytest.mark.parametrize("seed", range(6))
def test_tilted_planes_and_clouds_match_eigh(seed):
    rng = np.random.default_rng(100 + seed)
    # a plane with a random normal and slight noise: the normal is recovered
    nrm = rng.standard_normal(3)
    nrm /= np.linalg.norm(nrm)
    basis = np.linalg.svd(np.eye(3) - np.outer(nrm, nrm))[0][:, :2]
    pts = (rng.standard_normal((60, 2)) @ basis.T + 1e-4 * rng.standard_normal((60, 1)) * nrm) * 3e-10
    n = O.best_plane(pts.astype(np.float32)).astype(np.float64)
    assert abs(np.linalg.norm(n) - 1.0) < 1e-6
    assert abs(abs(n @ nrm) - 1.0) < 1e-5
    # an anisotropic cloud: the smallest-eigenvalue eigenvector of the FP64
    # scatter matrix (LAPACK) up to sign
    cloud = (rng.standard_normal((80, 3)) * [3.0, 2.0, 0.5]) @ np.linalg.qr(rng.standard_normal((3, 3)))[0]
    cloud = (cloud * 1e-10).astype(np.float32)
    w, v = np.linalg.eigh(cloud.astype(np.float64).T @ cloud.astype(np.float64))
    n = O.best_plane(cloud).astype(np.float64)
    assert abs(abs(n @ v[:, 0]) - 1.0) < 1e-5


def test_normals_point_to_vacuum(tip_small):
    o = O.Oracle(O.F32, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(FIXTURE_SPECIES))
    o.sync()
    o.surface_init(0)
    xyz = tip_small["xyz"].astype(np.float64)
    surf = o.surface_nodes()
    ptr, idx = o.csr()
    for node in surf[:: max(1, len(surf) // 40)]:
        n = o.surface_normal(int(node)).astype(np.float64)
        assert abs(np.linalg.norm(n) - 1.0) < 1e-6
        # vacuum nodes within three hops, as in the reference
        shell = {int(node)}
        front = {int(node)}
        for _ in range(3):
            front = {int(v) for u in front for v in idx[ptr[u]:ptr[u + 1]]} - shell
            shell |= front
        shell.discard(int(node))
        vac = [v for v in shell if tip_small["ids"][v] == 0]
        toward = (xyz[vac] - xyz[int(node)]).sum(0)
        assert n @ toward >= -1e-6 * np.linalg.norm(toward)
    # a vacuum node is not a surface site: zero vector
    vac_node = int(np.nonzero(tip_small["ids"] == 0)[0][0])
    assert np.array_equal(o.surface_normal(vac_node), np.zeros(3, np.float32))


@pytest.mark.parametrize("seed", range(4))
def test_library_best_plane_equals_oracle(seed):
    """The product's host restatement (csrc/host/normal.cpp, through the C
    ABI) and the oracle's give the same bits."""
    import tapsim_modified_amd as tsg
    rng = np.random.default_rng(seed)
    for k in (0, 1, 2, 3, 7, 40, 200):
        off = (rng.standard_normal((k, 3)) * [2.0, 1.0, 0.3] * 1e-10).astype(np.float32)
        a, b = tsg.best_plane(off), O.best_plane(off)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (k, a, b)
