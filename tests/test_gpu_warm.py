"""The learned warm start in the event loop, through the C ABI: events record
their shell graphs, the network trains on the GPU, and from the same state
the warm-started relaxation reaches the same residual target with fewer
correction steps while the evaporation sequence stays the same."""
import sys

import numpy as np
import pytest

import tapsim_modified_amd as tsg

pytestmark = pytest.mark.gpu

SPECIES = {
    0: dict(), 1: dict(dirichlet=1), 2: dict(neumann=1), 3: dict(dirichlet=1),
    10: dict(removable=1, dirichlet=1, phi=5000.0, mass=55.845, evap_charge=2, evap_field=3.0e10, evap_energy=1.0),
}


def log(msg):
    print("[gpu-warm] " + msg, flush=True)
    sys.stdout.flush()


def small_tip():
    p = tsg.SpecimenParams()
    p.a, p.radius, p.shank = 0.287e-9, 2.5e-9, 3e-9
    for k, v in enumerate((20e-9, 20e-9, 40e-9)):
        p.box[k] = v
    p.shell, p.vacuum_target, p.decay, p.face_spacing = 2.0, 20000, 3.0, 4e-9
    p.cluster_density, p.cluster_radius, p.seed = 0.0, 2e-9, 77
    p.id_vacuum, p.id_top, p.id_bottom, p.id_sides, p.id_matrix, p.id_solute = 0, 1, 2, 3, 10, 10
    xyz, ids, _ = tsg.generate_specimen(p)
    return xyz, ids


def options(method):
    e = tsg.EvapOptions()
    e.prob_mode, e.evap_mode, e.vacuum_id = 0, 0, 0
    e.relax_threshold, e.relax_shell, e.relax_cycle, e.relax_queue = 1e-2, 3, 5, 5
    e.relax_global_cycles, e.refresh_interval, e.refresh_threshold = 1, 1, 1e-8
    e.exact_probability = 1
    e.method, e.accel_tol = method, 1e-8
    return e


def test_learned_warm_start_gain():
    xyz, ids = small_tip()
    mesh = tsg.Mesh(xyz)
    species = tsg.species_table(SPECIES)
    grids = [tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64) for _ in range(2)]
    for g in grids:
        g.accel_setup(0.5)
        g.accel_relax(1e-9)
    learned, baseline = grids
    assert np.array_equal(learned.phi(0), baseline.phi(0))
    net = tsg.WarmStartNet(max_batch_rows=32768, data_rows=1 << 20, data_edges=1 << 24, seed=1)
    learned.attach(net, record=True, predict=False)
    n_train = 160
    ev_l = learned.evaporate(options(1), n_train)
    ev_b = baseline.evaporate(options(1), n_train)
    assert [e.node for e in ev_l] == [e.node for e in ev_b]
    assert np.array_equal(learned.phi(0), baseline.phi(0))   # deterministic device path
    graphs, rows, edges = net.dataset()
    log("recorded %d graphs, %d rows, %d edges" % (graphs, rows, edges))
    assert graphs == n_train
    untrained = net.evaluate(n_train - 20, n_train)
    first, last = net.train(0, n_train - 20, epochs=80, batch_graphs=16, lr=2e-3, seed=5)
    held = net.evaluate(n_train - 20, n_train)
    log("train loss %.4g -> %.4g, held-out %.4g (untrained %.4g)" % (first, last, held, untrained))
    assert held < 0.5 * untrained

    learned.attach(net, record=False, predict=True)
    k = 40
    ev_l = learned.evaporate(options(2), k)
    ev_b = baseline.evaporate(options(1), k)
    steps_l = sum(e.global_iter for e in ev_l)
    steps_b = sum(e.global_iter for e in ev_b)
    used = sum(1 for e in ev_l if e.warm_start == 1)
    ratio = np.median([e.residual_warm / e.residual_cold for e in ev_l if e.residual_cold > 0])
    log("correction steps: learned %d, without %d (%.1f%% fewer); prediction kept in %d/%d events; "
        "median residual ratio %.3g" % (steps_l, steps_b, 100.0 * (1 - steps_l / steps_b), used, k, ratio))
    assert [e.node for e in ev_l] == [e.node for e in ev_b]
    assert used >= k // 2
    assert steps_l < steps_b
