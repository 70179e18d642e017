"""GPU parity on the BASELINE workloads through the C ABI:

* c1 (BCC [110] tip r 5 nm, 10 nm shank + cap, 60x60x120 nm graded vacuum):
  the plain FP64 path against the CPU oracle event by event (bit-exact), the
  committed 1000-event plain sequence (tests/golden/c1_sequence.npz, written by
  tests/golden/make_config_sequences.py) reproduced, and the mixed and
  learned+mixed paths keeping that sequence for all 1000 events;
* c3 (the c1 geometry, two-phase: solute at 25 V/nm in 2 nm clusters): the
  mixed path's sequence equal to the plain path's, its potentials within
  1e-9 relative of the CPU oracle's tightly relaxed solution;
* c2 (r 15 nm, 40 nm shank, 3.0 M atoms + 1.5 M vacuum nodes): ten events of
  the plain, mixed and learned+mixed paths with one sequence, the mixed
  path's final residual checked by the oracle's FP64 sweep on the CPU.
"""
import os
import sys

import numpy as np
import pytest

import bench
import tapsim_modified_amd as tsg
from oracle import oracle as O
from tests.conftest import ROOT
from tests.golden.make_config_sequences import learned_run, options, relaxed_state

pytestmark = pytest.mark.gpu


def log(msg):
    print("[gpu-configs] " + msg, flush=True)
    sys.stdout.flush()


def oracle_threads():
    return max(1, min(16, bench.host_cores()))


@pytest.fixture(scope="module")
def c1():
    log("building c1")
    xyz, ids, counts, mesh, species, state0 = relaxed_state(tsg, "c1")
    log("c1: %d nodes %s" % (len(xyz), counts))
    return dict(xyz=xyz, ids=ids, counts=counts, mesh=mesh, species=species, state0=state0)


@pytest.fixture(scope="module")
def c1_seq():
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "c1_sequence.npz")))


def grid_from(w, phi=None):
    g = tsg.Grid3D(w["mesh"], w["ids"], w["species"], 0.0, tsg.PLAIN_F64)
    g.set_phi(*(phi if phi is not None else w["state0"]))
    return g


def test_c1_plain_fp64_matches_the_oracle(c1):
    """three plain events on the c1 tip, GPU against the oracle (its refresh
    relaxation on worker threads with the reference's chunking, and its
    flights' step-control root correctly rounded as the device's): the chosen
    nodes, the relaxations' iteration counts, the record, the flights' end
    points and both potential buffers bit-exact"""
    k = 3
    g = grid_from(c1)
    tsg.srand(99)
    eg = g.evaporate(options(0), k)
    tv, tn = c1["mesh"].tets()
    o = O.Oracle(O.F64, c1["xyz"], c1["ids"], tv, tn, O.species_table(bench.SPECIES))
    o.sync()
    o.set_phi(*c1["state0"])
    o.set_threads(oracle_threads())
    o.srand(99)
    try:
        O.set_root_mode(1)
        eo = o.evaporate(bench.oracle_options(O, options(0)), k)
    finally:
        O.set_root_mode(0)
    log("c1 plain: GPU %s oracle %s, refresh iterations %s" % ([e.node for e in eg], [e.node for e in eo],
                                                             [e.refresh_iter for e in eo]))
    assert [e.node for e in eg] == [e.node for e in eo]
    assert [e.local_iter for e in eg] == [e.local_iter for e in eo]
    assert [e.global_iter for e in eg] == [e.refresh_iter for e in eo]
    f32 = lambda v: np.asarray(v, np.float32).view(np.uint32)  # noqa: E731 (the oracle's record is float)
    assert np.array_equal(f32([e.pot_after for e in eg]), f32([e.pot_after for e in eo]))
    assert np.array_equal(f32([list(e.field_before) for e in eg]), f32([list(e.field_before) for e in eo]))
    assert np.array_equal(f32([list(e.stop) for e in eg]), f32([list(e.stop) for e in eo]))
    assert g.switches()[0] == o.switches()[0]
    assert np.array_equal(g.phi(0).view(np.uint64), o.phi(0).view(np.uint64))
    assert np.array_equal(g.phi(1).view(np.uint64), o.phi(1).view(np.uint64))


def test_c1_plain_reproduces_the_committed_sequence(c1, c1_seq):
    n = 60
    g = grid_from(c1)
    tsg.srand(99)
    ev = g.evaporate(options(0), n)
    assert np.array_equal([e.node for e in ev], c1_seq["plain_node"][:n])
    assert np.array_equal([[e.local_iter, e.global_iter] for e in ev], c1_seq["plain_iter"][:n])


def test_c1_device_surface_bookkeeping(c1):
    """the surface set kept on the device over 40 events of the mixed path,
    one event per call: the site list (rebuilt on the device after every
    take-away) equals the oracle's Surface_3d set driven with the same atoms;
    the record's apex is the first highest site of the list before the
    event, and its normal (the device kernel) is bit-identical to the host
    restatement's normal of the chosen site on that list (the host
    restatement is pinned to the reference's normals by the fixture tests)"""
    g = grid_from(c1)
    g.accel_setup(0.9)
    tv, tn = c1["mesh"].tets()
    o = O.Oracle(O.F64, c1["xyz"], c1["ids"], tv, tn, O.species_table(bench.SPECIES))
    o.sync()
    o.surface_init(0)
    g.surface_init(0)
    xyz = c1["xyz"]
    tsg.srand(99)
    opts = options(1)
    sizes = []
    for k in range(40):
        sites = g.surface_nodes()
        assert np.array_equal(sites, o.surface_nodes())
        g.surface_evaluate(True, 0)
        node, _ = g.candidate()
        normal = g.surface_normal(node)
        apex = xyz[sites[np.argmax(xyz[sites, 2])]]
        (e,) = g.evaporate(opts, 1)
        assert e.node == node
        assert np.array_equal(np.asarray(e.normal, np.float32).view(np.uint32), normal.view(np.uint32))
        assert np.array_equal(np.asarray(e.apex, np.float32).view(np.uint32), apex.view(np.uint32))
        o.take_away(e.node)
        assert e.surface_size == len(o.surface_nodes())
        sizes.append(e.surface_size)
    assert np.array_equal(g.surface_nodes(), o.surface_nodes())
    log("c1 surface: %d events, sites %d -> %d" % (len(sizes), sizes[0], sizes[-1]))


def test_c1_mixed_and_learned_keep_the_plain_sequence_for_1000_events(c1, c1_seq):
    """the mixed FP32/FP64 path (multigrid correction solve, residual <= 1e-8 V)
    and the learned warm start (trained on its first 200 events) evaporate the
    plain FP64 path's 1000 atoms in its order"""
    n = 1000
    assert len(c1_seq["plain_node"]) == n
    g = grid_from(c1)
    g.accel_setup(0.9)
    tsg.srand(99)
    me = []
    while len(me) < n:
        me += list(g.evaporate(options(1), min(250, n - len(me))))
    assert all(e.surface_size > 0 for e in me)
    mixed = np.array([e.node for e in me])
    first = np.where(mixed != c1_seq["plain_node"])[0]
    log("c1 mixed: first difference at %s" % (first[0] if len(first) else None))
    assert len(first) == 0
    tsg.srand(99)
    _, le = learned_run(c1["mesh"], c1["ids"], c1["species"], c1["state0"], n)
    learned = np.array([e.node for e in le])
    first = np.where(learned != c1_seq["plain_node"])[0]
    kept = sum(1 for e in le if e.warm_start == 1)
    log("c1 learned: first difference at %s; prediction kept in %d of %d events" % (
        first[0] if len(first) else None, kept, n - 200))
    assert len(first) == 0
    assert kept >= (n - 200) // 2


def test_c3_two_phase_mixed_against_the_oracle():
    """c3: the c1 tip with solute clusters.  Mixed and plain evaporate the
    same 100 atoms; the mixed potentials lie within 1e-9 relative of the
    oracle's FP64 Jacobi relaxation of the same system continued from them
    to a deviation of 1e-11 V on the CPU"""
    xyz, ids, counts, mesh, species, state0 = relaxed_state(tsg, "c3")
    atoms = counts["atoms"]
    log("c3: %d atoms, %d solute (%.2f%%)" % (atoms, counts["solute"], 100.0 * counts["solute"] / max(1, atoms)))
    assert counts["solute"] > 0
    w = dict(mesh=mesh, ids=ids, species=species, state0=state0)
    n = 100
    p = grid_from(w)
    tsg.srand(99)
    pe = p.evaporate(options(0), n)
    m = grid_from(w)
    m.accel_setup(0.9)
    tsg.srand(99)
    me = m.evaporate(options(1), n)
    assert [e.node for e in pe] == [e.node for e in me]
    log("c3: %d of %d evaporated atoms are solute" % (sum(int(ids[e.node]) == 11 for e in me), n))
    phi_m = m.phi(m.switches()[0])
    phi_p = p.phi((p.switches()[0] + 1) % 2)  # the buffer potential() reads
    tv, tn = mesh.tets()
    o = O.Oracle(O.F64, xyz, m.ids(), tv, tn, O.species_table(bench.SPECIES))
    o.sync()
    o.set_threads(oracle_threads())
    o.set_phi(phi_m, phi_m, 0, 0)
    res = o.relax(1)
    o.set_phi(phi_m, phi_m, 0, 0)
    it, _, _ = o.relax_threshold(1e-11, 1, 0)
    phi_t = o.phi(o.switches()[0])
    scale = np.max(np.abs(phi_t))
    err = np.max(np.abs(phi_m - phi_t)) / scale
    err_plain = np.max(np.abs(phi_p - phi_t)) / scale
    log("c3: mixed residual on the CPU %.3g V; relative error of the mixed potentials %.3g, of the plain "
        "path's %.3g (oracle to 1e-11 V in %d sweeps)" % (res, err, err_plain, it))
    assert res <= 1e-8
    assert err <= 1e-9


def test_c2_events_one_sequence():
    """c2 (r 15 nm, 40 nm shank): plain, mixed and learned+mixed evaporate the
    same 20 atoms (the learned path records the first 10 and predicts the
    next 10); the oracle's FP64 sweep on the CPU confirms the mixed path's
    residual target"""
    xyz, ids, counts, mesh, species, state0 = relaxed_state(tsg, "c2")
    log("c2: %d nodes %s" % (len(xyz), counts))
    w = dict(mesh=mesh, ids=ids, species=species, state0=state0)
    n = 20
    p = grid_from(w)
    tsg.srand(99)
    pe = p.evaporate(options(0), n)
    m = grid_from(w)
    m.accel_setup(0.9)
    tsg.srand(99)
    me = m.evaporate(options(1), n)
    tsg.srand(99)
    _, le = learned_run(mesh, ids, species, state0, n, record=n // 2)
    log("c2 sequences: plain %s mixed %s learned %s; predictions kept %d" % (
        [e.node for e in pe], [e.node for e in me], [e.node for e in le], sum(e.warm_start == 1 for e in le)))
    assert [e.node for e in pe] == [e.node for e in me] == [e.node for e in le]
    phi_m = m.phi(m.switches()[0])
    tv, tn = mesh.tets()
    o = O.Oracle(O.F64, xyz, m.ids(), tv, tn, O.species_table(bench.SPECIES))
    o.sync_from_couplings(m.couplings())
    o.set_phi(phi_m, phi_m, 0, 0)
    res, _ = o.relax_threaded(1, oracle_threads())
    log("c2: mixed path's final residual on the CPU oracle %.3g V" % res)
    assert res <= 1e-8


def test_c5_full_pipeline_events(capsys, monkeypatch):
    """c5 (r 25 nm, 60 nm shank: 12.75 M atoms, 3 % solute in clusters,
    4.67 M vacuum nodes, 107.6 M tets): plain FP64, mixed and learned+mixed
    evaporate the same 6 atoms, every ion drawn from its species' isotope
    table and flown by the RK4 integrator to the detector plane 10 cm above
    the mesh; the oracle's FP64 sweep on the CPU confirms the mixed path's
    residual target.  The steps report their progress as they go (the
    mesher's every 2 M points, a line per phase): the test takes minutes."""
    monkeypatch.setenv("TSG_VERBOSE", "1")
    with capsys.disabled():
        xyz, ids, counts, mesh, species, state0 = relaxed_state(tsg, "c5")
        log("c5: %d nodes %s, %d tets, relaxed" % (len(xyz), counts, mesh.m))
    zmax = float(np.max(xyz[:, 2]))

    def grid(method):
        g = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
        g.set_phi(*state0)
        for sid, (m, q, w) in bench.ION_TABLES.items():
            g.set_ion_table(sid, m, q, w)
        if method:
            g.accel_setup(0.9)
        return g

    def opts(method):
        e = bench.event_options(tsg, True, zmax)
        if method:
            e.method, e.accel_tol, e.accel_solver = method, 1e-8, 1
        return e

    n = 6
    p = grid(0)
    tsg.srand(99)
    pe = p.evaporate(opts(0), n)
    with capsys.disabled():
        log("c5 plain: %s" % [e.node for e in pe])
    m = grid(1)
    tsg.srand(99)
    me = m.evaporate(opts(1), n)
    lg = grid(2)
    net = tsg.WarmStartNet(max_batch_rows=65536, data_rows=14000 * 16, data_edges=12 * 14000 * 16, seed=1)
    lg.attach(net, record=True, predict=False)
    tsg.srand(99)
    le = lg.evaporate(opts(1), n // 2)
    graphs, _, _ = net.dataset()
    net.train(0, graphs, epochs=20, batch_graphs=16, lr=2e-3, seed=5)
    lg.attach(net, record=False, predict=True)
    le += lg.evaporate(opts(2), n - n // 2)
    log("c5 sequences: plain %s mixed %s learned %s; ions %s" % (
        [e.node for e in pe], [e.node for e in me], [e.node for e in le],
        [(round(e.ion_mass, 3), e.ion_charge) for e in le]))
    assert [e.node for e in pe] == [e.node for e in me] == [e.node for e in le]
    zdet = np.float32(zmax + bench.DETECTOR_DISTANCE)
    for e in pe + me + le:
        assert e.traj_status == 5 and abs(e.stop[2] - zdet) <= 1e-6 * zdet
        masses, charges, weights = bench.ION_TABLES[int(ids[e.node])]
        j = O.ion_pick(bench.ION_SEED, e.event_index, weights)
        assert e.ion_charge == charges[j] and e.ion_mass == np.float32(masses[j])
    phi_m = m.phi(m.switches()[0])
    with capsys.disabled():
        log("c5: sequences equal, flights on the detector plane; the oracle's check follows")
    tv, tn = mesh.tets()
    o = O.Oracle(O.F64, xyz, m.ids(), tv, tn, O.species_table(bench.SPECIES))
    o.sync_from_couplings(m.couplings())
    o.set_phi(phi_m, phi_m, 0, 0)
    res, _ = o.relax_threaded(1, oracle_threads())
    with capsys.disabled():
        log("c5: mixed path's final residual on the CPU oracle %.3g V" % res)
    assert res <= 1e-8
