"""The oracle's float restatement against the reference's own outputs
(tests/golden/tip_small.npz, written by the reference built from its sources).
Bit-exact unless stated."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import FIXTURE_SPECIES

# the C library's rand(), whose state the oracle's draws advance
libc = ctypes.CDLL(None)
libc.rand.restype = ctypes.c_int

F1E2 = float(np.float32(1e-2))
F1E3 = float(np.float32(1e-3))


def make(fx, mode=O.F32):
    return O.Oracle(mode, fx["xyz"], fx["ids"], fx["tets"], fx["ref_neigh"], O.species_table(FIXTURE_SPECIES))


def bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_topology(tip_small):
    o = make(tip_small)
    ptr, idx = o.csr()
    assert np.array_equal(o.assoc(), tip_small["ref_assoc"])
    assert np.array_equal(o.boundary(), tip_small["ref_boundary"])
    assert np.array_equal(ptr, tip_small["ref_ptr"]) and np.array_equal(idx, tip_small["ref_idx"])


def test_fast_sync_couplings_and_areas(tip_small):
    o = make(tip_small)
    o.sync()
    assert bits_equal(o.couplings(), tip_small["ref_coupling"])
    assert np.array_equal(o.flags(), tip_small["ref_flags"])
    ptr, idx = o.csr()
    a1 = [o.voronoi_area(i, idx[k], 1.0) for i in range(o.n) for k in range(ptr[i], ptr[i + 1])]
    assert bits_equal(a1, tip_small["ref_area_one"])


def test_jacobi_cycles_and_fields(tip_small):
    o = make(tip_small)
    o.sync()
    devs = [o.relax(1) for _ in range(6)]
    assert np.array_equal(np.array(devs), tip_small["ref_dev_r6"])
    assert bits_equal(o.phi(0), tip_small["ref_phi0_r6"]) and bits_equal(o.phi(1), tip_small["ref_phi1_r6"])
    f1 = [o.field_o1(i) for i in range(o.n)]
    assert bits_equal(f1, tip_small["ref_field_r6"])
    f2 = [o.field_o2(i) for i in range(o.n)]
    assert bits_equal([f[0] for f in f2], tip_small["ref_field2_r6"])
    assert bits_equal([f[1] for f in f2], tip_small["ref_charge2_r6"])


def test_threshold_queue_and_inplace_cycles(tip_small):
    o = make(tip_small)
    o.sync()
    it, _, _ = o.relax_threshold(0.5, 1, 4)
    assert it == tip_small["ref_thr_iter"][0]
    assert bits_equal(o.phi(0), tip_small["ref_phi0_thr"]) and bits_equal(o.phi(1), tip_small["ref_phi1_thr"])
    devs = [o.relax(4) for _ in range(3)]
    assert np.array_equal(np.array(devs), tip_small["ref_dev_c4"][-3:])
    assert bits_equal(o.phi(0), tip_small["ref_phi0_c4"]) and bits_equal(o.phi(1), tip_small["ref_phi1_c4"])


def test_resync_and_local_relax(tip_small):
    node = int(tip_small["local_node"])
    o = make(tip_small)
    o.sync()
    for _ in range(40):
        o.relax(1)
    o.set_id(node, 0)
    o.resync(node)
    assert bits_equal(o.couplings(), tip_small["ref_coupling_resync"])
    it = o.local_relax_threshold(node, 3, F1E2, 5, 5)
    assert it == tip_small["ref_local_iter"][0]
    assert bits_equal(o.phi(0), tip_small["ref_phi0_loc"]) and bits_equal(o.phi(1), tip_small["ref_phi1_loc"])


def evap_opts():
    return O.OraEvapOpts(prob_mode=0, evap_mode=0, vacuum_id=0, relax_threshold=1e-2, relax_shell=3, relax_cycle=5,
                         relax_queue=5, relax_global_cycles=1, refresh_interval=5, refresh_threshold=1e-3,
                         refresh_cycle=1, refresh_queue=0, trajectory=1, integrator=1, stepper=1, dt0=1e-14,
                         dt_min=1e-16, err_limit=1e-6, voltage_queue=250, shrink_limit=1.0, inplace_order=0)


def test_evaporation_sequence(tip_small):
    """25 events of Process::evaporation (steps I-VI): sequence, selection
    data, trajectory end points, voltage and time scale bit-exact, and the
    rand() stream left where the reference leaves it."""
    o = make(tip_small)
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.srand(99)
    ev = o.evaporate(evap_opts(), 25)
    F = tip_small["ref_evap_f"]
    assert np.array_equal([e.node for e in ev], tip_small["ref_evap_node"])
    assert bits_equal([e.probability for e in ev], F[:, 0])
    assert bits_equal([e.pot_before for e in ev], F[:, 1])
    assert bits_equal([list(e.field_before) for e in ev], F[:, 2:5])
    assert bits_equal([e.pot_after for e in ev], F[:, 5])
    assert bits_equal([list(e.field_after) for e in ev], F[:, 6:9])
    assert bits_equal([list(e.start) for e in ev], F[:, 11:14])
    assert bits_equal([list(e.stop) for e in ev], F[:, 14:17])
    assert np.array_equal([e.traj_len for e in ev], tip_small["ref_evap_trajlen"])
    assert np.array_equal([[e.local_iter, e.refresh_iter] for e in ev], tip_small["ref_evap_iter"])
    assert bits_equal([list(e.apex) for e in ev], F[:, 18:21])
    assert bits_equal([list(e.normal) for e in ev], tip_small["ref_evap_normal"])
    assert bits_equal([e.voltage for e in ev], F[:, 9])
    assert bits_equal([e.time_scale for e in ev], F[:, 10])
    assert bits_equal([e.tof for e in ev], F[:, 17])
    assert bits_equal(o.phi(0), tip_small["ref_phi0_evap"]) and bits_equal(o.phi(1), tip_small["ref_phi1_evap"])
    # the walks' first faces drew from rand() exactly as often as the reference's
    assert libc.rand() == tip_small["ref_evap_next_rand"][0]


def test_monte_carlo_thermal_evaporation(tip_small):
    """evapMonteCarlo with thermal start velocities at 60 K: the choice, the
    velocity and every walk step draw from rand() in the reference's order, so
    the sequence, the flights and the stream position after 25 events are the
    reference's."""
    o = O.Oracle(O.F32, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(FIXTURE_SPECIES), temperature=60.0)
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.srand(99)
    opts = evap_opts()
    opts.evap_mode = 1
    opts.thermal_velocity = 1
    ev = o.evaporate(opts, 25)
    F = tip_small["ref_mc_f"]
    assert np.array_equal([e.node for e in ev], tip_small["ref_mc_node"])
    assert bits_equal([e.probability for e in ev], F[:, 0])
    assert bits_equal([list(e.start) for e in ev], F[:, 11:14])
    assert bits_equal([list(e.stop) for e in ev], F[:, 14:17])
    assert np.array_equal([e.traj_len for e in ev], tip_small["ref_mc_trajlen"])
    assert bits_equal([e.voltage for e in ev], F[:, 9])
    assert bits_equal([e.time_scale for e in ev], F[:, 10])
    assert libc.rand() == tip_small["ref_mc_next_rand"][0]


def test_trajectory(tip_small):
    node = int(tip_small["local_node"])
    o = make(tip_small)
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.set_id(node, 0)
    o.resync(node)
    o.relax(1)
    o.srand(99)
    st, tr = o.trajectory(node, 2.0 * np.float32(1.60217733e-19), np.float32(55.845) * np.float32(1.660538782e-27))
    ref = tip_small["ref_traj"]
    assert st == tip_small["ref_traj_status"][0]
    assert tr.shape == ref.shape
    # every phase bit-exact, the extrapolated one included (its walk starts
    # from the last integrated point's tet, -1 outside the mesh)
    assert bits_equal(tr, ref)


def test_threaded_relax_equals_serial(tip_small):
    """The CPU baseline's threaded sweep (the reference's chunked scheme)
    gives the serial relax(1)'s potentials and deviation bit for bit."""
    a = O.Oracle(O.F64, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(FIXTURE_SPECIES))
    b = O.Oracle(O.F64, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(FIXTURE_SPECIES))
    a.sync()
    b.sync()
    for _ in range(7):
        d = a.relax(1)
    dt, sec = b.relax_threaded(7, 4)
    assert d == dt and sec > 0
    assert a.switches() == b.switches()
    assert np.array_equal(a.phi(0), b.phi(0)) and np.array_equal(a.phi(1), b.phi(1))


@pytest.mark.parametrize("mode", [O.F32, O.F64])
def test_threaded_threshold_relaxation_equals_serial(tip_small, mode):
    """Threshold relaxations with cycleSize 1 on worker threads (the event
    loop's refresh relaxation of the CPU baseline): the serial iterations,
    potentials and fault, and in float the reference's own"""
    def make_o(threads):
        o = O.Oracle(mode, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                     O.species_table(FIXTURE_SPECIES))
        o.sync()
        o.set_threads(threads)
        return o
    a, b = make_o(1), make_o(5)
    assert a.relax_threshold(0.5, 1, 4) == b.relax_threshold(0.5, 1, 4)
    if mode == O.F32:
        assert b.relax_threshold(0.0, 1, 0)[0] > 0
        a.relax_threshold(0.0, 1, 0)
    assert np.array_equal(a.phi(0), b.phi(0)) and np.array_equal(a.phi(1), b.phi(1))
    if mode == O.F32:
        assert bits_equal(b.phi(0), a.phi(0))
    # a stagnating queue: the same RelaxationFault on both
    assert a.relax_threshold(-1.0, 1, 25) == b.relax_threshold(-1.0, 1, 25)
    assert a.switches() == b.switches()
    # the float threaded sweep of the CPU calibration
    c = make_o(1)
    d1, _ = c.relax_threaded(9, 3)
    e = make_o(1)
    for _ in range(9):
        d2 = e.relax(1)
    assert d1 == d2 and np.array_equal(c.phi(0), e.phi(0)) and np.array_equal(c.phi(1), e.phi(1))


def test_threaded_event_loop_equals_serial(tip_small):
    """The event loop with its refresh relaxation on worker threads gives the
    serial loop's events (the CPU baseline times the threaded one)"""
    def run(threads):
        o = make(tip_small)
        o.sync()
        for _ in range(300):
            o.relax(1)
        o.set_threads(threads)
        o.srand(99)
        return o.evaporate(evap_opts(), 8), o.phi(0)
    e1, p1 = run(1)
    e4, p4 = run(4)
    assert [e.node for e in e1] == [e.node for e in e4]
    assert [e.refresh_iter for e in e1] == [e.refresh_iter for e in e4]
    assert bits_equal(p1, p4)
    assert np.array_equal([e.node for e in e1], tip_small["ref_evap_node"][:8])


def test_fieldlines(tip_small):
    """Fieldline_3d (init + compute) from six points above the tip, one after
    the other on one rand() stream: to the mesh's end (a), then limited to
    2 nm (b).  Every point, field and tet bit-exact, the statuses, and the
    stream position after each set."""
    o = make(tip_small)
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.srand(99)
    for tag, distance in (("a", -1.0), ("b", 2e-9)):
        info = tip_small["ref_fieldline_%s_info" % tag]
        ref = tip_small["ref_fieldline_%s" % tag]
        k = 0
        for i, start in enumerate(tip_small["fieldline_start"]):
            st, pts = o.fieldline(start, np.float32(distance), 5000, 1000)
            assert st == info[i, 0] and len(pts) == info[i, 1]
            assert bits_equal(pts, ref[k:k + len(pts)])
            k += len(pts)
        # the reference driver draws once to record the position
        assert libc.rand() == tip_small["ref_fieldline_%s_next_rand" % tag][0]


def space_charge_species():
    sp = {k: dict(v) for k, v in FIXTURE_SPECIES.items()}
    sp[0]["charge_density"] = 2.0e6   # C/m^3, the vacuum of the space-charge fixture
    return sp


def test_space_charge(tip_small):
    """computeCharge with a charged vacuum: every node's charge after fastSync
    and after fastResync of an evaporated atom, and six Jacobi cycles with
    the charges, bit-exact."""
    o = O.Oracle(O.F32, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(space_charge_species()))
    o.sync()
    assert bits_equal(o.charges(), tip_small["ref_charge_q"])
    for _ in range(6):
        o.relax(1)
    assert bits_equal(o.phi(0), tip_small["ref_phi0_q6"]) and bits_equal(o.phi(1), tip_small["ref_phi1_q6"])
    o2 = O.Oracle(O.F32, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                  O.species_table(space_charge_species()))
    o2.sync()
    for _ in range(40):
        o2.relax(1)
    node = int(tip_small["local_node"])
    o2.set_id(node, 0)
    o2.resync(node)
    assert bits_equal(o2.charges(), tip_small["ref_charge_q_resync"])
    assert o2.charges()[node] != 0.0


def test_update_coordination(tip_small):
    """The fork's update_coordination once and twelve times (ids wrap as a
    signed char): every id as the reference's."""
    o = make(tip_small)
    o.sync()
    o.update_coordination()
    assert np.array_equal(o.ids(), tip_small["ref_coord_ids1"])
    for _ in range(11):
        o.update_coordination()
    assert np.array_equal(o.ids(), tip_small["ref_coord_ids12"])


def test_evaporation_with_update_coordination(tip_small):
    """Process::evaporation with the fork's update_coordination after the
    initial surface evaluation and after every take-away, ids 10..127
    defined: six events' sequence, ids, selection data and flights, and every
    node's id afterwards, bit-exact; with only the standard species the first
    event fails at the configuration lookup, as the reference's does."""
    from tests.conftest import coordination_species
    o = O.Oracle(O.F32, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(coordination_species()))
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.srand(99)
    opts = evap_opts()
    opts.update_coordination = 1
    ev = o.evaporate(opts, 6)
    F = tip_small["ref_coord_evap_f"]
    assert np.array_equal([e.node for e in ev], tip_small["ref_coord_evap_node"])
    assert np.array_equal([e.id for e in ev], tip_small["ref_coord_evap_id"])
    assert bits_equal([e.probability for e in ev], F[:, 0])
    assert bits_equal([list(e.stop) for e in ev], F[:, 14:17])
    assert np.array_equal(o.ids(), tip_small["ref_coord_evap_ids"])
    o2 = make(tip_small)
    o2.sync()
    with pytest.raises(O.OracleError):
        o2.evaporate(opts, 1)
