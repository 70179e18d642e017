"""GPU parity: the HIP path through the C ABI against the reference fixtures
(float-faithful mode) and against the oracle (plain FP64 mode).  Bit-exact
unless stated."""
import ctypes
import sys

import numpy as np
import pytest

import tapsim_modified_amd as tsg
from oracle import oracle as O
from tests.conftest import FIXTURE_SPECIES

pytestmark = pytest.mark.gpu

# the C library's rand(), which the oracle draws from
libc = ctypes.CDLL(None)
libc.rand.restype = ctypes.c_int

F1E2 = float(np.float32(1e-2))
F1E3 = float(np.float32(1e-3))


def log(msg):
    print("[gpu-test] " + msg, flush=True)
    sys.stdout.flush()


def grid(fx, mode):
    mesh = tsg.Mesh(fx["xyz"], fx["tets"], fx["ref_neigh"])
    return tsg.Grid3D(mesh, fx["ids"], tsg.species_table(FIXTURE_SPECIES), 0.0, mode)


def oracle(fx, mode):
    return O.Oracle(mode, fx["xyz"], fx["ids"], fx["tets"], fx["ref_neigh"], O.species_table(FIXTURE_SPECIES))


def same_bits(a, b, dtype):
    a = np.asarray(a, dtype)
    b = np.asarray(b, dtype)
    it = np.uint32 if dtype == np.float32 else np.uint64
    return np.array_equal(a.view(it), b.view(it))


def test_geometry_precompute_f32_matches_reference(tip_small):
    log("geometry f32 tip_small")
    g = grid(tip_small, tsg.FAITHFUL_F32)
    assert same_bits(g.couplings(), tip_small["ref_coupling"], np.float32)
    h32, h64, o32 = g.areas()
    assert same_bits(o32, tip_small["ref_area_one"], np.float32)
    assert np.array_equal(g.flags(), tip_small["ref_flags"])


def test_geometry_precompute_f64_matches_oracle(tip_small):
    log("geometry f64 tip_small")
    g = grid(tip_small, tsg.PLAIN_F64)
    o = oracle(tip_small, O.F64)
    o.sync()
    assert same_bits(g.couplings(), o.couplings(), np.float64)


@pytest.mark.parametrize("mode", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_jacobi_cycles(tip_small, mode):
    log("jacobi mode %d" % mode)
    g = grid(tip_small, mode)
    o = oracle(tip_small, O.F32 if mode == tsg.FAITHFUL_F32 else O.F64)
    o.sync()
    dt = np.float32 if mode == tsg.FAITHFUL_F32 else np.float64
    for k in range(6):
        dg = g.relax(1)
        do = o.relax(1)
        assert dg == do, (k, dg, do)
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    assert g.switches()[0] == o.switches()[0]
    if mode == tsg.FAITHFUL_F32:
        assert same_bits(g.phi(0), tip_small["ref_phi0_r6"], np.float32)
        assert np.array_equal([g.relax(1) for _ in range(0)], [])
    nodes = np.arange(g.n, dtype=np.int32)
    fg = g.field_o1(nodes)
    fo = np.array([o.field_o1(i) for i in range(g.n)])
    assert same_bits(fg, fo, dt)
    if mode == tsg.FAITHFUL_F32:
        assert same_bits(fg, tip_small["ref_field_r6"], np.float32)


@pytest.mark.parametrize("mode", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_threshold_relaxation_with_queue(tip_small, mode):
    log("threshold mode %d" % mode)
    g = grid(tip_small, mode)
    o = oracle(tip_small, O.F32 if mode == tsg.FAITHFUL_F32 else O.F64)
    o.sync()
    dt = np.float32 if mode == tsg.FAITHFUL_F32 else np.float64
    it_g = g.relax_threshold(0.5, 1, 4)
    it_o, _, _ = o.relax_threshold(0.5, 1, 4)
    assert it_g == it_o
    if mode == tsg.FAITHFUL_F32:
        assert it_g == tip_small["ref_thr_iter"][0]
        assert same_bits(g.phi(0), tip_small["ref_phi0_thr"], np.float32)
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    # a tighter threshold from there on
    it_g = g.relax_threshold(1e-4, 1, 0)
    it_o, _, _ = o.relax_threshold(1e-4, 1, 0)
    assert it_g == it_o
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)


@pytest.mark.parametrize("mode", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_resync_and_local_relax(tip_small, mode):
    log("local relax mode %d" % mode)
    node = int(tip_small["local_node"])
    g = grid(tip_small, mode)
    o = oracle(tip_small, O.F32 if mode == tsg.FAITHFUL_F32 else O.F64)
    o.sync()
    dt = np.float32 if mode == tsg.FAITHFUL_F32 else np.float64
    for _ in range(40):
        g.relax(1)
        o.relax(1)
    g.set_id(node, 0)
    o.set_id(node, 0)
    o.resync(node)
    assert same_bits(g.couplings(), o.couplings(), dt)
    it_g = g.local_relax(node, 3, F1E2, 5, 5)
    it_o = o.local_relax_threshold(node, 3, F1E2, 5, 5)
    assert it_g == it_o
    assert g.switches() == o.switches()
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    if mode == tsg.FAITHFUL_F32:
        assert same_bits(g.couplings(), tip_small["ref_coupling_resync"], np.float32)
        assert it_g == tip_small["ref_local_iter"][0]
        assert same_bits(g.phi(0), tip_small["ref_phi0_loc"], np.float32)
        assert same_bits(g.phi(1), tip_small["ref_phi1_loc"], np.float32)
    # the global cycles after the local relaxation read the new row
    for _ in range(3):
        assert g.relax(1) == o.relax(1)
    assert same_bits(g.phi(0), o.phi(0), dt)


def evap_opts(vac=0):
    e = tsg.EvapOptions()
    e.prob_mode, e.evap_mode, e.vacuum_id = 0, 0, vac
    e.relax_threshold, e.relax_shell, e.relax_cycle, e.relax_queue = F1E2, 3, 5, 5
    e.relax_global_cycles = 1
    e.refresh_interval, e.refresh_threshold, e.refresh_cycle, e.refresh_queue = 5, F1E3, 1, 0
    e.exact_probability = 1
    return e


def ora_opts():
    return O.OraEvapOpts(prob_mode=0, evap_mode=0, vacuum_id=0, relax_threshold=1e-2, relax_shell=3, relax_cycle=5,
                         relax_queue=5, relax_global_cycles=1, refresh_interval=5, refresh_threshold=1e-3,
                         refresh_cycle=1, refresh_queue=0, trajectory=0, integrator=1, stepper=1, dt0=1e-14,
                         dt_min=1e-16, err_limit=1e-6, voltage_queue=250, shrink_limit=1.0, inplace_order=0)


@pytest.mark.parametrize("mode", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_evaporation_sequence(tip_small, mode):
    log("evaporation mode %d" % mode)
    g = grid(tip_small, mode)
    o = oracle(tip_small, O.F32 if mode == tsg.FAITHFUL_F32 else O.F64)
    o.sync()
    for _ in range(300):
        g.relax(1)
        o.relax(1)
    evg = g.evaporate(evap_opts(), 25)
    evo = o.evaporate(ora_opts(), 25)
    assert [e.node for e in evg] == [e.node for e in evo]
    assert [e.probability for e in evg] == [e.probability for e in evo]
    fdt = np.float32
    assert same_bits([e.pot_before for e in evg], [e.pot_before for e in evo], fdt)
    assert same_bits([list(e.field_before) for e in evg], [list(e.field_before) for e in evo], fdt)
    assert same_bits([e.pot_after for e in evg], [e.pot_after for e in evo], fdt)
    assert [e.local_iter for e in evg] == [e.local_iter for e in evo]
    assert [e.global_iter for e in evg] == [e.refresh_iter for e in evo]
    assert same_bits([list(e.apex) for e in evg], [list(e.apex) for e in evo], fdt)
    assert same_bits([list(e.normal) for e in evg], [list(e.normal) for e in evo], fdt)
    dt = np.float32 if mode == tsg.FAITHFUL_F32 else np.float64
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    if mode == tsg.FAITHFUL_F32:
        F = tip_small["ref_evap_f"]
        assert np.array_equal([e.node for e in evg], tip_small["ref_evap_node"])
        assert same_bits([e.probability for e in evg], F[:, 0], np.float32)
        assert same_bits([e.pot_before for e in evg], F[:, 1], np.float32)
        assert same_bits([list(e.field_before) for e in evg], F[:, 2:5], np.float32)
        assert same_bits([e.pot_after for e in evg], F[:, 5], np.float32)
        assert same_bits([list(e.normal) for e in evg], tip_small["ref_evap_normal"], np.float32)
        assert same_bits(g.phi(0), tip_small["ref_phi0_evap"], np.float32)


def test_accelerated_relaxation_residual_and_error(tip_small):
    """The mixed FP32/FP64 relaxation reaches max|J(phi)-phi| <= 1e-8 V, checked
    by one FP64 Jacobi sweep of the oracle on the device result, and lands
    within 1e-6 V of the oracle's solution relaxed to 1e-13 V."""
    log("accelerated relaxation")
    g = grid(tip_small, tsg.PLAIN_F64)
    rho = g.accel_setup(0.5)
    assert 0.5 < rho < 1.0
    outer, inner, res = g.accel_relax(1e-8)
    assert res <= 1e-8 and outer >= 1
    phi = g.phi(0)
    assert np.array_equal(phi, g.phi(1))
    o = oracle(tip_small, O.F64)
    o.sync()
    o.set_phi(phi, phi, 0, 0)
    dev = o.relax(1)
    assert dev <= 1e-8, dev
    ex = oracle(tip_small, O.F64)
    ex.sync()
    ex.relax_threshold(1e-13, 1, 0)
    err = np.max(np.abs(phi - ex.phi(ex.switches()[0])))
    log("accelerated: rho %.6f outer %d inner %d residual %.2e error %.2e" % (rho, outer, inner, res, err))
    assert err < 1e-6


def test_accelerated_events_track_plain_sequence(tip_small):
    """Events with the accelerated relaxation: every relaxation ends below the
    residual target and the evaporation sequence follows the plain FP64
    path's (both relaxed to 1e-8 V; near ties may legitimately differ)."""
    log("accelerated events")
    g = grid(tip_small, tsg.PLAIN_F64)
    g.accel_setup(0.5)
    g.accel_relax(1e-10)
    p0, p1 = g.phi(0), g.phi(1)
    ga = grid(tip_small, tsg.PLAIN_F64)
    ga.set_phi(p0, p1, *g.switches())
    op = evap_opts()
    op.refresh_interval, op.refresh_threshold = 1, 1e-8
    oa = evap_opts()
    oa.method, oa.accel_tol = 1, 1e-8
    ep = g.evaporate(op, 15)
    ea = ga.evaporate(oa, 15)
    np_, na = [e.node for e in ep], [e.node for e in ea]
    log("sequences plain %s accel %s" % (np_, na))
    assert np_ == na


def test_trajectory_matches_reference(tip_small):
    """Ion trajectory (Cash-Karp O5, error-restricted stepper) from the apex
    atom after its removal, in the float-faithful grid.  Against the oracle
    computing the step-size root (tol/err)^(1/4) as the device does
    (correctly rounded): every phase bit-exact.  Against the reference, whose
    root is glibc's powf (different in the last bit for about one argument
    in 1500): bit-exact when the oracle's two root modes agree on this
    flight, else within 1e-4 relative at its end."""
    node = int(tip_small["local_node"])
    log("trajectory from node %d" % node)
    g = grid(tip_small, tsg.FAITHFUL_F32)
    o = oracle(tip_small, O.F32)
    o.sync()
    for _ in range(300):
        g.relax(1)
        o.relax(1)
    g.set_id(node, 0)
    g.relax(1)
    o.set_id(node, 0)
    o.resync(node)
    o.relax(1)
    q = 2.0 * np.float32(1.60217733e-19)
    m = np.float32(55.845) * np.float32(1.660538782e-27)
    (st, tr), = g.trajectories([node], q, m, integrator=1, stepper=1, cap=8192)
    try:
        O.set_root_mode(1)
        sto, tro = o.trajectory(node, q, m)
    finally:
        O.set_root_mode(0)
    st0, tr0 = o.trajectory(node, q, m)
    ref = tip_small["ref_traj"]
    log("status %d (ref %d), %d phases (ref %d); oracle root modes agree: %s"
        % (st, tip_small["ref_traj_status"][0], len(tr), len(ref), same_bits(tro, tr0, np.float32)))
    assert st == sto and same_bits(tr, tro, np.float32)
    assert st == tip_small["ref_traj_status"][0]
    if same_bits(tro, tr0, np.float32):
        assert same_bits(tr, ref, np.float32)
    else:
        assert abs(len(tr) - len(ref)) <= max(3, len(ref) // 20)
        span = np.max(np.abs(ref[:, 1:4]))
        assert np.max(np.abs(tr[-1, 1:4] - ref[-1, 1:4])) <= 1e-4 * span


def test_trajectory_batch_matches_single(tip_small):
    """A batch of ions (one device thread each) gives the same trajectories
    as one ion at a time."""
    log("trajectory batch")
    g = grid(tip_small, tsg.PLAIN_F64)
    for _ in range(200):
        g.relax(1)
    g.surface_init(0)
    nodes = g.surface_nodes()[::max(1, len(g.surface_nodes()) // 24)][:24]
    q = 2.0 * np.float32(1.60217733e-19)
    m = np.float32(55.845) * np.float32(1.660538782e-27)
    batch = g.trajectories(nodes, q, m, cap=4096)
    for k in (0, len(nodes) // 2, len(nodes) - 1):
        (st, tr), = g.trajectories([nodes[k]], q, m, cap=4096)
        assert st == batch[k][0]
        assert np.array_equal(tr.view(np.uint32), batch[k][1].view(np.uint32))
    log("statuses %s, lengths %s" % ([b[0] for b in batch], [len(b[1]) for b in batch]))


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("precision", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_probability_models(tip_small, mode, precision):
    """Surface probabilities of the four models of evap_compute_probabilities
    against the oracle: bit-exact, except probBoltzmann's exponential (the
    device exp is not glibc's: within 4 ulp)."""
    log("probability model %d precision %d" % (mode, precision))
    mesh = tsg.Mesh(tip_small["xyz"], tip_small["tets"], tip_small["ref_neigh"])
    g = tsg.Grid3D(mesh, tip_small["ids"], tsg.species_table(FIXTURE_SPECIES), 300.0, precision)
    o = O.Oracle(O.F32 if precision == tsg.FAITHFUL_F32 else O.F64, tip_small["xyz"], tip_small["ids"],
                 tip_small["tets"], tip_small["ref_neigh"], O.species_table(FIXTURE_SPECIES), temperature=300.0)
    o.sync()
    for _ in range(120):
        g.relax(1)
        o.relax(1)
    g.surface_init(0)
    o.surface_init(0)
    assert np.array_equal(g.surface_nodes(), o.surface_nodes())
    ref_g = g.surface_evaluate(True, mode)
    ref_o = o.specific_fields()
    o.compute_probabilities(mode)
    pg, po = g.surface_probabilities(), o.surface_probabilities()
    assert np.float32(ref_g) == np.float32(ref_o)
    if mode == 1:
        np.testing.assert_allclose(pg, po, rtol=5e-7 if precision == tsg.FAITHFUL_F32 else 1e-14, atol=0)
    else:
        assert same_bits(pg, po, np.float32)


def test_monte_carlo_choice(tip_small):
    """evapMonteCarlo draws from the process's rand(): with the same seed the
    device event loop picks the oracle's sequence."""
    log("monte carlo choice")
    g = grid(tip_small, tsg.PLAIN_F64)
    o = oracle(tip_small, O.F64)
    o.sync()
    for _ in range(300):
        g.relax(1)
        o.relax(1)
    eg, eo = evap_opts(), ora_opts()
    eg.evap_mode, eo.evap_mode = 1, 1
    tsg.srand(99)
    evg = g.evaporate(eg, 12)
    o.srand(99)
    evo = o.evaporate(eo, 12)
    assert [e.node for e in evg] == [e.node for e in evo]
    assert same_bits(g.phi(0), o.phi(0), np.float64)


def flight_opts(evap_mode=0, thermal=0):
    e = evap_opts()
    e.evap_mode = evap_mode
    e.trajectory = 1
    e.thermal_velocity = thermal
    return e


def flight_ora_opts(evap_mode=0, thermal=0):
    o = ora_opts()
    o.evap_mode, o.trajectory, o.thermal_velocity = evap_mode, 1, thermal
    return o


def flight_fields(ev):
    return dict(node=[x.node for x in ev], start=[list(x.start) for x in ev], stop=[list(x.stop) for x in ev],
                len=[x.traj_len for x in ev], voltage=[x.voltage for x in ev], ts=[x.time_scale for x in ev],
                tof=[x.tof for x in ev], apex=[list(x.apex) for x in ev], normal=[list(x.normal) for x in ev])


def run_flights(fx, evap_mode, thermal, temperature):
    """Device (float-faithful) and oracle runs of 25 events with flights from
    rand() seeded 99: the oracle once with the device's root (mode 1) and once
    with glibc's (mode 0, the reference's arithmetic)."""
    mesh = tsg.Mesh(fx["xyz"], fx["tets"], fx["ref_neigh"])
    g = tsg.Grid3D(mesh, fx["ids"], tsg.species_table(FIXTURE_SPECIES), temperature, tsg.FAITHFUL_F32)
    for _ in range(300):
        g.relax(1)
    tsg.srand(99)
    ev = g.evaporate(flight_opts(evap_mode, thermal), 25)
    dev_next = int(tsg.rand(1)[0])
    out = []
    for mode in (1, 0):
        o = O.Oracle(O.F32, fx["xyz"], fx["ids"], fx["tets"], fx["ref_neigh"], O.species_table(FIXTURE_SPECIES),
                     temperature=temperature)
        o.sync()
        for _ in range(300):
            o.relax(1)
        o.srand(99)
        try:
            O.set_root_mode(mode)
            evo = o.evaporate(flight_ora_opts(evap_mode, thermal), 25)
        finally:
            O.set_root_mode(0)
        out.append((evo, libc.rand()))
    return ev, dev_next, out


def check_flights(ev, dev_next, out, nodes, F, next_rand, label):
    (ev1, next1), (ev0, next0) = out
    # the device against the oracle with the device's root: everything, and
    # the position of the rand() stream afterwards
    d, o1 = flight_fields(ev), flight_fields(ev1)
    assert d["node"] == o1["node"] and d["len"] == o1["len"]
    for k in ("start", "stop", "voltage", "ts", "tof", "apex", "normal"):
        assert same_bits(d[k], o1[k], np.float32), k
    assert dev_next == next1
    # against the reference: bit-exact up to the first flight in which
    # glibc's powf and the correctly rounded root differ
    diff = [x.traj_root_diff for x in ev0]
    first = next((i for i, c in enumerate(diff) if c), len(ev))
    log("%s: root differences per event %s; device next rand %d, reference %d" % (label, diff, dev_next, next_rand))
    assert np.array_equal(d["node"][:first + 1], nodes[:first + 1])
    assert same_bits(d["start"][:first + 1], F[:first + 1, 11:14], np.float32)
    assert same_bits(d["apex"][:first + 1], F[:first + 1, 18:21], np.float32)
    assert same_bits(d["stop"][:first], F[:first, 14:17], np.float32)
    assert same_bits(d["voltage"][:first], F[:first, 9], np.float32)
    assert same_bits(d["ts"][:first], F[:first, 10], np.float32)
    assert same_bits(d["tof"][:first], F[:first, 17], np.float32)
    if first == len(ev):
        assert dev_next == next_rand
    return first


def test_evaporation_with_trajectories_matches_reference(tip_small):
    """Steps I-VI with the ion flights on (float-faithful grid), the walks
    continuing the process's rand() stream.  Device = oracle with the
    device's fourth root, bit for bit (sequence, flights, voltage, time scale,
    time of flight, apex, normal, rand() position); device = reference up to
    the first flight where glibc's powf root differs, and the sequence, start
    points, apex and normals throughout."""
    log("evaporation with trajectories")
    ev, dev_next, out = run_flights(tip_small, 0, 0, 0.0)
    F = tip_small["ref_evap_f"]
    check_flights(ev, dev_next, out, tip_small["ref_evap_node"], F, int(tip_small["ref_evap_next_rand"][0]), "maximum")
    assert np.array_equal([x.node for x in ev], tip_small["ref_evap_node"])
    assert same_bits([list(x.start) for x in ev], F[:, 11:14], np.float32)
    assert same_bits([list(x.apex) for x in ev], F[:, 18:21], np.float32)
    assert same_bits([list(x.normal) for x in ev], tip_small["ref_evap_normal"], np.float32)
    np.testing.assert_allclose([x.voltage for x in ev], F[:, 9], rtol=2e-3)
    np.testing.assert_allclose([x.time_scale for x in ev], F[:, 10], rtol=2e-3)


def test_monte_carlo_thermal_flights_match_reference(tip_small):
    """evapMonteCarlo with thermal start velocities (60 K): the choice, the
    velocity draws and every walk step take rand() in the reference's order
    on the device.  Device = oracle with the device's root bit for bit; device
    = reference up to the first flight where glibc's powf root differs."""
    log("monte carlo + thermal velocities with trajectories")
    ev, dev_next, out = run_flights(tip_small, 1, 1, 60.0)
    first = check_flights(ev, dev_next, out, tip_small["ref_mc_node"], tip_small["ref_mc_f"],
                          int(tip_small["ref_mc_next_rand"][0]), "monte carlo")
    log("monte carlo: reference sequence followed for %d events" % sum(
        1 for a, b in zip([x.node for x in ev], tip_small["ref_mc_node"]) if a == b))


@pytest.mark.parametrize("mode", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_initial_potentials_match_oracle(tip_small, mode):
    """randomReset (same rand() seed) and preset_sphericalPotential: device
    grid and oracle bit for bit."""
    log("initial potentials mode %d" % mode)
    g = grid(tip_small, mode)
    o = oracle(tip_small, O.F32 if mode == tsg.FAITHFUL_F32 else O.F64)
    o.sync()
    dt = np.float32 if mode == tsg.FAITHFUL_F32 else np.float64
    tsg.srand(11)
    g.reset_potentials(1)
    o.srand(11)
    o.random_reset()
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    c = tip_small["xyz"].mean(0).astype(np.float32)
    g.reset_potentials(2, c, 50.0, 4500.0)
    o.spherical_preset(c, 50.0, 4500.0)
    assert same_bits(g.phi(0), o.phi(0), dt)
    g.reset_potentials(0)
    assert same_bits(g.phi(0), np.array([FIXTURE_SPECIES[i].get("phi", 0.0) for i in tip_small["ids"]]), dt)


def test_fieldlines_match_reference(tip_small):
    """Field lines (Fieldline_3d) from six points above the tip, one rand()
    stream: to the mesh's end, then limited to 2 nm.  Device = oracle with
    the device's cube root (correctly rounded) bit for bit, the stream
    position included; device = reference wherever the oracle's two root
    modes agree; the parallel mode (one thread per line, no stream) gives
    the same lines."""
    log("field lines")
    g = grid(tip_small, tsg.FAITHFUL_F32)
    o = oracle(tip_small, O.F32)
    o.sync()
    for _ in range(300):
        g.relax(1)
        o.relax(1)
    starts = tip_small["fieldline_start"]
    tsg.srand(99)
    dev = {t: g.fieldlines(starts, np.float32(d), 5000, 1000) for t, d in (("a", -1.0), ("b", 2e-9))}
    dev_next = int(tsg.rand(1)[0])
    ora = {}
    for mode in (1, 0):
        o.srand(99)
        try:
            O.set_root_mode(mode)
            ora[mode] = {t: [o.fieldline(s, np.float32(d), 5000, 1000) for s in starts]
                         for t, d in (("a", -1.0), ("b", 2e-9))}
        finally:
            O.set_root_mode(0)
        if mode == 1:
            ora_next = libc.rand()
    assert dev_next == ora_next
    for t in "ab":
        info = tip_small["ref_fieldline_%s_info" % t]
        ref = tip_small["ref_fieldline_%s" % t]
        k = 0
        for i, ((st, pts), (sto, pto), (st0, pt0)) in enumerate(zip(dev[t], ora[1][t], ora[0][t])):
            assert st == sto and same_bits(pts, pto, np.float32), (t, i)
            if st0 == sto and same_bits(pt0, pto, np.float32):
                assert st == info[i, 0] and same_bits(pts, ref[k:k + info[i, 1]], np.float32), (t, i)
            k += info[i, 1]
        par = g.fieldlines(starts, np.float32(-1.0 if t == "a" else 2e-9), 5000, 1000, process_rand=False)
        for (st, pts), (stp, ptp) in zip(dev[t], par):
            assert st == stp and same_bits(pts, ptp, np.float32)
    log("field line statuses %s / %s" % ([s for s, _ in dev["a"]], [s for s, _ in dev["b"]]))


@pytest.mark.parametrize("mode", [tsg.FAITHFUL_F32, tsg.PLAIN_F64])
def test_space_charge(tip_small, mode):
    """computeCharge with a charged vacuum (2e6 C/m^3): the device's cell
    volumes give every charge after fastSync and after the fastResync of an
    evaporated atom, and six Jacobi cycles with them, bit-exact against the
    oracle (both precisions) and the reference (float)."""
    log("space charge mode %d" % mode)
    sp = {k: dict(v) for k, v in FIXTURE_SPECIES.items()}
    sp[0]["charge_density"] = 2.0e6
    mesh = tsg.Mesh(tip_small["xyz"], tip_small["tets"], tip_small["ref_neigh"])
    g = tsg.Grid3D(mesh, tip_small["ids"], tsg.species_table(sp), 0.0, mode)
    om = O.F32 if mode == tsg.FAITHFUL_F32 else O.F64
    o = O.Oracle(om, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"], O.species_table(sp))
    o.sync()
    dt = np.float32 if mode == tsg.FAITHFUL_F32 else np.float64
    assert same_bits(g.charges(), o.charges(), dt)
    for _ in range(6):
        g.relax(1)
        o.relax(1)
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    node = int(tip_small["local_node"])
    g.set_id(node, 0)
    o.set_id(node, 0)
    o.resync(node)
    assert same_bits(g.charges(), o.charges(), dt) and g.charges()[node] != 0.0
    if mode == tsg.FAITHFUL_F32:
        g2 = tsg.Grid3D(mesh, tip_small["ids"], tsg.species_table(sp), 0.0, mode)
        assert same_bits(g2.charges(), tip_small["ref_charge_q"], np.float32)
        for _ in range(6):
            g2.relax(1)
        assert same_bits(g2.phi(0), tip_small["ref_phi0_q6"], np.float32)


def test_update_coordination(tip_small):
    """The fork's update_coordination on the device grid's ids, once and
    twelve times (signed-char wrap): the reference's ids.  Events with it (ids
    10..127 defined): sequence, ids, probabilities and flights as the oracle
    with the device's root and the reference's sequence and ids; with only
    the standard species the first event fails at the configuration lookup."""
    from tests.conftest import coordination_species
    log("update_coordination")
    g = grid(tip_small, tsg.FAITHFUL_F32)
    g.update_coordination()
    assert np.array_equal(g.ids(), tip_small["ref_coord_ids1"])
    for _ in range(11):
        g.update_coordination()
    assert np.array_equal(g.ids(), tip_small["ref_coord_ids12"])
    sp = coordination_species()
    mesh = tsg.Mesh(tip_small["xyz"], tip_small["tets"], tip_small["ref_neigh"])
    g = tsg.Grid3D(mesh, tip_small["ids"], tsg.species_table(sp), 0.0, tsg.FAITHFUL_F32)
    for _ in range(300):
        g.relax(1)
    tsg.srand(99)
    e = flight_opts()
    e.update_coordination = 1
    ev = g.evaporate(e, 6)
    o = O.Oracle(O.F32, tip_small["xyz"], tip_small["ids"], tip_small["tets"], tip_small["ref_neigh"],
                 O.species_table(sp))
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.srand(99)
    oo = flight_ora_opts()
    oo.update_coordination = 1
    try:
        O.set_root_mode(1)
        evo = o.evaporate(oo, 6)
    finally:
        O.set_root_mode(0)
    d, o1 = flight_fields(ev), flight_fields(evo)
    assert d["node"] == o1["node"] and [x.id for x in ev] == [x.id for x in evo]
    for k in ("start", "stop", "voltage", "ts", "tof"):
        assert same_bits(d[k], o1[k], np.float32), k
    assert np.array_equal(g.ids(), o.ids())
    assert np.array_equal(d["node"], tip_small["ref_coord_evap_node"])
    assert np.array_equal([x.id for x in ev], tip_small["ref_coord_evap_id"])
    assert same_bits([x.probability for x in ev], tip_small["ref_coord_evap_f"][:, 0], np.float32)
    assert np.array_equal(g.ids(), tip_small["ref_coord_evap_ids"])
    g2 = grid(tip_small, tsg.FAITHFUL_F32)
    with pytest.raises(RuntimeError):
        g2.evaporate(e, 1)
