"""GPU parity for the parts pinned by tests/golden/pinning.npz (the
reference's own outputs) and the oracle: relax(cycleSize > 1) with its
in-place cycles, RelaxationFault through the C ABI, randomReset on the
process's rand() stream, the probability models 1-3 in the event loop,
field_o2, the Euler integrator and the none / geometry-restricted / adaptive
steppers, and the event counter across chunked calls.  Bit-exact unless
stated."""
import os
import sys

import numpy as np
import pytest

import tapsim_modified_amd as tsg
from oracle import oracle as O
from tests.conftest import FIXTURE_SPECIES, ROOT

pytestmark = pytest.mark.gpu

# temperature of the Boltzmann fixture (tests/golden/make_golden.py)
BOLTZMANN_T = 20000.0
Q = 2.0 * np.float32(1.60217733e-19)
M = np.float32(55.845) * np.float32(1.660538782e-27)


def log(msg):
    print("[gpu-pin] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pin():
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "pinning.npz")))


def grid(fx, mode, temperature=0.0):
    mesh = tsg.Mesh(fx["xyz"], fx["tets"], fx["ref_neigh"])
    return tsg.Grid3D(mesh, fx["ids"], tsg.species_table(FIXTURE_SPECIES), temperature, mode)


def oracle(fx, mode, temperature=0.0):
    o = O.Oracle(mode, fx["xyz"], fx["ids"], fx["tets"], fx["ref_neigh"], O.species_table(FIXTURE_SPECIES),
                 temperature=temperature)
    o.sync()
    return o


def same_bits(a, b, dtype):
    a = np.asarray(a, dtype)
    b = np.asarray(b, dtype)
    it = np.uint32 if dtype == np.float32 else np.uint64
    return a.shape == b.shape and np.array_equal(a.view(it), b.view(it))


def precisions():
    return [(tsg.FAITHFUL_F32, O.F32, np.float32), (tsg.PLAIN_F64, O.F64, np.float64)]


@pytest.mark.parametrize("mode,omode,dt", precisions())
def test_relax_cycle_size_above_one(tip_small, pin, mode, omode, dt):
    """relax(0.5, 4, 5) and relax(100): odd cycles Jacobi, even cycles in
    place in ascending node order, the switch held for each call"""
    log("relax(C>1) mode %d" % mode)
    g = grid(tip_small, mode)
    o = oracle(tip_small, omode)
    it_g = g.relax_threshold(0.5, 4, 5)
    it_o, _, _ = o.relax_threshold(0.5, 4, 5)
    assert it_g == it_o
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    assert g.switches()[0] == o.switches()[0]
    dg, do = g.relax(100), o.relax(100)
    assert dg == do
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    if mode == tsg.FAITHFUL_F32:
        assert it_g == pin["ref_thr4_iter"][0]
        assert dg == pin["ref_dev_c100"][0]
        assert same_bits(g.phi(0), pin["ref_phi0_c100"], dt) and same_bits(g.phi(1), pin["ref_phi1_c100"], dt)
    log("relax(0.5, 4, 5): %d iterations; relax(100) deviation %.6g" % (it_g, dg))


def test_relax_four_cycles_against_reference(tip_small):
    """relax(4) three times after relax(0.5, 1, 4), float, against the
    reference's deviations and both buffers"""
    g = grid(tip_small, tsg.FAITHFUL_F32)
    g.relax_threshold(0.5, 1, 4)
    devs = [g.relax(4) for _ in range(3)]
    assert np.array_equal(np.array(devs), tip_small["ref_dev_c4"][-3:])
    assert same_bits(g.phi(0), tip_small["ref_phi0_c4"], np.float32)
    assert same_bits(g.phi(1), tip_small["ref_phi1_c4"], np.float32)


def test_relaxation_fault_through_the_abi(tip_small, pin):
    """a negative threshold from a float fixed point: tsg_relax_threshold
    reports the RelaxationFault (iteration, deviation, slope) of the reference,
    for cycleSize 1 and 4"""
    g = grid(tip_small, tsg.FAITHFUL_F32)
    for _ in range(2000):
        g.relax(1)
    with pytest.raises(tsg.RelaxationFault) as e:
        g.relax_threshold(-1.0, 1, 25)
    assert -1 - e.value.iteration == pin["ref_fault_iter"][0]
    assert (e.value.deviation, e.value.slope) == tuple(pin["ref_fault"])
    with pytest.raises(tsg.RelaxationFault) as e:
        g.relax_threshold(-2.0, 4, 5)
    assert -1 - e.value.iteration == pin["ref_fault_iter"][1]
    assert (e.value.deviation, e.value.slope) == tuple(pin["ref_fault4"])
    assert same_bits(g.phi(0), pin["ref_phi0_fault"], np.float32)
    assert same_bits(g.phi(1), pin["ref_phi1_fault"], np.float32)
    # without raising: the iteration count comes back as -1 - iteration
    assert g.relax_threshold(-1.0, 1, 25, raise_fault=False) == pin["ref_fault_iter"][0]


@pytest.mark.parametrize("mode,omode,dt", precisions())
def test_random_reset(tip_small, pin, mode, omode, dt):
    g = grid(tip_small, mode)
    tsg.srand(7)
    g.reset_potentials(1)
    o = oracle(tip_small, omode)
    o.srand(7)
    o.random_reset()
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
    if mode == tsg.FAITHFUL_F32:
        assert same_bits(g.phi(0), pin["ref_phi0_rr"], dt) and same_bits(g.phi(1), pin["ref_phi1_rr"], dt)


def evap_opts(prob):
    e = tsg.EvapOptions()
    e.prob_mode, e.evap_mode, e.vacuum_id = prob, 0, 0
    e.relax_threshold, e.relax_shell, e.relax_cycle, e.relax_queue = float(np.float32(1e-2)), 3, 5, 5
    e.relax_global_cycles, e.refresh_interval = 1, 5
    e.refresh_threshold, e.refresh_cycle, e.refresh_queue = float(np.float32(1e-3)), 1, 0
    e.exact_probability = 1
    e.trajectory, e.integrator, e.stepper = 1, 1, 1
    return e


@pytest.mark.parametrize("prob", [1, 2, 3])
def test_probability_models_in_event_loop(tip_small, pin, prob):
    """10 events with probBoltzmann, probLinear_force and
    probVoronoiFluxForce in the float grid against the reference: sequence,
    relaxation counts, fields bit-exact; probabilities bit-exact for models 2
    and 3, within 8 ulp for the Boltzmann model (the device exponential is
    not glibc's)"""
    log("probability model %d in the event loop" % prob)
    T = BOLTZMANN_T if prob == 1 else 0.0
    g = grid(tip_small, tsg.FAITHFUL_F32, T)
    for _ in range(300):
        g.relax(1)
    tsg.srand(99)
    ev = g.evaporate(evap_opts(prob), 10)
    F = pin["ref_p%d_f" % prob]
    assert [e.node for e in ev] == list(pin["ref_p%d_node" % prob])
    assert np.array_equal([[e.local_iter, e.global_iter] for e in ev], pin["ref_p%d_iter" % prob])
    assert same_bits([e.pot_before for e in ev], F[:, 1], np.float32)
    assert same_bits([list(e.field_before) for e in ev], F[:, 2:5], np.float32)
    p = np.array([e.probability for e in ev], np.float32)
    if prob == 1:
        ulp = np.abs(p.view(np.int32).astype(np.int64) - F[:, 0].view(np.int32).astype(np.int64))
        assert ulp.max() <= 8, ulp
    else:
        assert same_bits(p, F[:, 0], np.float32)


@pytest.mark.parametrize("mode,omode,dt", precisions())
def test_field_o2(tip_small, mode, omode, dt):
    """tsg_field_o2 of every node against the oracle (and the reference's
    field and charge in float)"""
    g = grid(tip_small, mode)
    o = oracle(tip_small, omode)
    for _ in range(6):
        g.relax(1)
        o.relax(1)
    nodes = np.arange(g.n, dtype=np.int32)
    fg, qg = g.field_o2(nodes)
    fo = np.array([o.field_o2(i)[0] for i in range(g.n)])
    qo = np.array([o.field_o2(i)[1] for i in range(g.n)])
    assert same_bits(fg, fo, dt) and same_bits(qg, qo, dt)
    if mode == tsg.FAITHFUL_F32:
        assert same_bits(fg, tip_small["ref_field2_r6"], np.float32)
        assert same_bits(qg, tip_small["ref_charge2_r6"], np.float32)


@pytest.mark.parametrize("tag,integrator,stepper", [("o1", 0, 1), ("s0", 1, 0), ("s2", 1, 2), ("s3", 1, 3),
                                                    ("o5", 1, 1)])
def test_trajectory_integrators_and_steppers(tip_small, pin, tag, integrator, stepper):
    """one flight per integrator / stepper from the apex atom after its
    removal: every phase and the final error estimate equal the oracle with
    the device's correctly rounded step-control root, and the reference where
    the oracle's two root modes agree (the none and geometry-restricted
    steppers take no root)"""
    node = int(tip_small["local_node"])
    log("trajectory %s from node %d" % (tag, node))
    g = grid(tip_small, tsg.FAITHFUL_F32)
    o = oracle(tip_small, O.F32)
    for _ in range(300):
        g.relax(1)
        o.relax(1)
    g.set_id(node, 0)
    g.relax(1)
    o.set_id(node, 0)
    o.resync(node)
    o.relax(1)
    (st, tr, est), = g.trajectories([node], Q, M, integrator=integrator, stepper=stepper, cap=8192, with_error=True)
    try:
        O.set_root_mode(1)
        sto, tro, esto = o.trajectory(node, Q, M, integrator=integrator, stepper=stepper, with_error=True)
    finally:
        O.set_root_mode(0)
    st0, tr0, est0 = o.trajectory(node, Q, M, integrator=integrator, stepper=stepper, with_error=True)
    assert st == sto and same_bits(tr, tro, np.float32)
    assert same_bits(est, esto, np.float32)
    ref = pin["ref_traj_" + tag]
    assert st == pin["ref_traj_status_" + tag][0]
    if same_bits(tro, tr0, np.float32) and same_bits(esto, est0, np.float32):
        assert same_bits(tr, ref, np.float32)
        assert same_bits(est, pin["ref_traj_err_" + tag], np.float32)
    log("%s: status %d, %d phases" % (tag, st, len(tr)))


def test_event_counter_across_chunked_calls(tip_small):
    """events evaporated in chunks get the event indices and refresh
    relaxations of one uninterrupted run (eventIndex = initEventCnt + events,
    refresh when (initEventCnt + eventIndex) % interval == 0)"""
    def run(chunks):
        g = grid(tip_small, tsg.FAITHFUL_F32)
        for _ in range(300):
            g.relax(1)
        e = evap_opts(0)
        e.trajectory = 0
        e.refresh_interval, e.init_event_cnt = 3, 2
        out = []
        for k in chunks:
            out += g.evaporate(e, k)
        return out, g.phi(0)
    one, p1 = run([10])
    parts, p2 = run([4, 1, 5])
    assert [e.event_index for e in one] == list(range(3, 13))
    assert [e.event_index for e in parts] == list(range(3, 13))
    assert [e.node for e in one] == [e.node for e in parts]
    refresh = [e.event_index for e in one if e.global_iter != 0]
    assert refresh == [i for i in range(3, 13) if (2 + i) % 3 == 0]
    assert [e.global_iter for e in one] == [e.global_iter for e in parts]
    assert same_bits(p1, p2, np.float32)
    # the oracle counts the same way
    o = oracle(tip_small, O.F32)
    for _ in range(300):
        o.relax(1)
    oo = O.OraEvapOpts(prob_mode=0, evap_mode=0, vacuum_id=0, relax_threshold=1e-2, relax_shell=3, relax_cycle=5,
                       relax_queue=5, relax_global_cycles=1, refresh_interval=3, refresh_threshold=1e-3,
                       refresh_cycle=1, refresh_queue=0, trajectory=0, integrator=1, stepper=1, dt0=1e-14,
                       dt_min=1e-16, err_limit=1e-6, voltage_queue=250, shrink_limit=1.0, inplace_order=0,
                       init_event_cnt=2)
    eo = o.evaporate(oo, 4) + o.evaporate(oo, 6)
    assert [e.event_index for e in eo] == list(range(3, 13))
    assert [e.node for e in eo] == [e.node for e in one]
    assert [e.refresh_iter for e in eo] == [e.global_iter for e in one]


@pytest.mark.parametrize("mode,omode,dt", precisions())
def test_local_relax_of_a_shell_with_neumann_rows(tip_small, mode, omode, dt):
    """localRelax around an atom whose 3-shell holds Neumann nodes (the box
    floor): a Neumann row reads higher-index neighbours that do not read it,
    so its in-place cycles must see their old values"""
    node = 1141
    g = grid(tip_small, mode)
    o = oracle(tip_small, omode)
    for _ in range(300):
        g.relax(1)
        o.relax(1)
    fl = o.flags()
    shell = o.local_nodes(node, 3)
    assert ((fl[shell] & 4) != 0).sum() > 0   # the case this test is for
    g.surface_init(0)
    o.surface_init(0)
    g.take_away(node)
    o.take_away(node)
    thr = float(np.float32(1e-2))
    assert g.local_relax(node, 3, thr, 5, 5) == o.local_relax_threshold(node, 3, thr, 5, 5)
    assert g.switches() == o.switches()
    assert same_bits(g.phi(0), o.phi(0), dt) and same_bits(g.phi(1), o.phi(1), dt)
