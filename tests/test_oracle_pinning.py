"""The oracle against the reference's own outputs for the parts of the path
that tests/golden/tip_small.npz leaves open (tests/golden/pinning.npz,
written by the reference built from its sources, tests/golden/make_golden.py
make_pinning): relax(cycleSize > 1) with its in-place cycles inside a
threshold relaxation and on its own, RelaxationFault, randomReset, the
probability models 1-3 in the event loop, the Euler integrator and the none,
geometry-restricted and adaptive steppers.  Bit-exact unless stated."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import FIXTURE_SPECIES, ROOT


@pytest.fixture(scope="module")
def pin():
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "pinning.npz")))


def make(fx, mode=O.F32, temperature=0.0):
    return O.Oracle(mode, fx["xyz"], fx["ids"], fx["tets"], fx["ref_neigh"], O.species_table(FIXTURE_SPECIES),
                    temperature=temperature)


def bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def evap_opts(prob):
    return O.OraEvapOpts(prob_mode=prob, evap_mode=0, vacuum_id=0, relax_threshold=1e-2, relax_shell=3,
                         relax_cycle=5, relax_queue=5, relax_global_cycles=1, refresh_interval=5,
                         refresh_threshold=1e-3, refresh_cycle=1, refresh_queue=0, trajectory=1, integrator=1,
                         stepper=1, dt0=1e-14, dt_min=1e-16, err_limit=1e-6, voltage_queue=250, shrink_limit=1.0,
                         inplace_order=0)


def test_threshold_relaxation_with_inplace_cycles(tip_small, pin):
    """relax(0.5, 4, 5): each relax(4) call runs Jacobi, in place, Jacobi, in
    place (the table's switch held for the call); then one relax(100)"""
    o = make(tip_small)
    o.sync()
    it, _, _ = o.relax_threshold(0.5, 4, 5)
    assert it == pin["ref_thr4_iter"][0]
    assert bits_equal(o.phi(0), pin["ref_phi0_t4"]) and bits_equal(o.phi(1), pin["ref_phi1_t4"])
    assert bits_equal([o.potential(i) for i in range(o.n)], pin["ref_pot_t4"])
    d = o.relax(100)
    assert d == pin["ref_dev_c100"][0]
    assert bits_equal(o.phi(0), pin["ref_phi0_c100"]) and bits_equal(o.phi(1), pin["ref_phi1_c100"])


def test_relaxation_fault(tip_small, pin):
    """From a float fixed point a negative threshold never ends the loop; a
    full queue of equal deviations has slope 0 and throws RelaxationFault at
    the reference's iteration, with its deviation and slope"""
    o = make(tip_small)
    o.sync()
    for _ in range(2000):
        o.relax(1)
    it, fd, fs = o.relax_threshold(-1.0, 1, 25)
    assert it == pin["ref_fault_iter"][0]           # -1 - RelaxationFault::iteration()
    assert (fd, fs) == tuple(pin["ref_fault"])
    it, fd, fs = o.relax_threshold(-2.0, 4, 5)
    assert it == pin["ref_fault_iter"][1]
    assert (fd, fs) == tuple(pin["ref_fault4"])
    assert bits_equal(o.phi(0), pin["ref_phi0_fault"]) and bits_equal(o.phi(1), pin["ref_phi1_fault"])


def test_random_reset(tip_small, pin):
    """randomReset after srand(7): rand()/RAND_MAX scaled into the species'
    potential range for every non-Dirichlet node, in node order"""
    o = make(tip_small)
    o.sync()
    o.srand(7)
    o.random_reset()
    assert bits_equal(o.phi(0), pin["ref_phi0_rr"]) and bits_equal(o.phi(1), pin["ref_phi1_rr"])


def test_random_reset_fixture_separates_draw_orders(tip_small, pin):
    """the fixture separates the reference's draw order (one rand() per
    non-Dirichlet node, none for Dirichlet nodes) from the plausible variant
    that draws for every node; (in float, rand()/RAND_MAX equals
    rand()/(RAND_MAX+1): RAND_MAX rounds to 2^31)"""
    import ctypes
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    o = make(tip_small)
    o.sync()
    flags = o.flags()
    libc.srand(7)
    wrong = np.zeros(o.n, np.float32)
    for i in range(o.n):
        v = np.float32(libc.rand()) / np.float32(2147483647)
        wrong[i] = np.float32(5000.0) if flags[i] & 2 else v * np.float32(5000.0)
    ref = pin["ref_phi0_rr"]
    assert (wrong.view(np.uint32) != ref.view(np.uint32)).sum() > o.n // 2
    assert np.all((ref >= 0.0) & (ref <= 5000.0))


@pytest.mark.parametrize("prob", [1, 2, 3])
def test_probability_models_in_event_loop(tip_small, pin, prob):
    """10 events with probBoltzmann (at 20000 K, so that the exponentials of
    this small tip's fields stay finite), probLinear_force (field_o2) and
    probVoronoiFluxForce: the sequence, the probabilities and the record"""
    from tests.golden.make_golden import BOLTZMANN_T
    o = make(tip_small, temperature=BOLTZMANN_T if prob == 1 else 0.0)
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.srand(99)
    ev = o.evaporate(evap_opts(prob), 10)
    F = pin["ref_p%d_f" % prob]
    assert np.array_equal([e.node for e in ev], pin["ref_p%d_node" % prob])
    assert np.array_equal([[e.local_iter, e.refresh_iter] for e in ev], pin["ref_p%d_iter" % prob])
    assert bits_equal([e.probability for e in ev], F[:, 0])
    assert bits_equal([e.pot_before for e in ev], F[:, 1])
    assert bits_equal([list(e.field_before) for e in ev], F[:, 2:5])
    assert bits_equal([e.scaling_ref for e in ev], F[:, 21])


@pytest.mark.parametrize("tag,integrator,stepper", [("o1", 0, 1), ("s0", 1, 0), ("s2", 1, 2), ("s3", 1, 3),
                                                    ("o5", 1, 1)])
def test_trajectory_integrators_and_steppers(tip_small, pin, tag, integrator, stepper):
    """every phase, the status and the final error estimate (the root-sum
    of the steps' estimates, moved along by the extrapolation) of one flight"""
    node = int(tip_small["local_node"])
    o = make(tip_small)
    o.sync()
    for _ in range(300):
        o.relax(1)
    o.set_id(node, 0)
    o.resync(node)
    o.relax(1)
    o.srand(99)
    q = 2.0 * np.float32(1.60217733e-19)
    m = np.float32(55.845) * np.float32(1.660538782e-27)
    st, tr, est = o.trajectory(node, q, m, integrator=integrator, stepper=stepper, with_error=True)
    assert st == pin["ref_traj_status_" + tag][0]
    assert bits_equal(tr, pin["ref_traj_" + tag])
    assert bits_equal(est, pin["ref_traj_err_" + tag])
