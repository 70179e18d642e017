"""ctypes front end of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker.  See tapsim_oracle.h for what it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
ROOT = os.path.dirname(HERE)

F32, F64 = 0, 1


class OraSpecies(ctypes.Structure):
    _fields_ = [
        ("defined", ctypes.c_int), ("charge_density", ctypes.c_float), ("epsilon", ctypes.c_float),
        ("removable", ctypes.c_int), ("neumann", ctypes.c_int), ("dirichlet", ctypes.c_int),
        ("phi", ctypes.c_float), ("mass", ctypes.c_float), ("evap_charge", ctypes.c_int),
        ("evap_field", ctypes.c_float), ("evap_energy", ctypes.c_float),
    ]


class OraEvapOpts(ctypes.Structure):
    _fields_ = [
        ("prob_mode", ctypes.c_int), ("evap_mode", ctypes.c_int), ("vacuum_id", ctypes.c_int),
        ("relax_threshold", ctypes.c_float), ("relax_shell", ctypes.c_int), ("relax_cycle", ctypes.c_int),
        ("relax_queue", ctypes.c_int), ("relax_global_cycles", ctypes.c_int),
        ("refresh_interval", ctypes.c_int), ("refresh_threshold", ctypes.c_float), ("refresh_cycle", ctypes.c_int),
        ("refresh_queue", ctypes.c_int), ("trajectory", ctypes.c_int), ("integrator", ctypes.c_int),
        ("stepper", ctypes.c_int), ("dt0", ctypes.c_float), ("dt_min", ctypes.c_float), ("err_limit", ctypes.c_float),
        ("voltage_queue", ctypes.c_uint), ("shrink_limit", ctypes.c_float), ("inplace_order", ctypes.c_int),
        ("thermal_velocity", ctypes.c_int), ("update_coordination", ctypes.c_int), ("init_event_cnt", ctypes.c_int),
        ("detector_z", ctypes.c_float), ("ion_table", ctypes.c_int), ("ion_seed", ctypes.c_uint),
    ]


class OraEvent(ctypes.Structure):
    _fields_ = [
        ("node", ctypes.c_int), ("id", ctypes.c_int), ("probability", ctypes.c_float),
        ("pot_before", ctypes.c_float), ("field_before", ctypes.c_float * 3),
        ("pot_after", ctypes.c_float), ("field_after", ctypes.c_float * 3),
        ("voltage", ctypes.c_float), ("time_scale", ctypes.c_float),
        ("start", ctypes.c_float * 3), ("stop", ctypes.c_float * 3), ("tof", ctypes.c_float),
        ("traj_len", ctypes.c_int), ("traj_status", ctypes.c_int),
        ("local_iter", ctypes.c_longlong), ("refresh_iter", ctypes.c_longlong),
        ("surface_size", ctypes.c_int), ("scaling_ref", ctypes.c_float), ("apex", ctypes.c_float * 3),
        ("normal", ctypes.c_float * 3), ("traj_root_diff", ctypes.c_int), ("event_index", ctypes.c_longlong),
        ("ion_mass", ctypes.c_float), ("ion_charge", ctypes.c_int),
        ("err_pos", ctypes.c_float * 3), ("err_vel", ctypes.c_float * 3),
    ]


_P = ctypes.c_void_p
_SIGS = {
    "ora_create": ([ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_int, _P, _P, _P, ctypes.c_float], _P),
    "ora_destroy": ([_P], None),
    "ora_last_error": ([], ctypes.c_char_p),
    "ora_sync": ([_P], ctypes.c_int),
    "ora_set_couplings": ([_P, _P], ctypes.c_int),
    "ora_resync": ([_P, ctypes.c_int], ctypes.c_int),
    "ora_set_id": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "ora_nnz": ([_P], ctypes.c_int),
    "ora_get_csr": ([_P, _P, _P], None),
    "ora_get_coupling": ([_P, _P], None),
    "ora_get_charge": ([_P, _P], None),
    "ora_get_flags": ([_P, _P], None),
    "ora_get_ids": ([_P, _P], None),
    "ora_get_assoc": ([_P, _P], None),
    "ora_get_boundary": ([_P, _P], None),
    "ora_get_phi": ([_P, ctypes.c_int, _P], None),
    "ora_set_phi": ([_P, _P, _P, ctypes.c_int, ctypes.c_int], None),
    "ora_phi_switch": ([_P], ctypes.c_int),
    "ora_local_switch": ([_P], ctypes.c_int),
    "ora_potential": ([_P, ctypes.c_int], ctypes.c_double),
    "ora_set_inplace_order": ([_P, ctypes.c_int, _P, ctypes.c_int], None),
    "ora_relax": ([_P, ctypes.c_int, _P], ctypes.c_double),
    "ora_relax_threshold": ([_P, ctypes.c_double, ctypes.c_int, ctypes.c_int, _P, _P, _P], ctypes.c_longlong),
    "ora_find_local_nodes": ([_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int], ctypes.c_int),
    "ora_local_relax": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P], ctypes.c_double),
    "ora_local_relax_threshold": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _P],
                                  ctypes.c_longlong),
    "ora_field_o1": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "ora_field_o2": ([_P, ctypes.c_int, _P, _P], ctypes.c_int),
    "ora_voronoi_area": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_double], ctypes.c_double),
    "ora_voronoi_volume": ([_P, ctypes.c_int, ctypes.c_double], ctypes.c_double),
    "ora_distance": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_double),
    "ora_surface_init": ([_P, ctypes.c_int], ctypes.c_int),
    "ora_surface_size": ([_P], ctypes.c_int),
    "ora_surface_nodes": ([_P, _P], None),
    "ora_surface_probabilities": ([_P, _P], None),
    "ora_compute_probabilities": ([_P, ctypes.c_int], ctypes.c_int),
    "ora_specific_fields": ([_P], ctypes.c_float),
    "ora_find_candidate": ([_P, ctypes.c_int], ctypes.c_int),
    "ora_take_away": ([_P, ctypes.c_int], ctypes.c_int),
    "ora_surface_normal": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "ora_best_plane": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "ora_srand": ([ctypes.c_uint], None),
    "ora_update_coordination": ([_P], None),
    "ora_find_tetrahedron": ([_P, _P, ctypes.c_int], ctypes.c_int),
    "ora_locate": ([_P, _P], ctypes.c_int),
    "ora_set_root_mode": ([ctypes.c_int], None),
    "ora_fieldline": ([_P, _P, ctypes.c_float, ctypes.c_ulong, ctypes.c_ulong, _P, ctypes.c_int, _P], ctypes.c_int),
    "ora_tet_weights": ([_P, ctypes.c_int, _P, _P], None),
    "ora_point_potential": ([_P, _P, ctypes.c_int], ctypes.c_float),
    "ora_point_field": ([_P, _P, ctypes.c_int, _P], ctypes.c_int),
    "ora_trajectory": ([_P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                        ctypes.c_float, ctypes.c_float, _P, ctypes.c_int, _P, _P], ctypes.c_int),
    "ora_evaporate": ([_P, _P, ctypes.c_int, _P, _P], ctypes.c_int),
    "ora_random_reset": ([_P], None),
    "ora_spherical_preset": ([_P, _P, ctypes.c_float, ctypes.c_float], None),
    "ora_relax_threaded": ([_P, ctypes.c_int, ctypes.c_int, _P, _P], ctypes.c_double),
    "ora_set_threads": ([_P, ctypes.c_int], None),
    "ora_mpnn": ([_P, _P, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "ora_bf16": ([ctypes.c_float], ctypes.c_float),
    "ora_set_ion_table": ([_P, ctypes.c_int, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "ora_ion_pick": ([ctypes.c_uint, ctypes.c_longlong, ctypes.c_int, _P], ctypes.c_int),
    "ora_rk4_analytic": ([ctypes.c_int, _P, ctypes.c_float, _P, _P, ctypes.c_float, ctypes.c_int, _P, _P],
                         ctypes.c_int),
}

_lib = None


def build():
    """Compiles liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-B", "-f", "oracle/Makefile"], cwd=ROOT, check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def species_table(entries):
    table = (OraSpecies * 128)()
    for sid, kw in entries.items():
        s = table[sid]
        s.defined = 1
        s.epsilon = 1.0
        for k, v in kw.items():
            setattr(s, k, v)
    return table


class OracleError(RuntimeError):
    pass


class Oracle:
    """One restated System_3d (config + geometry + grid) in a given precision."""

    def __init__(self, mode, xyz, ids, tets, neighbours, species, temperature=0.0):
        lib = load()
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self.ids0 = np.ascontiguousarray(ids, dtype=np.int16)
        self.tv = np.ascontiguousarray(tets, dtype=np.int32)
        self.tn = np.ascontiguousarray(neighbours, dtype=np.int32)
        self.species = species
        self.n = len(self.xyz)
        self.m = len(self.tv)
        self._h = lib.ora_create(mode, self.n, _ptr(self.xyz), _ptr(self.ids0), self.m, _ptr(self.tv), _ptr(self.tn),
                                 ctypes.byref(species), float(temperature))
        if not self._h:
            raise OracleError(lib.ora_last_error().decode())
        self.mode = mode
        self.nnz = lib.ora_nnz(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ora_destroy(self._h)
            self._h = None

    def _chk(self, rc):
        if rc:
            raise OracleError("oracle error %d: %s" % (rc, load().ora_last_error().decode()))

    def sync(self):
        self._chk(load().ora_sync(self._h))

    def sync_from_couplings(self, coup):
        c = np.ascontiguousarray(coup, dtype=np.float64)
        assert len(c) == self.nnz
        self._chk(load().ora_set_couplings(self._h, _ptr(c)))

    def resync(self, node):
        self._chk(load().ora_resync(self._h, node))

    def set_id(self, node, sid):
        self._chk(load().ora_set_id(self._h, node, sid))

    def csr(self):
        ptr = np.zeros(self.n + 1, np.int32)
        idx = np.zeros(self.nnz, np.int32)
        load().ora_get_csr(self._h, _ptr(ptr), _ptr(idx))
        return ptr, idx

    def couplings(self):
        out = np.zeros(self.nnz, np.float64)
        load().ora_get_coupling(self._h, _ptr(out))
        return out

    def charges(self):
        out = np.zeros(self.n, np.float64)
        load().ora_get_charge(self._h, _ptr(out))
        return out

    def flags(self):
        out = np.zeros(self.n, np.uint8)
        load().ora_get_flags(self._h, _ptr(out))
        return out

    def assoc(self):
        out = np.zeros(self.n, np.int32)
        load().ora_get_assoc(self._h, _ptr(out))
        return out

    def boundary(self):
        out = np.zeros(self.n, np.uint8)
        load().ora_get_boundary(self._h, _ptr(out))
        return out

    def phi(self, buffer):
        out = np.zeros(self.n, np.float64)
        load().ora_get_phi(self._h, buffer, _ptr(out))
        return out

    def set_phi(self, p0, p1, sw=0, lsw=0):
        p0 = np.ascontiguousarray(p0, dtype=np.float64)
        p1 = np.ascontiguousarray(p1, dtype=np.float64)
        load().ora_set_phi(self._h, _ptr(p0), _ptr(p1), sw, lsw)

    def switches(self):
        return load().ora_phi_switch(self._h), load().ora_local_switch(self._h)

    def potential(self, node):
        return load().ora_potential(self._h, node)

    def set_inplace_order(self, kind, colour=None, ncolours=1):
        c = None if colour is None else np.ascontiguousarray(colour, dtype=np.int32)
        load().ora_set_inplace_order(self._h, kind, None if c is None else _ptr(c), ncolours)

    def relax(self, cycles=1):
        err = ctypes.c_int()
        d = load().ora_relax(self._h, cycles, ctypes.byref(err))
        self._chk(err.value)
        return d

    def random_reset(self):
        load().ora_random_reset(self._h)

    def spherical_preset(self, center, pmin, pmax):
        c = np.ascontiguousarray(center, np.float32)
        load().ora_spherical_preset(self._h, _ptr(c), pmin, pmax)

    def set_threads(self, threads):
        """worker threads of the cycleSize-1 threshold relaxations (the event
        loop's refresh relaxation): the reference's chunked threading"""
        load().ora_set_threads(self._h, threads)

    def relax_threaded(self, sweeps, threads):
        """`sweeps` relax(1) calls on `threads` workers (FP64): (deviation, seconds)."""
        sec = ctypes.c_double()
        err = ctypes.c_int()
        dev = load().ora_relax_threaded(self._h, sweeps, threads, ctypes.byref(sec), ctypes.byref(err))
        self._chk(err.value)
        return dev, sec.value

    def relax_threshold(self, threshold, cycles=1, queue=0):
        err, fd, fs = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        it = load().ora_relax_threshold(self._h, threshold, cycles, queue, ctypes.byref(fd), ctypes.byref(fs),
                                        ctypes.byref(err))
        self._chk(err.value)
        return it, fd.value, fs.value

    def local_nodes(self, node, order):
        out = np.zeros(self.n, np.int32)
        c = load().ora_find_local_nodes(self._h, node, order, _ptr(out), self.n)
        return out[:c].copy()

    def local_relax(self, node, order, cycles):
        err = ctypes.c_int()
        d = load().ora_local_relax(self._h, node, order, cycles, ctypes.byref(err))
        self._chk(err.value)
        return d

    def local_relax_threshold(self, node, order, threshold, cycles, queue=0):
        err = ctypes.c_int()
        it = load().ora_local_relax_threshold(self._h, node, order, threshold, cycles, queue, ctypes.byref(err))
        self._chk(err.value)
        return it

    def field_o1(self, node):
        out = np.zeros(3, np.float64)
        self._chk(load().ora_field_o1(self._h, node, _ptr(out)))
        return out

    def field_o2(self, node):
        out = np.zeros(3, np.float64)
        q = ctypes.c_double()
        self._chk(load().ora_field_o2(self._h, node, _ptr(out), ctypes.byref(q)))
        return out, q.value

    def voronoi_area(self, n1, n2, inf):
        return load().ora_voronoi_area(self._h, n1, n2, inf)

    def voronoi_volume(self, node, inf):
        return load().ora_voronoi_volume(self._h, node, inf)

    def surface_init(self, vacuum_id):
        self._chk(load().ora_surface_init(self._h, vacuum_id))

    def surface_nodes(self):
        c = load().ora_surface_size(self._h)
        out = np.zeros(c, np.int32)
        load().ora_surface_nodes(self._h, _ptr(out))
        return out

    def surface_probabilities(self):
        c = load().ora_surface_size(self._h)
        out = np.zeros(c, np.float32)
        load().ora_surface_probabilities(self._h, _ptr(out))
        return out

    def surface_normal(self, node):
        """Surface_3d::Table::normal (surface_3d.cpp:123-169)"""
        out = np.zeros(3, np.float32)
        self._chk(load().ora_surface_normal(self._h, node, _ptr(out)))
        return out

    def compute_probabilities(self, mode=0):
        self._chk(load().ora_compute_probabilities(self._h, mode))

    def specific_fields(self):
        return load().ora_specific_fields(self._h)

    def candidate(self, mode=0):
        return load().ora_find_candidate(self._h, mode)

    def take_away(self, node):
        self._chk(load().ora_take_away(self._h, node))

    def ids(self):
        out = np.zeros(self.n, np.int16)
        load().ora_get_ids(self._h, _ptr(out))
        return out

    def update_coordination(self):
        load().ora_update_coordination(self._h)

    def srand(self, seed):
        load().ora_srand(seed)

    def find_tetrahedron(self, pos, guess=None):
        """findTetrahedron(pos, guess); without a guess the one-argument form
        (tetGuess's random sample, then the walk)"""
        p = np.ascontiguousarray(pos, dtype=np.float32)
        if guess is None:
            return load().ora_locate(self._h, _ptr(p))
        return load().ora_find_tetrahedron(self._h, _ptr(p), guess)

    def fieldline(self, start, distance=-1.0, iterations=5000, stepper_limit=1000, cap=100000):
        """Fieldline_3d from `start`: (status, points (k, 7): x, y, z, Ex, Ey, Ez, tet)"""
        p = np.ascontiguousarray(start, np.float32)
        out = np.zeros((cap, 7), np.float32)
        ln = ctypes.c_int()
        st = load().ora_fieldline(self._h, _ptr(p), distance, iterations, stepper_limit, _ptr(out), cap,
                                  ctypes.byref(ln))
        if st < 0:
            raise OracleError(load().ora_last_error().decode())
        return st, out[:min(ln.value, cap)].copy()

    def trajectory(self, node, charge, mass, integrator=1, stepper=1, dt0=1e-14, dt_min=1e-16, err=1e-6, cap=200000,
                   with_error=False):
        """(status, phases (k, 8): t, x, y, z, vx, vy, vz, tet) of one flight;
        with_error adds the final error estimate (position xyz, velocity xyz)"""
        out = np.zeros((cap, 8), np.float32)
        est = np.zeros(6, np.float32)
        ln = ctypes.c_int()
        st = load().ora_trajectory(self._h, node, charge, mass, integrator, stepper, dt0, dt_min, err, _ptr(out), cap,
                                   ctypes.byref(ln), _ptr(est))
        if with_error:
            return st, out[:min(ln.value, cap)].copy(), est
        return st, out[:min(ln.value, cap)].copy()

    def set_ion_table(self, species, mass, charge, weight):
        """the species' ion table (mass amu, charge e, weight per entry; at most 8)"""
        m = np.ascontiguousarray(mass, np.float32)
        q = np.ascontiguousarray(charge, np.int32)
        w = np.ascontiguousarray(weight, np.float32)
        self._chk(load().ora_set_ion_table(self._h, species, len(m), _ptr(m), _ptr(q), _ptr(w)))

    def evaporate(self, opts, k):
        out = (OraEvent * max(1, k))()
        done = ctypes.c_int()
        self._chk(load().ora_evaporate(self._h, ctypes.byref(opts), k, out, ctypes.byref(done)))
        return [out[i] for i in range(done.value)]


# ---------------------------------------------------------------- warm-start network

MPNN_IN, MPNN_HID, MPNN_LAYERS = 32, 64, 3
MPNN_W_IN = 0
MPNN_B_IN = MPNN_IN * MPNN_HID
MPNN_LAYER0 = MPNN_B_IN + MPNN_HID
MPNN_LAYER_STRIDE = 2 * MPNN_HID * MPNN_HID + MPNN_HID
MPNN_W_OUT = MPNN_LAYER0 + MPNN_LAYERS * MPNN_LAYER_STRIDE
MPNN_B_OUT = MPNN_W_OUT + MPNN_HID
MPNN_NPARAM = MPNN_B_OUT + 1


class _MpnnGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("a", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("t", ctypes.c_void_p), ("mask", ctypes.c_void_p)]


def mpnn(ptr, col, a, x, t, mask, params, emulate=True, grad=True):
    """The warm-start network on one (or a block-diagonal batch of) graph(s):
    returns (loss, y, grad or None) in double."""
    lib = load()
    ptr = np.ascontiguousarray(ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    a = np.ascontiguousarray(a, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    mask = np.ascontiguousarray(mask, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    assert x.shape == (len(ptr) - 1, MPNN_IN) and params.shape == (MPNN_NPARAM,)
    g = _MpnnGraph(len(ptr) - 1, _ptr(ptr), _ptr(col), _ptr(a), _ptr(x), _ptr(t), _ptr(mask))
    loss = ctypes.c_double()
    y = np.zeros(len(ptr) - 1)
    gr = np.zeros(MPNN_NPARAM) if grad else None
    rc = lib.ora_mpnn(ctypes.byref(g), _ptr(params), 1 if emulate else 0, ctypes.byref(loss), _ptr(y),
                      _ptr(gr) if grad else None)
    if rc:
        raise OracleError("ora_mpnn failed (%d)" % rc)
    return loss.value, y, gr


def bf16(v):
    """Round-to-nearest-even to bfloat16, as float32 values."""
    v = np.ascontiguousarray(v, np.float32)
    u = v.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def best_plane(offsets):
    """Geometry_3d::find_best_plane (geometry_3d.cpp:1806-1870) of float
    offsets (k, 3): unit eigenvector of the smallest scatter eigenvalue."""
    off = np.ascontiguousarray(offsets, np.float32).reshape(-1, 3)
    out = np.zeros(3, np.float32)
    rc = load().ora_best_plane(_ptr(off), len(off), _ptr(out))
    if rc:
        raise RuntimeError(load().ora_last_error().decode())
    return out


def ion_pick(seed, event_index, weight):
    """the index of the ion drawn for an event from a table with these weights"""
    w = np.ascontiguousarray(weight, np.float32)
    return load().ora_ion_pick(seed, event_index, len(w), _ptr(w))


def rk4_analytic(kind, g, qm, p0, v0, delta, steps):
    """the fourth-order Runge-Kutta step of the flights in an analytic field
    (kind 0: E = g; kind 1: E = -(g0 x, g1 y, g2 z)): (t, p, v) after `steps`
    steps of `delta` and the last step's error estimate (p, v)"""
    f = lambda v: np.ascontiguousarray(v, np.float32)  # noqa: E731
    out = np.zeros(7, np.float32)
    err = np.zeros(6, np.float32)
    load().ora_rk4_analytic(kind, _ptr(f(g)), qm, _ptr(f(p0)), _ptr(f(v0)), delta, steps, _ptr(out), _ptr(err))
    return out[0], out[1:4].copy(), out[4:7].copy(), err[:3].copy(), err[3:].copy()


def set_root_mode(mode):
    """The trajectories' step-size root: 0 glibc's powf(x, 0.25f) as the
    reference, 1 the correctly rounded fourth root that the device uses."""
    load().ora_set_root_mode(mode)
