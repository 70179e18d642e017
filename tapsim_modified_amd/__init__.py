"""MI355X (gfx950) per-event electrostatics of TAPSim.

Python mirror of the reference's C++ call surface for this path
(Grid_3d::Table, Surface_3d::Table, Process::evaporation in
ashtonmv/tapsim_modified tapsim/grid_3d.h, surface_3d.h, process.h) on top of
the C ABI in include/tapsim_gpu.h.  The work runs in the in-tree HIP library
libtapsim_gpu.so; there is no CPU fallback: a missing library or GPU raises.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "load", "TsgError", "RelaxationFault", "SpecimenParams", "Species", "Mesh", "Grid3D",
    "generate_specimen", "EvapOptions", "PLAIN_F64", "FAITHFUL_F32",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtapsim_gpu.so")

PLAIN_F64 = 0
FAITHFUL_F32 = 1

FLAG_BOUNDARY, FLAG_DIRICHLET, FLAG_NEUMANN = 1, 2, 4

_lib = None


class TsgError(RuntimeError):
    """A failed call of the C ABI (status code and message)."""

    def __init__(self, code, msg):
        super().__init__("tsg error %d: %s" % (code, msg))
        self.code = code


class RelaxationFault(TsgError):
    """Grid_3d::Table::RelaxationFault (tapsim/grid_3d.h:103-129)."""

    def __init__(self, iteration, deviation, slope):
        TsgError.__init__(self, -5, "relaxation stagnated at iteration %d (deviation %g, slope %g)"
                          % (iteration, deviation, slope))
        self.iteration, self.deviation, self.slope = iteration, deviation, slope


class SpecimenParams(ctypes.Structure):
    _fields_ = [
        ("a", ctypes.c_double), ("radius", ctypes.c_double), ("shank", ctypes.c_double),
        ("box", ctypes.c_double * 3), ("shell", ctypes.c_double), ("vacuum_target", ctypes.c_longlong),
        ("decay", ctypes.c_double), ("face_spacing", ctypes.c_double), ("cluster_density", ctypes.c_double),
        ("cluster_radius", ctypes.c_double), ("seed", ctypes.c_ulonglong),
        ("id_vacuum", ctypes.c_int), ("id_top", ctypes.c_int), ("id_bottom", ctypes.c_int),
        ("id_sides", ctypes.c_int), ("id_matrix", ctypes.c_int), ("id_solute", ctypes.c_int),
    ]


class Species(ctypes.Structure):
    """Configuration::NodeData (tapsim/configuration.h:100-150)."""
    _fields_ = [
        ("defined", ctypes.c_int), ("charge_density", ctypes.c_float), ("epsilon", ctypes.c_float),
        ("removable", ctypes.c_int), ("neumann", ctypes.c_int), ("dirichlet", ctypes.c_int),
        ("phi", ctypes.c_float), ("mass", ctypes.c_float), ("evap_charge", ctypes.c_int),
        ("evap_field", ctypes.c_float), ("evap_energy", ctypes.c_float),
    ]


class EvapOptions(ctypes.Structure):
    """The hot-path knobs of Process::EvaporationOptions (tapsim/process.h:39-122)."""
    _fields_ = [
        ("prob_mode", ctypes.c_int), ("evap_mode", ctypes.c_int), ("vacuum_id", ctypes.c_int),
        ("relax_threshold", ctypes.c_double), ("relax_shell", ctypes.c_int), ("relax_cycle", ctypes.c_int),
        ("relax_queue", ctypes.c_int), ("relax_global_cycles", ctypes.c_int), ("refresh_interval", ctypes.c_int),
        ("refresh_threshold", ctypes.c_double), ("refresh_cycle", ctypes.c_int), ("refresh_queue", ctypes.c_int),
        ("exact_probability", ctypes.c_int), ("method", ctypes.c_int), ("accel_tol", ctypes.c_double),
        ("trajectory", ctypes.c_int), ("integrator", ctypes.c_int), ("stepper", ctypes.c_int),
        ("dt0", ctypes.c_float), ("dt_min", ctypes.c_float), ("err_limit", ctypes.c_float),
        ("voltage_queue", ctypes.c_int), ("shrink_limit", ctypes.c_float), ("update_coordination", ctypes.c_int),
        ("thermal_velocity", ctypes.c_int), ("accel_solver", ctypes.c_int), ("init_event_cnt", ctypes.c_int),
        ("detector_z", ctypes.c_float), ("ion_table", ctypes.c_int), ("ion_seed", ctypes.c_uint),
    ]

    def __init__(self, **kw):
        super().__init__()
        self.integrator, self.stepper = 1, 1
        self.dt0, self.dt_min, self.err_limit = 1e-14, 1e-16, 1e-6
        self.voltage_queue, self.shrink_limit = 250, 1.0
        self.thermal_velocity = 0
        self.update_coordination = 0
        self.accel_tol = 1e-8
        for k, v in kw.items():
            setattr(self, k, v)


class Event(ctypes.Structure):
    _fields_ = [
        ("node", ctypes.c_int), ("id", ctypes.c_int), ("probability", ctypes.c_float),
        ("pot_before", ctypes.c_double), ("field_before", ctypes.c_double * 3),
        ("pot_after", ctypes.c_double), ("field_after", ctypes.c_double * 3),
        ("local_iter", ctypes.c_longlong), ("global_iter", ctypes.c_longlong),
        ("surface_size", ctypes.c_int), ("scaling_ref", ctypes.c_double), ("near_tie", ctypes.c_int),
        ("warm_start", ctypes.c_int), ("residual_cold", ctypes.c_double), ("residual_warm", ctypes.c_double),
        ("start", ctypes.c_float * 3), ("stop", ctypes.c_float * 3), ("apex", ctypes.c_float * 3),
        ("normal", ctypes.c_float * 3),
        ("voltage", ctypes.c_float), ("time_scale", ctypes.c_float), ("tof", ctypes.c_float),
        ("traj_len", ctypes.c_int), ("traj_status", ctypes.c_int), ("event_index", ctypes.c_longlong),
        ("ion_mass", ctypes.c_float), ("ion_charge", ctypes.c_int),
        ("err_pos", ctypes.c_float * 3), ("err_vel", ctypes.c_float * 3),
    ]


# exported symbols and their argument types
_P = ctypes.c_void_p
_SIGS = {
    "tsg_last_error": ([], ctypes.c_char_p),
    "tsg_version": ([], ctypes.c_int),
    "tsg_free": ([_P], None),
    "tsg_specimen_generate": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "tsg_mesh_create": ([ctypes.c_int, _P, _P], ctypes.c_int),
    "tsg_mesh_create_from_tets": ([ctypes.c_int, _P, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "tsg_mesh_info": ([_P, _P, _P, _P], ctypes.c_int),
    "tsg_mesh_get_tets": ([_P, _P, _P], ctypes.c_int),
    "tsg_mesh_get_csr": ([_P, _P, _P], ctypes.c_int),
    "tsg_mesh_get_nodeinfo": ([_P, _P, _P], ctypes.c_int),
    "tsg_mesh_destroy": ([_P], None),
    "tsg_grid_create": ([_P, _P, _P, ctypes.c_float, ctypes.c_int, _P], ctypes.c_int),
    "tsg_grid_destroy": ([_P], None),
    "tsg_grid_info": ([_P, _P, _P, _P], ctypes.c_int),
    "tsg_relax_cycles": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_relax_threshold": ([_P, ctypes.c_double, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P], ctypes.c_int),
    "tsg_local_relax": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "tsg_get_phi": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_set_phi": ([_P, _P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "tsg_phi_switches": ([_P, _P, _P], ctypes.c_int),
    "tsg_potential": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_field_o1": ([_P, ctypes.c_int, _P, _P], ctypes.c_int),
    "tsg_field_o2": ([_P, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "tsg_set_ion_table": ([_P, ctypes.c_int, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "tsg_grid_extents": ([_P, _P], ctypes.c_int),
    "tsg_evaporate_traj": ([_P, _P, ctypes.c_int, _P, _P, ctypes.c_int, _P, _P], ctypes.c_int),
    "tsg_field_o1_noalloc": ([_P, ctypes.c_int, _P, _P], ctypes.c_int),
    "tsg_get_couplings": ([_P, _P], ctypes.c_int),
    "tsg_get_charges": ([_P, _P], ctypes.c_int),
    "tsg_get_flags": ([_P, _P], ctypes.c_int),
    "tsg_get_ids": ([_P, _P], ctypes.c_int),
    "tsg_get_areas": ([_P, _P, _P, _P], ctypes.c_int),
    "tsg_set_id": ([_P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "tsg_surface_init": ([_P, ctypes.c_int], ctypes.c_int),
    "tsg_surface_size": ([_P, _P], ctypes.c_int),
    "tsg_surface_nodes": ([_P, _P], ctypes.c_int),
    "tsg_surface_evaluate": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_surface_values": ([_P, _P], ctypes.c_int),
    "tsg_surface_probabilities": ([_P, _P], ctypes.c_int),
    "tsg_surface_candidate": ([_P, _P, _P], ctypes.c_int),
    "tsg_take_away": ([_P, ctypes.c_int], ctypes.c_int),
    "tsg_surface_normal": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_update_coordination": ([_P], ctypes.c_int),
    "tsg_best_plane": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_rand": ([ctypes.c_int, _P], ctypes.c_int),
    "tsg_fieldlines": ([_P, ctypes.c_int, _P, ctypes.c_float, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int,
                        ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "tsg_evaporate": ([_P, _P, ctypes.c_int, _P, _P], ctypes.c_int),
    "tsg_sweep_stats": ([_P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "tsg_accel_setup": ([_P, ctypes.c_double, _P], ctypes.c_int),
    "tsg_time_kernel": ([_P, ctypes.c_int, ctypes.c_int, _P, _P], ctypes.c_int),
    "tsg_accel_relax": ([_P, ctypes.c_double, _P, _P, _P], ctypes.c_int),
    "tsg_accel_multigrid": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_phase_times": ([_P, _P, ctypes.c_int, _P], ctypes.c_int),
    "tsg_mpnn_create": ([ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_uint, _P], ctypes.c_int),
    "tsg_mpnn_destroy": ([_P], None),
    "tsg_mpnn_set_params": ([_P, _P], ctypes.c_int),
    "tsg_mpnn_get_params": ([_P, _P], ctypes.c_int),
    "tsg_mpnn_eval": ([_P, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "tsg_mpnn_add_graph": ([_P, ctypes.c_int, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "tsg_mpnn_dataset": ([_P, _P, _P, _P], ctypes.c_int),
    "tsg_mpnn_clear": ([_P], ctypes.c_int),
    "tsg_mpnn_train": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint, _P,
                        _P], ctypes.c_int),
    "tsg_mpnn_evaluate": ([_P, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "tsg_grid_attach_mpnn": ([_P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "tsg_surface_evaluate_model": ([_P, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "tsg_srand": ([ctypes.c_uint], None),
    "tsg_reset_potentials": ([_P, ctypes.c_int, _P, ctypes.c_float, ctypes.c_float], ctypes.c_int),
    "tsg_copy_probe": ([ctypes.c_longlong, ctypes.c_int, _P], ctypes.c_int),
    "tsg_trajectories": ([_P, ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_int, _P, _P, _P, _P], ctypes.c_int),
}


def load():
    """Loads libtapsim_gpu.so (built by __graft_entry__.build()); raises when absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TsgError(-2, "libtapsim_gpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGS)


def copy_probe(nbytes, reps=1):
    """Average time (ms) of a plain device copy of nbytes (the counter calibration kernel)."""
    ms = ctypes.c_double()
    _chk(load().tsg_copy_probe(nbytes, reps, ctypes.byref(ms)))
    return ms.value


def best_plane(offsets):
    """Geometry_3d::find_best_plane (geometry_3d.cpp:1806-1870) of (k, 3)
    float offsets: unit normal of the least-squares plane (host code)."""
    off = np.ascontiguousarray(offsets, np.float32).reshape(-1, 3)
    out = np.zeros(3, np.float32)
    _chk(load().tsg_best_plane(_ptr(off), len(off), _ptr(out)))
    return out


def srand(seed):
    """Seeds the process's rand() stream (the Monte Carlo choice, randomReset,
    thermal start velocities and the point-location walks draw from it)."""
    load().tsg_srand(seed)


def rand(count=1):
    """The next draws of the process's rand() stream."""
    out = np.zeros(count, np.int32)
    _chk(load().tsg_rand(count, _ptr(out)))
    return out


def _chk(rc):
    if rc != 0:
        raise TsgError(rc, load().tsg_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def generate_specimen(params):
    """Seeded BCC [110] tip node set; returns (xyz float32 (n,3), ids int16 (n,), counts)."""
    lib = load()
    n = ctypes.c_int()
    xyz = ctypes.POINTER(ctypes.c_float)()
    ids = ctypes.POINTER(ctypes.c_short)()
    counts = (ctypes.c_longlong * 5)()
    _chk(lib.tsg_specimen_generate(ctypes.byref(params), ctypes.byref(n), ctypes.byref(xyz), ctypes.byref(ids), counts))
    try:
        X = np.ctypeslib.as_array(xyz, (n.value * 3,)).copy().reshape(-1, 3)
        I = np.ctypeslib.as_array(ids, (n.value,)).copy()
    finally:
        lib.tsg_free(ctypes.cast(xyz, ctypes.c_void_p))
        lib.tsg_free(ctypes.cast(ids, ctypes.c_void_p))
    keys = ("atoms", "solute", "vacuum_lattice", "vacuum_random", "box")
    return X, I, dict(zip(keys, list(counts)))


class Mesh:
    """Delaunay tetrahedralisation of a node set plus its topology
    (Geometry_3d::Table after tetrahedralize, tapsim/geometry_3d.cpp:759-810)."""

    def __init__(self, xyz, tets=None, neighbours=None):
        lib = load()
        self._lib = lib
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        h = ctypes.c_void_p()
        if tets is None:
            _chk(lib.tsg_mesh_create(len(self.xyz), _ptr(self.xyz), ctypes.byref(h)))
        else:
            tv = np.ascontiguousarray(tets, dtype=np.int32)
            tn = np.ascontiguousarray(neighbours, dtype=np.int32)
            _chk(lib.tsg_mesh_create_from_tets(len(self.xyz), _ptr(self.xyz), len(tv), _ptr(tv), _ptr(tn), ctypes.byref(h)))
        self._h = h
        n, m, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        _chk(lib.tsg_mesh_info(h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(nnz)))
        self.n, self.m, self.nnz = n.value, m.value, nnz.value

    def tets(self):
        tv = np.zeros((self.m, 4), np.int32)
        tn = np.zeros((self.m, 4), np.int32)
        _chk(load().tsg_mesh_get_tets(self._h, _ptr(tv), _ptr(tn)))
        return tv, tn

    def csr(self):
        ptr = np.zeros(self.n + 1, np.int32)
        idx = np.zeros(self.nnz, np.int32)
        _chk(load().tsg_mesh_get_csr(self._h, _ptr(ptr), _ptr(idx)))
        return ptr, idx

    def nodeinfo(self):
        assoc = np.zeros(self.n, np.int32)
        bnd = np.zeros(self.n, np.uint8)
        _chk(load().tsg_mesh_get_nodeinfo(self._h, _ptr(assoc), _ptr(bnd)))
        return assoc, bnd

    def __del__(self):
        lib = getattr(self, "_lib", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.tsg_mesh_destroy(self._h)
            self._h = None


def species_table(entries):
    """entries: {id: dict(field=value)} -> ctypes array of 128 Species."""
    table = (Species * 128)()
    for sid, kw in entries.items():
        s = table[sid]
        s.defined = 1
        s.epsilon = 1.0
        for k, v in kw.items():
            setattr(s, k, v)
    return table


class Grid3D:
    """Device grid with the reference's Grid_3d::Table semantics
    (tapsim/grid_3d.h:100-295): two potential buffers, the _phiSwitch /
    _phiLocalSwitch rules and one-cycle-stale potential()."""

    def __init__(self, mesh, ids, species, temperature=0.0, mode=PLAIN_F64):
        lib = load()
        self._lib = lib
        self.mesh = mesh
        self.ids0 = np.ascontiguousarray(ids, dtype=np.int16)
        self.species = species
        self.mode = mode
        h = ctypes.c_void_p()
        _chk(lib.tsg_grid_create(mesh._h, _ptr(self.ids0), ctypes.byref(species), float(temperature), mode, ctypes.byref(h)))
        self._h = h
        n, nnz, nrows = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
        _chk(lib.tsg_grid_info(h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nrows)))
        self.n, self.nnz = n.value, nnz.value

    def __del__(self):
        # the grid goes before the network it may hold (a raw pointer on the
        # device side); the library handle is the one loaded at construction
        lib = getattr(self, "_lib", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.tsg_grid_destroy(self._h)
            self._h = None
        self._net = None

    # -- Grid_3d::Table::relax / localRelax
    def relax(self, cycles=1):
        dev = ctypes.c_double()
        _chk(load().tsg_relax_cycles(self._h, cycles, ctypes.byref(dev)))
        return dev.value

    def relax_threshold(self, threshold, cycles=1, queue=0, raise_fault=True):
        it, fd, fs, ld = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _chk(load().tsg_relax_threshold(self._h, threshold, cycles, queue, ctypes.byref(it), ctypes.byref(fd),
                                        ctypes.byref(fs), ctypes.byref(ld)))
        if it.value < 0 and raise_fault:
            raise RelaxationFault(-1 - it.value, fd.value, fs.value)
        return it.value

    def local_relax(self, node, order, threshold, cycles, queue=0):
        it = ctypes.c_longlong()
        _chk(load().tsg_local_relax(self._h, node, order, threshold, cycles, queue, ctypes.byref(it)))
        return it.value

    # -- state
    def phi(self, buffer):
        out = np.zeros(self.n, np.float64)
        _chk(load().tsg_get_phi(self._h, buffer, _ptr(out)))
        return out

    def set_phi(self, phi0, phi1, phi_switch=0, local_switch=0):
        p0 = np.ascontiguousarray(phi0, dtype=np.float64)
        p1 = np.ascontiguousarray(phi1, dtype=np.float64)
        _chk(load().tsg_set_phi(self._h, _ptr(p0), _ptr(p1), phi_switch, local_switch))

    def switches(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _chk(load().tsg_phi_switches(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def potential(self, node):
        out = ctypes.c_double()
        _chk(load().tsg_potential(self._h, node, ctypes.byref(out)))
        return out.value

    def field_o1(self, nodes):
        nodes = np.ascontiguousarray(np.atleast_1d(nodes), dtype=np.int32)
        out = np.zeros((len(nodes), 3), np.float64)
        _chk(load().tsg_field_o1(self._h, len(nodes), _ptr(nodes), _ptr(out)))
        return out

    def field_o1_noalloc(self, nodes):
        """field_o1 through tsg_field_o1_noalloc (staging in the library's
        static device array, no device allocation)"""
        nodes = np.ascontiguousarray(np.atleast_1d(nodes), dtype=np.int32)
        out = np.zeros((len(nodes), 3), np.float64)
        _chk(load().tsg_field_o1_noalloc(self._h, len(nodes), _ptr(nodes), _ptr(out)))
        return out

    def field_o2(self, nodes):
        """Grid_3d::Table::field_o2: (fields (k, 3), charges (k,)) of the quadratic fit"""
        nodes = np.ascontiguousarray(np.atleast_1d(nodes), dtype=np.int32)
        out = np.zeros((len(nodes), 3), np.float64)
        q = np.zeros(len(nodes), np.float64)
        _chk(load().tsg_field_o2(self._h, len(nodes), _ptr(nodes), _ptr(out), _ptr(q)))
        return out, q

    def charges(self):
        """Node::charge of every node (computeCharge)"""
        out = np.zeros(self.n, np.float64)
        _chk(load().tsg_get_charges(self._h, _ptr(out)))
        return out

    def couplings(self):
        out = np.zeros(self.nnz, np.float64)
        _chk(load().tsg_get_couplings(self._h, _ptr(out)))
        return out

    def flags(self):
        out = np.zeros(self.n, np.uint8)
        _chk(load().tsg_get_flags(self._h, _ptr(out)))
        return out

    def ids(self):
        out = np.zeros(self.n, np.int16)
        _chk(load().tsg_get_ids(self._h, _ptr(out)))
        return out

    def areas(self):
        h32 = np.zeros(self.nnz, np.float32)
        h64 = np.zeros(self.nnz, np.float64)
        o32 = np.zeros(self.nnz, np.float32)
        _chk(load().tsg_get_areas(self._h, _ptr(h32), _ptr(h64), _ptr(o32)))
        return h32, h64, o32

    def set_id(self, node, sid):
        _chk(load().tsg_set_id(self._h, node, sid))

    # -- Surface_3d
    def surface_init(self, vacuum_id):
        _chk(load().tsg_surface_init(self._h, vacuum_id))

    def surface_nodes(self):
        c = ctypes.c_int()
        _chk(load().tsg_surface_size(self._h, ctypes.byref(c)))
        out = np.zeros(c.value, np.int32)
        _chk(load().tsg_surface_nodes(self._h, _ptr(out)))
        return out

    def surface_evaluate(self, exact_sum=True, mode=0):
        """Specific fields and the probability model's values (mode 0 linear
        field, 1 Boltzmann, 2 linear force, 3 Voronoi flux force); returns
        the scaling reference max |E|/F."""
        ref = ctypes.c_double()
        _chk(load().tsg_surface_evaluate_model(self._h, mode, 1 if exact_sum else 0, ctypes.byref(ref)))
        return ref.value

    def surface_values(self):
        out = np.zeros(len(self.surface_nodes()), np.float64)
        _chk(load().tsg_surface_values(self._h, _ptr(out)))
        return out

    def surface_probabilities(self):
        out = np.zeros(len(self.surface_nodes()), np.float32)
        _chk(load().tsg_surface_probabilities(self._h, _ptr(out)))
        return out

    def candidate(self):
        node, tie = ctypes.c_int(), ctypes.c_int()
        _chk(load().tsg_surface_candidate(self._h, ctypes.byref(node), ctypes.byref(tie)))
        return node.value, tie.value

    def take_away(self, node):
        _chk(load().tsg_take_away(self._h, node))

    def update_coordination(self):
        """the fork's update_coordination on the grid's ids"""
        _chk(load().tsg_update_coordination(self._h))

    def surface_normal(self, node):
        """Surface_3d::Table::normal(node) (surface_3d.cpp:123-169)"""
        out = np.zeros(3, np.float32)
        _chk(load().tsg_surface_normal(self._h, node, _ptr(out)))
        return out

    def sweep_stats(self):
        b, ms = ctypes.c_double(), ctypes.c_double()
        r, e, ref, sw, la = (ctypes.c_longlong() for _ in range(5))
        _chk(load().tsg_sweep_stats(self._h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(e), ctypes.byref(ref),
                                    ctypes.byref(ms), ctypes.byref(sw), ctypes.byref(la)))
        return dict(bytes=b.value, rows=r.value, entries=e.value, referenced=ref.value, ms=ms.value,
                    sweeps=sw.value, launched=la.value)

    def time_kernel(self, kind, reps=200):
        ms, b = ctypes.c_double(), ctypes.c_double()
        _chk(load().tsg_time_kernel(self._h, kind, reps, ctypes.byref(ms), ctypes.byref(b)))
        return ms.value, b.value

    def accel_setup(self, safety=0.5):
        rho = ctypes.c_double()
        _chk(load().tsg_accel_setup(self._h, safety, ctypes.byref(rho)))
        return rho.value

    def accel_relax(self, tol=1e-8):
        outer, inner, res = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_double()
        _chk(load().tsg_accel_relax(self._h, tol, ctypes.byref(outer), ctypes.byref(inner), ctypes.byref(res)))
        return outer.value, inner.value, res.value

    PHASES = ("choice", "take_away", "warm_start", "relaxation", "surface", "flight")

    def phase_times(self):
        """host wall time (ms) of the event loop's phases, accumulated, and the
        number of events timed"""
        ms = np.zeros(6, np.float64)
        ev = ctypes.c_longlong()
        _chk(load().tsg_phase_times(self._h, ms.ctypes.data_as(_P), 6, ctypes.byref(ev)))
        return dict(zip(self.PHASES, ms.tolist())), ev.value

    def accel_multigrid(self, on=True, eager=False):
        """select the multigrid-preconditioned CG correction solve (True) or
        the Chebyshev-Jacobi steps (False); eager launches every kernel on its
        own instead of replaying the captured iteration graph; returns the
        coarse levels built"""
        lv = ctypes.c_int()
        mode = (2 if eager else 1) if on else 0
        _chk(load().tsg_accel_multigrid(self._h, mode, ctypes.byref(lv)))
        return lv.value

    def trajectories(self, nodes, charge, mass, integrator=1, stepper=1, dt0=1e-14, dt_min=1e-16, err=1e-6,
                     cap=4096, with_error=False):
        """Ion trajectories from the given nodes (tsg_trajectories): a list of
        (status, phases[len, 8]) per ion, phases truncated to cap; with_error
        adds each flight's final error estimate (position xyz, velocity xyz)."""
        nodes = np.ascontiguousarray(np.atleast_1d(nodes), np.int32)
        k = len(nodes)
        charge = np.ascontiguousarray(np.broadcast_to(np.float32(charge), (k,)), np.float32)
        mass = np.ascontiguousarray(np.broadcast_to(np.float32(mass), (k,)), np.float32)
        out = np.zeros((k, cap, 8), np.float32)
        ln = np.zeros(k, np.int32)
        st = np.zeros(k, np.int32)
        est = np.zeros((k, 6), np.float32)
        _chk(load().tsg_trajectories(self._h, k, _ptr(nodes), _ptr(charge), _ptr(mass), integrator, stepper, dt0, dt_min,
                                     err, cap, _ptr(out), _ptr(ln), _ptr(st), _ptr(est)))
        if with_error:
            return [(int(st[i]), out[i, :min(int(ln[i]), cap)].copy(), est[i].copy()) for i in range(k)]
        return [(int(st[i]), out[i, :min(int(ln[i]), cap)].copy()) for i in range(k)]

    def fieldlines(self, starts, distance=-1.0, iterations=5000, stepper_limit=1000, process_rand=True, cap=4096):
        """Field lines (tsg_fieldlines) from (k, 3) start points: a list of
        (status, points[len, 7]) with points (x, y, z, Ex, Ey, Ez, tet)."""
        starts = np.ascontiguousarray(np.asarray(starts, np.float32).reshape(-1, 3))
        k = len(starts)
        out = np.zeros((k, cap, 7), np.float32)
        ln = np.zeros(k, np.int32)
        st = np.zeros(k, np.int32)
        _chk(load().tsg_fieldlines(self._h, k, _ptr(starts), distance, iterations, stepper_limit, int(process_rand), cap,
                                   _ptr(out), _ptr(ln), _ptr(st)))
        return [(int(st[i]), out[i, :min(int(ln[i]), cap)].copy()) for i in range(k)]

    def reset_potentials(self, kind=0, center=None, pmin=0.0, pmax=0.0):
        """0 species presets, 1 random (randReset), 2 radial ramp (preset_sphericalPotential)."""
        c = np.ascontiguousarray(center if center is not None else (0.0, 0.0, 0.0), np.float32)
        _chk(load().tsg_reset_potentials(self._h, kind, _ptr(c), pmin, pmax))

    def attach(self, net, record=True, predict=True):
        """Attaches a WarmStartNet (None detaches): events of methods 1 and 2
        record training graphs, method 2 uses the prediction."""
        self._net = net
        _chk(load().tsg_grid_attach_mpnn(self._h, net.handle if net is not None else None, int(record), int(predict)))

    def extents(self):
        """Geometry_3d::Table min, max, constrained min, constrained max (4 x 3 floats)"""
        out = np.zeros(12, np.float32)
        _chk(load().tsg_grid_extents(self._h, _ptr(out)))
        return out.reshape(4, 3)

    def set_ion_table(self, species, mass, charge, weight):
        """the species' ion table for the flights (mass amu, charge e, weight
        per entry; at most 8; empty clears it), used with EvapOptions.ion_table"""
        m = np.ascontiguousarray(mass, np.float32)
        q = np.ascontiguousarray(charge, np.int32)
        w = np.ascontiguousarray(weight, np.float32)
        _chk(load().tsg_set_ion_table(self._h, species, len(m), _ptr(m), _ptr(q), _ptr(w)))

    def evaporate(self, options, k, trajectories=0):
        """Process::evaporation steps I-VI for up to k events (tsg_evaporate);
        trajectories > 0 also returns each event's flight phases (at most that
        many per event, t unscaled) for the trajectory stream
        (tsg_evaporate_traj): (events, [phases (len, 8) per event])"""
        out = (Event * max(1, k))()
        done = ctypes.c_int()
        if trajectories <= 0:
            _chk(load().tsg_evaporate(self._h, ctypes.byref(options), k, out, ctypes.byref(done)))
            return [out[i] for i in range(done.value)]
        cap = int(trajectories)
        phases = np.zeros((max(1, k), cap, 8), np.float32)
        lens = np.zeros(max(1, k), np.int32)
        _chk(load().tsg_evaporate_traj(self._h, ctypes.byref(options), k, out, ctypes.byref(done), cap, _ptr(phases),
                                       _ptr(lens)))
        ev = [out[i] for i in range(done.value)]
        return ev, [phases[i, :min(int(lens[i]), cap)].copy() for i in range(done.value)]


# ---------------------------------------------------------------- learned warm start

MPNN_IN, MPNN_HID, MPNN_LAYERS = 32, 64, 3
MPNN_NPARAM = MPNN_IN * MPNN_HID + MPNN_HID + MPNN_LAYERS * (2 * MPNN_HID * MPNN_HID + MPNN_HID) + MPNN_HID + 1


class WarmStartNet:
    """The warm-start network and its training store in device memory
    (tsg_mpnn_*): forward/backward on host graphs, training on recorded
    events."""

    def __init__(self, max_batch_rows=65536, data_rows=1 << 22, data_edges=1 << 26, seed=1):
        h = ctypes.c_void_p()
        self._lib = load()
        _chk(self._lib.tsg_mpnn_create(max_batch_rows, data_rows, data_edges, seed, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        # at interpreter teardown the module's globals may already be gone:
        # only the library handle loaded at construction is used here
        lib = getattr(self, "_lib", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.tsg_mpnn_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.zeros(MPNN_NPARAM, np.float32)
        _chk(load().tsg_mpnn_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (MPNN_NPARAM,)
        _chk(load().tsg_mpnn_set_params(self._h, _ptr(p)))

    @staticmethod
    def _graph(ptr, col, a, x, t, mask):
        ptr = np.ascontiguousarray(ptr, np.int32)
        n = len(ptr) - 1
        return (n, ptr, np.ascontiguousarray(col, np.int32), np.ascontiguousarray(a, np.float32),
                np.ascontiguousarray(x, np.float32).reshape(n, MPNN_IN), np.ascontiguousarray(t, np.float32),
                np.ascontiguousarray(mask, np.float32))

    def eval(self, ptr, col, a, x, t, mask, grad=True):
        """(loss, y, grad) of one host graph."""
        n, ptr, col, a, x, t, mask = self._graph(ptr, col, a, x, t, mask)
        loss = ctypes.c_double()
        y = np.zeros(n, np.float32)
        g = np.zeros(MPNN_NPARAM, np.float32) if grad else None
        _chk(load().tsg_mpnn_eval(self._h, n, _ptr(ptr), _ptr(col), _ptr(a), _ptr(x), _ptr(t), _ptr(mask),
                                  ctypes.byref(loss), _ptr(y), _ptr(g) if grad else None))
        return loss.value, y, g

    def add_graph(self, ptr, col, a, x, t, mask):
        n, ptr, col, a, x, t, mask = self._graph(ptr, col, a, x, t, mask)
        _chk(load().tsg_mpnn_add_graph(self._h, n, _ptr(ptr), _ptr(col), _ptr(a), _ptr(x), _ptr(t), _ptr(mask)))

    def dataset(self):
        g, r, e = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
        _chk(load().tsg_mpnn_dataset(self._h, ctypes.byref(g), ctypes.byref(r), ctypes.byref(e)))
        return g.value, r.value, e.value

    def clear(self):
        _chk(load().tsg_mpnn_clear(self._h))

    def train(self, g_begin, g_end, epochs, batch_graphs=16, lr=1e-3, seed=1):
        """Adam over graphs [g_begin, g_end) (tsg_mpnn_train), the rate lr
        decayed over the epochs on a cosine; the mean batch loss of the
        first and the last epoch"""
        f, l = ctypes.c_double(), ctypes.c_double()
        _chk(load().tsg_mpnn_train(self._h, g_begin, g_end, epochs, batch_graphs, lr, seed, ctypes.byref(f),
                                   ctypes.byref(l)))
        return f.value, l.value

    def evaluate(self, g_begin, g_end):
        v = ctypes.c_double()
        _chk(load().tsg_mpnn_evaluate(self._h, g_begin, g_end, ctypes.byref(v)))
        return v.value
