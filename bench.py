"""Benchmark: evaporation events/s of the per-event electrostatics on one
MI355X (plain FP64 path), with the relaxation sweep's HBM roofline and the
CPU oracle timed beside it.

A step is one evaporation event: choose the atom (first maximum of the
surface probabilities), remove it and re-couple its cell, relax the 3-shell
locally, relax globally to a Jacobi-scaled residual of 1e-8 V, and evaluate
every surface field and probability.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# BCC [110] tips (a = 0.287 nm), Fe-like matrix at 5 kV Dirichlet, grounded box
WORKLOADS = {
    # cylinder r 5 nm, 10 nm + hemisphere, ~89 k atoms, ~150 k graded vacuum, 60x60x120 nm box
    "c1": dict(radius=5e-9, shank=10e-9, box=(60e-9, 60e-9, 120e-9), vacuum=150000, face=6e-9),
    # r 15 nm, shank 40 nm, ~3.0 M atoms + ~1.2 M vacuum
    "c2": dict(radius=15e-9, shank=40e-9, box=(150e-9, 150e-9, 300e-9), vacuum=1200000, face=15e-9),
    # small case for quick checks
    "s": dict(radius=2e-9, shank=3e-9, box=(20e-9, 20e-9, 40e-9), vacuum=15000, face=4e-9),
}

SPECIES = {
    0: dict(),                                       # Vacuum
    1: dict(dirichlet=1),                            # Top, 0 V
    2: dict(neumann=1),                              # Bottom
    3: dict(dirichlet=1),                            # Sides, 0 V
    10: dict(removable=1, dirichlet=1, phi=5000.0, mass=55.845, evap_charge=2, evap_field=3.0e10, evap_energy=1.0),
    11: dict(removable=1, dirichlet=1, phi=5000.0, mass=58.69, evap_charge=2, evap_field=2.5e10, evap_energy=1.0),
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_specimen(tsg, w, seed=20261203):
    p = tsg.SpecimenParams()
    p.a, p.radius, p.shank = 0.287e-9, w["radius"], w["shank"]
    for k in range(3):
        p.box[k] = w["box"][k]
    p.shell, p.vacuum_target, p.decay, p.face_spacing = 2.0, w["vacuum"], 3.0, w["face"]
    p.cluster_density, p.cluster_radius, p.seed = 0.0, 2e-9, seed
    p.id_vacuum, p.id_top, p.id_bottom, p.id_sides, p.id_matrix, p.id_solute = 0, 1, 2, 3, 10, 11
    return tsg.generate_specimen(p)


def event_options(tsg):
    e = tsg.EvapOptions()
    e.prob_mode, e.evap_mode, e.vacuum_id = 0, 0, 0
    e.relax_threshold, e.relax_shell, e.relax_cycle, e.relax_queue = float(np.float32(1e-2)), 3, 5, 5
    e.relax_global_cycles = 1
    e.refresh_interval, e.refresh_threshold, e.refresh_cycle, e.refresh_queue = 1, 1e-8, 1, 0
    e.exact_probability = 1
    return e


def cpu_baseline(tsg, grid, xyz, ids, tets, neigh, sweeps_per_event, budget_s=20.0):
    """The FP64 oracle (one thread) on a bounded sample of the same workload:
    the Jacobi sweep and the surface evaluation are timed and combined with
    the GPU's sweeps per event (the two are bit-identical iterations)."""
    from oracle import oracle as O
    o = O.Oracle(O.F64, xyz, ids, tets, neigh, O.species_table(SPECIES))
    # couplings from the device precompute (the setup the reference's timer also excludes)
    o.sync_from_couplings(grid.couplings())
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < budget_s * 0.75 or k < 2:
        o.relax(1)
        k += 1
    t_sweep = (time.perf_counter() - t0) / k
    o.surface_init(0)
    t1 = time.perf_counter()
    o.specific_fields()
    o.compute_probabilities(0)
    t_surf = time.perf_counter() - t1
    per_event = sweeps_per_event * t_sweep + t_surf
    return dict(value=1.0 / per_event, unit="events/s", cores=1, kind="port",
                sample="%d FP64 Jacobi sweeps (%.3f s each) + one surface evaluation (%.3f s) of the same mesh, "
                       "combined with %.0f sweeps/event" % (k, t_sweep, t_surf, sweeps_per_event))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (see DESIGN.md, sharding)", file=sys.stderr)
        sys.exit(2)

    import tapsim_modified_amd as tsg
    tsg.load()
    w = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    xyz, ids, counts = build_specimen(tsg, w)
    log("specimen %s: %d nodes %s (%.1fs)" % (args.workload, len(xyz), counts, time.perf_counter() - t0))
    t0 = time.perf_counter()
    mesh = tsg.Mesh(xyz)
    log("mesh: %d tets (%.1fs)" % (mesh.m, time.perf_counter() - t0))
    t0 = time.perf_counter()
    grid = tsg.Grid3D(mesh, ids, tsg.species_table(SPECIES), 0.0, tsg.PLAIN_F64)
    log("grid: nnz %d (%.1fs)" % (grid.nnz, time.perf_counter() - t0))
    t0 = time.perf_counter()
    it = grid.relax_threshold(1e-8, 1, 0)
    log("initial relaxation to 1e-8 V: %d sweeps (%.1fs)" % (it, time.perf_counter() - t0))

    opts = event_options(tsg)
    for _ in range(args.warmup):
        grid.evaporate(opts, 1)
    s0 = grid.sweep_stats()
    t0 = time.perf_counter()
    events = []
    for k in range(args.steps):
        events += grid.evaporate(opts, 1)
        if (k + 1) % max(1, args.steps // 5) == 0:
            log("event %d/%d" % (k + 1, args.steps))
    dt = time.perf_counter() - t0
    s1 = grid.sweep_stats()
    sweeps = s1["sweeps"] - s0["sweeps"]
    sweep_ms = (s1["ms"] - s0["ms"]) / max(1, sweeps)
    achieved = s1["bytes"] / (sweep_ms * 1e-3) / 1e9
    res = {
        "metric": "evaporation events/s at residual <=1e-8 V (plain FP64 relaxation)",
        "value": args.steps / dt,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded BCC [110] tip, graded vacuum, own Delaunay mesh)",
        "config": {"workload": "plain FP64 events, %s: %d atoms + %d vacuum nodes, %d unknown rows, %d sweep entries"
                               % (args.workload, counts["atoms"], counts["vacuum_lattice"] + counts["vacuum_random"],
                                  s1["rows"], s1["entries"]),
                   "sweeps_per_event": sweeps / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_jacobi<double>", "bytes_per_sweep": s1["bytes"], "sweep_us": sweep_ms * 1e3},
    }
    if not args.no_cpu:
        tv, tn = mesh.tets()
        res["cpu_baseline"] = cpu_baseline(tsg, grid, xyz, ids, tv, tn, sweeps / args.steps)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
