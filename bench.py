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


def run_events(grid, opts, warmup, steps, tag):
    for _ in range(warmup):
        grid.evaporate(opts, 1)
    s0 = grid.sweep_stats()
    t0 = time.perf_counter()
    events = []
    for k in range(steps):
        events += grid.evaporate(opts, 1)
        if (k + 1) % max(1, steps // 4) == 0:
            log("%s event %d/%d" % (tag, k + 1, steps))
    dt = time.perf_counter() - t0
    s1 = grid.sweep_stats()
    return events, dt, s0, s1


def roofline(grid, kind, kernel):
    ms, nbytes = grid.time_kernel(kind, 200)
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "kernel": kernel, "bytes_per_launch": nbytes, "launch_us": ms * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=None, help="events of the plain FP64 comparison (default: steps)")
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
    species = tsg.species_table(SPECIES)
    grid = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
    log("grid: nnz %d (%.1fs)" % (grid.nnz, time.perf_counter() - t0))
    t0 = time.perf_counter()
    rho = grid.accel_setup(0.5)
    outer, inner, res = grid.accel_relax(1e-8)
    log("initial relaxation to 1e-8 V: rho %.8f, %d refinements, %d correction steps (%.1fs)"
        % (rho, outer, inner, time.perf_counter() - t0))
    phi0, phi1 = grid.phi(0), grid.phi(1)
    sw = grid.switches()

    # plain FP64: the reference's step III (local 3-shell relaxation, one
    # global cycle) followed by a relaxation to a deviation of 1e-8 V
    plain_steps = args.plain_steps if args.plain_steps is not None else args.steps
    popts = event_options(tsg)
    pev, pdt, ps0, ps1 = run_events(grid, popts, args.warmup, plain_steps, "plain")
    psweeps = (ps1["sweeps"] - ps0["sweeps"]) / max(1, plain_steps)
    roof_plain = roofline(grid, 0, "k_jacobi<double>")
    log("plain FP64: %.3f events/s, %.0f sweeps/event" % (plain_steps / pdt, psweeps))

    # accelerated: FP32 Chebyshev-Jacobi corrections in FP64 refinement, same start
    grid2 = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
    grid2.set_phi(phi0, phi1, sw[0], sw[1])
    rho2 = grid2.accel_setup(0.5)
    aopts = event_options(tsg)
    aopts.method, aopts.accel_tol = 1, 1e-8
    aev, adt, as0, as1 = run_events(grid2, aopts, args.warmup, args.steps, "mixed")
    roof_acc = roofline(grid2, 1, "k_cheb32")
    steps_per_event = sum(e.global_iter for e in aev) / max(1, len(aev))
    refinements = sum(e.local_iter for e in aev) / max(1, len(aev))
    pn = [e.node for e in pev]
    an = [e.node for e in aev]
    common = min(len(pn), len(an))
    first_diff = next((i for i in range(common) if pn[i] != an[i]), None)
    value = args.steps / adt
    res = {
        "metric": "evaporation events/s at residual <=1e-8 V (mixed FP32/FP64 relaxation; plain FP64 alongside)",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": adt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded BCC [110] tip, graded vacuum, own Delaunay mesh)",
        "config": {"workload": "%s: BCC [110] tip r %.0f nm shank %.0f nm, %d atoms + %d vacuum nodes, %d unknown rows"
                               % (args.workload, w["radius"] * 1e9, w["shank"] * 1e9, counts["atoms"],
                                  counts["vacuum_lattice"] + counts["vacuum_random"], ps1["rows"]),
                   "rho": rho2, "refinements_per_event": refinements, "correction_steps_per_event": steps_per_event},
        "plain_fp64": {"value": plain_steps / pdt, "unit": "events/s", "steps": plain_steps,
                       "ms_per_step": pdt / plain_steps * 1e3, "sweeps_per_event": psweeps},
        "speedup_vs_plain_fp64": value / (plain_steps / pdt),
        "sequence": {"compared": common, "first_difference": first_diff},
        "roofline": roof_acc,
        "roofline_plain_fp64": roof_plain,
    }
    if not args.no_cpu:
        tv, tn = mesh.tets()
        res["cpu_baseline"] = cpu_baseline(tsg, grid, xyz, ids, tv, tn, psweeps)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
