"""CPU-baseline calibration (test infrastructure, this container only): the
reference TAPSim built from its own sources (oracle/Makefile.ref ->
oracle/_ref/ref_driver) against the oracle's float-faithful restatement, a
fixed number of Jacobi sweeps (relax(1) calls) on the same mesh with the same
number of threads, both threaded as the reference's cwrapper_threadedRelax
(tapsim/grid_3d.cpp:1458-1514).  The potentials after the sweeps are
compared bit for bit, so both sides did the same arithmetic.

    python oracle/calibrate_cpu.py [--workload c1] [--sweeps 400] [--threads 1,8] [--out result.json]

The specimen and mesh come from the product's generator and mesher (host
code); the reference adopts the mesh through set_metaInformation.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# run as a script, its own folder would shadow the oracle package
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != os.path.dirname(os.path.abspath(__file__))]
sys.path.insert(0, ROOT)

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--sweeps", type=int, default=400)
    ap.add_argument("--threads", default="1,8")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-j8", "-f", "oracle/Makefile.ref"], cwd=ROOT, check=True)
    import bench
    import tapsim_modified_amd as tsg
    from oracle import oracle as O
    from tests.golden.make_golden import aly_text, write_input

    w = bench.WORKLOADS[args.workload]
    xyz, ids, counts = bench.build_specimen(tsg, w)
    mesh = tsg.Mesh(xyz)
    tv, tn = mesh.tets()
    names = {0: "Vacuum", 1: "Top", 2: "Bottom", 3: "Sides", 10: "Matrix", 11: "Solute"}
    species = {}
    for sid, s in bench.SPECIES.items():
        species[sid] = dict(name=names[sid], eps=s.get("epsilon", 1.0), rem=s.get("removable", 0),
                            neu=s.get("neumann", 0), dir=s.get("dirichlet", 0), phi=s.get("phi", 0.0),
                            mass=s.get("mass", 0.0), charge=s.get("evap_charge", 0), field=s.get("evap_field", 0.0),
                            energy=s.get("evap_energy", 0.0))
    rows = []
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(species))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        for T in [int(t) for t in args.threads.split(",")]:
            t0 = time.perf_counter()
            r = subprocess.run([DRIVER, os.path.join(wd, "in.bin"), os.path.join(wd, "cfg.aly"), wd, "mesh", "sync",
                                "threads:%d" % T, "relax:%d:1" % args.sweeps, "dumpphi:cal"],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stderr[-2000:])
            steps = np.fromfile(os.path.join(wd, "step_seconds.f64"), np.float64)
            ref_phi0 = np.fromfile(os.path.join(wd, "phi0_cal.f32"), np.float32)
            ref_phi1 = np.fromfile(os.path.join(wd, "phi1_cal.f32"), np.float32)
            o = O.Oracle(O.F32, xyz, ids, tv, tn, O.species_table(bench.SPECIES))
            o.sync()
            _, sec = o.relax_threaded(args.sweeps, T)
            same = (np.array_equal(o.phi(0).astype(np.float32).view(np.uint32), ref_phi0.view(np.uint32)) and
                    np.array_equal(o.phi(1).astype(np.float32).view(np.uint32), ref_phi1.view(np.uint32)))
            row = dict(threads=T, sweeps=args.sweeps, reference_ms_per_sweep=steps[3] / args.sweeps * 1e3,
                       oracle_ms_per_sweep=sec / args.sweeps * 1e3, ratio_reference_over_oracle=steps[3] / sec,
                       potentials_bit_exact=bool(same), reference_fastsync_s=float(steps[1]),
                       wall_s=time.perf_counter() - t0)
            print(json.dumps(row), flush=True)
            rows.append(row)
    res = dict(workload=args.workload, nodes=int(len(xyz)), tets=int(len(tv)), cpu=cpu_model(),
               host_cores=len(os.sched_getaffinity(0)), rows=rows)
    print(json.dumps(res))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
