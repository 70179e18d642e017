"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference TAPSim sources are compiled (never copied) by
oracle/Makefile.ref into oracle/_ref/ref_driver, which runs the reference's own
Grid_3d / Surface_3d / Process code on a small specimen generated and meshed
by this repository.  Only inputs and the reference's outputs are stored.

    python tests/golden/make_golden.py      (needs /root/reference and the built library)
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import tapsim_modified_amd as tsg  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
OUT = os.path.join(ROOT, "tests", "golden")

# species of the fixture (ids as in the reference's examples/example.aly)
SPECIES = {
    0: dict(name="Vacuum", eps=1.0, rem=0, neu=0, dir=0, phi=0.0),
    1: dict(name="Top", eps=1.0, rem=0, neu=0, dir=1, phi=0.0),
    2: dict(name="Bottom", eps=1.0, rem=0, neu=1, dir=0, phi=0.0),
    3: dict(name="Sides", eps=1.0, rem=0, neu=0, dir=1, phi=0.0),
    10: dict(name="Matrix", eps=1.0e6, rem=1, neu=0, dir=1, phi=5000.0, mass=55.845, charge=2, field=3.0e10, energy=1.0),
    11: dict(name="Solute", eps=1.0e6, rem=1, neu=0, dir=1, phi=5000.0, mass=58.69, charge=2, field=2.5e10, energy=1.0),
}


# temperature of the thermal start velocity fixture, K
THERMAL_T = 60.0
# charge density of the vacuum in the space-charge fixture, C/m^3
RHO_VACUUM = 2.0e6


def coordination_species():
    """SPECIES plus ids 12..127 as matrix atoms whose evaporation field grows
    with the id (the ids that the fork's update_coordination produces)"""
    sp = {k: dict(v) for k, v in SPECIES.items()}
    for k in range(12, 128):
        sp[k] = dict(name="Coord%d" % k, eps=1.0e6, rem=1, neu=0, dir=1, phi=5000.0, mass=55.845, charge=2,
                     field=3.0e10 * (1.0 + 0.002 * (k - 10)), energy=1.0)
    return sp


def aly_text(species, temperature=0.0):
    out = ["ASCII", "", "TEMPERATURE = %.5e" % temperature]
    for sid in sorted(species):
        s = species[sid]
        out += ["", "ID = %d" % sid, "NAME = %s" % s["name"], "CHARGE_DENSITY = %.8e" % s.get("rho", 0.0),
                "DIELECTRICITY = %.8e" % s["eps"], "REMOVABLE = %d" % s["rem"], "NEUMANN_BOUNDARY = %d" % s["neu"],
                "DIRICHLET_BOUNDARY = %d" % s["dir"], "POTENTIAL = %.8e" % s["phi"],
                "MASS = %.8e" % s.get("mass", 0.0), "EVAPORATION_CHARGE_STATE = %d" % s.get("charge", 0),
                "EVAPORATION_FIELD_STRENGTH = %.8e" % s.get("field", 0.0),
                "EVAPORATION_ACTIVATION_ENERGY = %.8e" % s.get("energy", 0.0)]
    return "\n".join(out) + "\n"


def write_input(path, xyz, ids, tets=None):
    with open(path, "wb") as f:
        f.write(b"TSGM")
        f.write(struct.pack("<iii", 1, len(xyz), 0 if tets is None else len(tets)))
        f.write(np.ascontiguousarray(xyz, "<f4").tobytes())
        f.write(np.ascontiguousarray(ids, "<i2").tobytes())
        if tets is not None:
            f.write(np.ascontiguousarray(tets, "<i4").tobytes())


def run(workdir, steps, tets=True):
    r = subprocess.run([DRIVER, os.path.join(workdir, "in.bin"), os.path.join(workdir, "cfg.aly"), workdir] + steps,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("ref_driver failed: %s\n%s" % (steps, r.stderr[-2000:]))
    return r


def rd(workdir, name, dtype):
    return np.fromfile(os.path.join(workdir, name), dtype=dtype)


def specimen(radius, shank, box, vac, spacing, seed):
    p = tsg.SpecimenParams()
    p.a = 0.287e-9
    p.radius, p.shank = radius, shank
    for k in range(3):
        p.box[k] = box[k]
    p.shell, p.vacuum_target, p.decay, p.face_spacing = 2.0, vac, 3.0, spacing
    p.cluster_density, p.cluster_radius, p.seed = 0.0, 2e-9, seed
    p.id_vacuum, p.id_top, p.id_bottom, p.id_sides, p.id_matrix, p.id_solute = 0, 1, 2, 3, 10, 11
    return tsg.generate_specimen(p)


def make_small():
    xyz, ids, counts = specimen(1.2e-9, 1.5e-9, (8e-9, 8e-9, 11e-9), 1500, 2.5e-9, 11)
    # a few solute atoms so that two evaporation fields take part
    rng = np.random.default_rng(5)
    atoms = np.where(ids == 10)[0]
    ids[rng.choice(atoms, size=len(atoms) // 10, replace=False)] = 11
    mesh = tsg.Mesh(xyz)
    tv, tn = mesh.tets()
    fx = dict(xyz=xyz, ids=ids, tets=tv, neigh=tn)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        # A) topology, couplings, a few Jacobi cycles, fields
        run(wd, ["mesh", "sync", "dumpmesh", "dumpgrid", "vorarea:one", "dumpphi:init", "relax:6:1", "dumpphi:r6",
                 "field:r6"])
        fx["ref_neigh"] = rd(wd, "neigh.i32", np.int32).reshape(-1, 4)
        fx["ref_assoc"] = rd(wd, "assoc.i32", np.int32)
        fx["ref_boundary"] = rd(wd, "boundary.u8", np.uint8)
        fx["ref_ptr"] = rd(wd, "nbr_ptr.i32", np.int32)
        fx["ref_idx"] = rd(wd, "nbr_idx.i32", np.int32)
        fx["ref_coupling"] = rd(wd, "coupling.f32", np.float32)
        fx["ref_flags"] = rd(wd, "flags.u8", np.uint8)
        fx["ref_area_one"] = rd(wd, "vorarea_one.f32", np.float32)
        fx["ref_dev_r6"] = rd(wd, "relax_dev.f64", np.float64)
        fx["ref_phi0_r6"] = rd(wd, "phi0_r6.f32", np.float32)
        fx["ref_phi1_r6"] = rd(wd, "phi1_r6.f32", np.float32)
        fx["ref_field_r6"] = rd(wd, "field_o1_r6.f32", np.float32).reshape(-1, 3)
        fx["ref_field2_r6"] = rd(wd, "field_o2_r6.f32", np.float32).reshape(-1, 3)
        fx["ref_charge2_r6"] = rd(wd, "charge_o2_r6.f32", np.float32)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        # B) threshold relaxation with a deviation queue, and in-place cycles
        run(wd, ["mesh", "sync", "relaxthr:0.5:1:4", "dumpphi:thr", "relax:3:4", "dumpphi:c4"])
        fx["ref_thr_iter"] = rd(wd, "relaxthr_iter.i64", np.int64)
        fx["ref_phi0_thr"] = rd(wd, "phi0_thr.f32", np.float32)
        fx["ref_phi1_thr"] = rd(wd, "phi1_thr.f32", np.float32)
        fx["ref_dev_c4"] = rd(wd, "relax_dev.f64", np.float64)
        fx["ref_phi0_c4"] = rd(wd, "phi0_c4.f32", np.float32)
        fx["ref_phi1_c4"] = rd(wd, "phi1_c4.f32", np.float32)
    # the atom with the largest z among the removable atoms (the apex)
    rem = np.where(np.isin(ids, [10, 11]))[0]
    node = int(rem[np.argmax(xyz[rem, 2])])
    fx["local_node"] = np.int32(node)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        # C) an atom becomes vacuum (fastResync), then localRelax on its shell
        run(wd, ["mesh", "sync", "relax:40:1", "evapnode:%d" % node, "dumpgrid", "local:%d:3:1e-2:5:5" % node,
                 "dumpphi:loc"])
        fx["ref_coupling_resync"] = rd(wd, "coupling.f32", np.float32)
        fx["ref_local_iter"] = rd(wd, "local_iter.i64", np.int64)
        fx["ref_phi0_loc"] = rd(wd, "phi0_loc.f32", np.float32)
        fx["ref_phi1_loc"] = rd(wd, "phi1_loc.f32", np.float32)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        # D) Process::evaporation: 25 events with local relaxation, one global
        #    cycle and a refresh relaxation every 5 events; Cash-Karp flights
        run(wd, ["mesh", "sync", "relax:300:1",
                 "evap:25:0:0:1e-2:3:5:5:1:5:1e-3:1:0:1:1", "dumpphi:evap"])
        ints = rd(wd, "evap_int.i32", np.int32).reshape(-1, 4)
        fx["ref_evap_node"], fx["ref_evap_id"] = ints[:, 0], ints[:, 1]
        fx["ref_evap_trajlen"], fx["ref_evap_surface"] = ints[:, 2], ints[:, 3]
        fx["ref_evap_iter"] = rd(wd, "evap_iter.i64", np.int64).reshape(-1, 2)
        fx["ref_evap_f"] = rd(wd, "evap_f.f32", np.float32).reshape(-1, 22)
        fx["ref_evap_normal"] = rd(wd, "evap_normal.f32", np.float32).reshape(-1, 3)
        fx["ref_evap_next_rand"] = rd(wd, "evap_next_rand.i32", np.int32)
        fx["ref_phi0_evap"] = rd(wd, "phi0_evap.f32", np.float32)
        fx["ref_phi1_evap"] = rd(wd, "phi1_evap.f32", np.float32)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES, temperature=THERMAL_T))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        # D2) the same with the Monte Carlo choice and thermal start velocities
        #     at THERMAL_T: every draw of rand() (choice, velocity, the walks'
        #     first faces) in the reference's order
        run(wd, ["mesh", "sync", "relax:300:1",
                 "evap:25:0:1:1e-2:3:5:5:1:5:1e-3:1:0:1:1:1"])
        ints = rd(wd, "evap_int.i32", np.int32).reshape(-1, 4)
        fx["ref_mc_node"], fx["ref_mc_trajlen"] = ints[:, 0], ints[:, 2]
        fx["ref_mc_f"] = rd(wd, "evap_f.f32", np.float32).reshape(-1, 22)
        fx["ref_mc_next_rand"] = rd(wd, "evap_next_rand.i32", np.int32)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        # E) one trajectory from the apex atom after relaxation (O5 + error-restricted)
        run(wd, ["mesh", "sync", "relax:300:1", "evapnode:%d" % node, "relax:1:1", "traj:t:%d:1:1" % node])
        fx["ref_traj"] = rd(wd, "traj_t.f32", np.float32).reshape(-1, 8)
        fx["ref_traj_status"] = rd(wd, "traj_status_t.i32", np.int32)
    # G) space charge: the vacuum species with a charge density (computeCharge's
    #    Voronoi cell volumes in fastSync and fastResync), then Jacobi cycles
    sq = {k: dict(v) for k, v in SPECIES.items()}
    sq[0]["rho"] = RHO_VACUUM
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(sq))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        run(wd, ["mesh", "sync", "dumpgrid", "relax:6:1", "dumpphi:q6"])
        fx["ref_charge_q"] = rd(wd, "charge.f32", np.float32)
        fx["ref_phi0_q6"] = rd(wd, "phi0_q6.f32", np.float32)
        fx["ref_phi1_q6"] = rd(wd, "phi1_q6.f32", np.float32)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(sq))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        run(wd, ["mesh", "sync", "relax:40:1", "evapnode:%d" % node, "dumpgrid"])
        fx["ref_charge_q_resync"] = rd(wd, "charge.f32", np.float32)
    # H) the fork's update_coordination applied once and twelve times (ids of every
    #    atom grow by their atom-neighbour counts and wrap as a signed char)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        run(wd, ["mesh", "sync", "coord", "dumpgrid"])
        fx["ref_coord_ids1"] = rd(wd, "ids.i16", np.int16)
        run(wd, ["mesh", "sync"] + ["coord"] * 12 + ["dumpgrid"])
        fx["ref_coord_ids12"] = rd(wd, "ids.i16", np.int16)
    # H2) Process::evaporation with the fork's update_coordination (6 events),
    #     every id from 10 to 127 defined, the evaporation field growing with
    #     the id, so that the coordination ids steer the sequence
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(coordination_species()))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        run(wd, ["mesh", "sync", "relax:300:1", "evap:6:0:0:1e-2:3:5:5:1:5:1e-3:1:0:1:1:0:1", "dumpgrid"])
        ints = rd(wd, "evap_int.i32", np.int32).reshape(-1, 4)
        fx["ref_coord_evap_node"], fx["ref_coord_evap_id"] = ints[:, 0], ints[:, 1]
        fx["ref_coord_evap_f"] = rd(wd, "evap_f.f32", np.float32).reshape(-1, 22)
        fx["ref_coord_evap_ids"] = rd(wd, "ids.i16", np.int16)
    # F) field lines (Fieldline_3d, debug.cpp's surface field lines) from the
    #    centroids of the highest vacuum-side tets of the six highest atoms:
    #    to the mesh's end, and limited to 2 nm
    top = rem[np.argsort(xyz[rem, 2])[-6:]]
    starts = []
    for a in top:
        inc = np.where((tv == a).any(1))[0]
        cz = xyz[tv[inc]].astype(np.float32).mean(1)
        starts.append(cz[np.argmax(cz[:, 2])])
    starts = np.array(starts, np.float32)
    fx["fieldline_start"] = starts
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)
        starts.tofile(os.path.join(wd, "a_start.f32"))
        starts.tofile(os.path.join(wd, "b_start.f32"))
        run(wd, ["mesh", "sync", "relax:300:1", "fieldlines:a:-1:5000:1000", "fieldlines:b:2e-9:5000:1000"])
        for t in "ab":
            fx["ref_fieldline_%s_info" % t] = rd(wd, "fieldlines_%s.i32" % t, np.int32).reshape(-1, 2)
            fx["ref_fieldline_%s" % t] = rd(wd, "fieldlines_%s.f32" % t, np.float32).reshape(-1, 7)
            fx["ref_fieldline_%s_next_rand" % t] = rd(wd, "fieldlines_next_rand_%s.i32" % t, np.int32)
    np.savez_compressed(os.path.join(OUT, "tip_small.npz"), **fx)
    print("tip_small: nodes %d tets %d counts %s" % (len(xyz), len(tv), counts))


def make_tetgen_mesh():
    """Delaunay fixture: TetGen's tetrahedralisation of random points in a
    unit-ish box (general position) -- my mesher must reproduce it exactly."""
    rng = np.random.default_rng(3)
    pts = (rng.random((600, 3)) * 4e-9).astype(np.float32)
    # hull corners jittered off the cube so that no five points are cospherical
    corners = np.array([[x, y, z] for x in (-1e-9, 5e-9) for y in (-1e-9, 5e-9) for z in (-1e-9, 5e-9)])
    corners = (corners + (rng.random(corners.shape) - 0.5) * 0.4e-9).astype(np.float32)
    xyz = np.vstack([corners, pts])
    ids = np.zeros(len(xyz), np.int16)
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids)
        run(wd, ["mesh", "dumpmesh"])
        tets = rd(wd, "tets.i32", np.int32).reshape(-1, 4)
    np.savez_compressed(os.path.join(OUT, "tetgen_random.npz"), xyz=xyz, tets=tets)
    print("tetgen_random: %d tets" % len(tets))


def make_formats():
    """File formats: the reference's own File_Io writes a node file (ASCII
    with numbers and potentials, BINARY with numbers) and rewrites two
    configurations (its example.aly and the fixture's); it reads node files
    written here (ASCII with comment lines and unnumbered nodes, BINARY with
    potentials) and reports what it read."""
    from tapsim_modified_amd import formats as F
    rng = np.random.default_rng(11)
    n = 300
    xyz = ((rng.random((n, 3)) - 0.5) * 8e-9).astype(np.float32)
    ids = rng.choice(np.array([0, 1, 3, 10, 11], np.int16), n)
    out = {}
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), xyz, ids)
        # the reference writes
        run(wd, ["writegeom:w_ascii.txt:0:1:1", "writegeom:w_bin.bin:1:1:0", "writeconfig:w_cfg.aly"])
        for k in ("w_ascii.txt", "w_bin.bin", "w_cfg.aly"):
            out[k.replace(".", "_")] = np.frombuffer(open(os.path.join(wd, k), "rb").read(), np.uint8)
        # the reference reads files written here
        numbers = (np.arange(n, dtype=np.uint32) * 3 + 7)
        numbers[rng.choice(n, 40, replace=False)] = F.INVALID_NUMBER
        phi = (rng.random(n) * 5000.0).astype(np.float32)
        F.write_node_file(os.path.join(wd, "r_ascii.txt"), xyz, ids, numbers, phi, "ASCII")
        body = open(os.path.join(wd, "r_ascii.txt"), "rb").read()
        body = b"# node file written by the fixture generator\n#\n" + body
        open(os.path.join(wd, "r_ascii.txt"), "wb").write(body)
        F.write_node_file(os.path.join(wd, "r_bin.bin"), xyz, ids, None, phi, "BINARY")
        run(wd, ["readgeom:r_ascii.txt:ra", "readgeom:r_bin.bin:rb"])
        for k in ("r_ascii.txt", "r_bin.bin"):
            out[k.replace(".", "_")] = np.frombuffer(open(os.path.join(wd, k), "rb").read(), np.uint8)
        for tag in ("ra", "rb"):
            out[tag + "_xyz"] = rd(wd, tag + "_xyz.f32", np.float32).reshape(-1, 3)
            out[tag + "_ids"] = rd(wd, tag + "_ids.i16", np.int16)
            out[tag + "_numbers"] = rd(wd, tag + "_numbers.u32", np.uint32)
            out[tag + "_phi"] = rd(wd, tag + "_phi.f32", np.float32)
            out[tag + "_flags"] = rd(wd, tag + "_flags.i32", np.int32)
        # the reference's own tapsim.ini (data)
        out["tapsim_ini"] = np.frombuffer(open("/root/reference/tapsim.ini", "rb").read(), np.uint8)
        # the reference's example configuration, as read and rewritten by it
        example = open("/root/reference/examples/example.aly", "rb").read()
        out["example_aly"] = np.frombuffer(example, np.uint8)
        open(os.path.join(wd, "cfg.aly"), "wb").write(example)
        write_input(os.path.join(wd, "in.bin"), xyz, np.zeros(n, np.int16))  # ids the example defines
        run(wd, ["writeconfig:x_cfg.aly"])
        out["x_cfg_aly"] = np.frombuffer(open(os.path.join(wd, "x_cfg.aly"), "rb").read(), np.uint8)
        # the reference's results stream of 12 synthetic events (27 floats each)
        K = 12
        rec = np.zeros((K, 27), np.float32)
        rec[:, 0] = np.arange(1, K + 1)
        rec[:, 1] = rng.choice([10, 11], K)
        rec[:, 2] = rng.integers(1, 100000, K)
        rec[:, 3:] = (rng.standard_normal((K, 24)) * 10.0 ** rng.integers(-12, 11, (K, 24))).astype(np.float32)
        rec[:, 10] = np.abs(rec[:, 10])
        rec.tofile(os.path.join(wd, "res_in.f32"))
        run(wd, ["writeresults:res_in.f32:res_ascii.txt:0", "writeresults:res_in.f32:res_bin.bin:1"])
        out["res_in"] = rec
        for k in ("res_ascii.txt", "res_bin.bin"):
            out[k.replace(".", "_")] = np.frombuffer(open(os.path.join(wd, k), "rb").read(), np.uint8)
        # the reference's trajectory stream (File_Io::LogTrajectory) of 10
        # synthetic flights, one file in each mode and ASCII in chunks of 4
        trng = np.random.default_rng(23)
        flat = []
        for e in range(10):
            P = int(trng.integers(1, 7))
            head = [e + 1, int(trng.integers(1, 100000)), float(trng.random() * 1e-3 + 1e-6),
                    [5, 1, 3, -1, 0, 5, 2, 4, 5, 6][e], 3.2043547e-19, 9.27e-26] + [P]
            head += list((trng.standard_normal(6) * 10.0 ** trng.integers(-14, -3, 6)))
            ph = (trng.standard_normal((P, 8)) * 10.0 ** trng.integers(-12, 5, (P, 8))).astype(np.float32)
            ph[:, 0] = np.abs(ph[:, 0])
            ph[:, 7] = trng.integers(-1, 2000000, P)
            flat += head + list(ph.ravel())
        traj = np.asarray(flat, np.float32)
        traj.tofile(os.path.join(wd, "traj_in.f32"))
        run(wd, ["writetraj:traj_in.f32:traj_ascii.txt:0:0", "writetraj:traj_in.f32:traj_bin.bin:1:0",
                 "writetraj:traj_in.f32:traj_chunk.txt:0:4"])
        out["traj_in"] = traj
        for k in ("traj_ascii.txt", "traj_bin.bin", "traj_chunk.txt.00000001", "traj_chunk.txt.00000004",
                  "traj_chunk.txt.00000008"):
            out[k.replace(".", "_")] = np.frombuffer(open(os.path.join(wd, k), "rb").read(), np.uint8)
    out["xyz"], out["ids"] = xyz, ids
    np.savez_compressed(os.path.join(OUT, "formats.npz"), **out)
    print("formats: %d nodes, %d bytes of reference output" % (n, sum(v.nbytes for k, v in out.items() if k.startswith(("w_", "x_")))))


# temperature of the Boltzmann-model fixture, K
BOLTZMANN_T = 20000.0


def make_pinning():
    """Fixtures for the parts of the path that the first fixture set leaves
    open, on the same small specimen (inputs from tip_small.npz): relax with
    cycleSize > 1 inside a threshold relaxation and relax(100), the
    RelaxationFault of a threshold relaxation that stagnates in float, the
    randomReset preset, the probability models 1-3 in the event loop, and
    trajectories with the Euler integrator and the none / geometry-restricted
    / adaptive steppers."""
    base = dict(np.load(os.path.join(OUT, "tip_small.npz")))
    xyz, ids, tv = base["xyz"], base["ids"], base["tets"]
    node = int(base["local_node"])
    fx = {}

    def fresh(wd, species=SPECIES, temperature=0.0):
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(species, temperature=temperature))
        write_input(os.path.join(wd, "in.bin"), xyz, ids, tv)

    with tempfile.TemporaryDirectory() as wd:
        fresh(wd)
        # relax(threshold, 4, 5) (odd cycles Jacobi, even cycles in place),
        # then relax(100) from there
        run(wd, ["mesh", "sync", "relaxthr:0.5:4:5", "dumpphi:t4", "relax:1:100", "dumpphi:c100"])
        fx["ref_thr4_iter"] = rd(wd, "relaxthr_iter.i64", np.int64)
        fx["ref_phi0_t4"] = rd(wd, "phi0_t4.f32", np.float32)
        fx["ref_phi1_t4"] = rd(wd, "phi1_t4.f32", np.float32)
        fx["ref_pot_t4"] = rd(wd, "potential_t4.f32", np.float32)
        fx["ref_dev_c100"] = rd(wd, "relax_dev.f64", np.float64)
        fx["ref_phi0_c100"] = rd(wd, "phi0_c100.f32", np.float32)
        fx["ref_phi1_c100"] = rd(wd, "phi1_c100.f32", np.float32)
    with tempfile.TemporaryDirectory() as wd:
        fresh(wd)
        # from a float fixed point (deviation exactly 0) a negative threshold
        # never stops the loop: a queue of 25 equal deviations has a zero
        # slope, which throws RelaxationFault; then the same with cycleSize 4
        run(wd, ["mesh", "sync", "relax:2000:1", "relaxthr:-1:1:25", "relaxthr:-2:4:5", "dumpphi:fault"])
        fx["ref_fault_iter"] = rd(wd, "relaxthr_iter.i64", np.int64)
        fx["ref_fault"] = rd(wd, "relaxthr_fault_-1.f64", np.float64)
        fx["ref_fault4"] = rd(wd, "relaxthr_fault_-2.f64", np.float64)
        fx["ref_phi0_fault"] = rd(wd, "phi0_fault.f32", np.float32)
        fx["ref_phi1_fault"] = rd(wd, "phi1_fault.f32", np.float32)
    with tempfile.TemporaryDirectory() as wd:
        fresh(wd)
        run(wd, ["mesh", "sync", "randreset:7", "dumpphi:rr"])
        fx["ref_phi0_rr"] = rd(wd, "phi0_rr.f32", np.float32)
        fx["ref_phi1_rr"] = rd(wd, "phi1_rr.f32", np.float32)
    for prob in (1, 2, 3):
        with tempfile.TemporaryDirectory() as wd:
            fresh(wd, temperature=BOLTZMANN_T if prob == 1 else 0.0)
            run(wd, ["mesh", "sync", "relax:300:1", "evap:10:%d:0:1e-2:3:5:5:1:5:1e-3:1:0:1:1" % prob])
            ints = rd(wd, "evap_int.i32", np.int32).reshape(-1, 4)
            fx["ref_p%d_node" % prob] = ints[:, 0]
            fx["ref_p%d_f" % prob] = rd(wd, "evap_f.f32", np.float32).reshape(-1, 22)
            fx["ref_p%d_iter" % prob] = rd(wd, "evap_iter.i64", np.int64).reshape(-1, 2)
    with tempfile.TemporaryDirectory() as wd:
        fresh(wd)
        run(wd, ["mesh", "sync", "relax:300:1", "evapnode:%d" % node, "relax:1:1", "traj:o1:%d:0:1" % node,
                 "traj:s0:%d:1:0" % node, "traj:s2:%d:1:2" % node, "traj:s3:%d:1:3" % node,
                 "traj:o5:%d:1:1" % node])
        for tag in ("o1", "s0", "s2", "s3", "o5"):
            fx["ref_traj_" + tag] = rd(wd, "traj_%s.f32" % tag, np.float32).reshape(-1, 8)
            fx["ref_traj_status_" + tag] = rd(wd, "traj_status_%s.i32" % tag, np.int32)
            # Trajectory_3d::error_estimate after the extrapolation: position xyz, velocity xyz
            fx["ref_traj_err_" + tag] = rd(wd, "traj_err_%s.f32" % tag, np.float32)
    np.savez_compressed(os.path.join(OUT, "pinning.npz"), **fx)
    print("pinning: %s" % ", ".join("%s%s" % (k, list(v.shape)) for k, v in sorted(fx.items())))


def make_dump():
    """The reference's dump (File_Io::writeSystem) of the small specimen after
    six Jacobi cycles, with last index 42 (tests/golden/dump.npz)"""
    base = dict(np.load(os.path.join(OUT, "tip_small.npz")))
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "cfg.aly"), "w").write(aly_text(SPECIES))
        write_input(os.path.join(wd, "in.bin"), base["xyz"], base["ids"], base["tets"])
        run(wd, ["mesh", "sync", "relax:6:1", "writedump:dump.bin:42"])
        fx = dict(dump=np.frombuffer(open(os.path.join(wd, "dump.bin"), "rb").read(), np.uint8),
                  cfg_aly=np.frombuffer(open(os.path.join(wd, "cfg.aly"), "rb").read(), np.uint8))
    np.savez_compressed(os.path.join(OUT, "dump.npz"), **fx)
    print("dump: %d bytes" % len(fx["dump"]))


if __name__ == "__main__":
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-f", "oracle/Makefile.ref"], cwd=ROOT, check=True)
    if len(sys.argv) > 1 and sys.argv[1] == "formats":
        make_formats()
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "pinning":
        make_pinning()
        sys.exit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "dump":
        make_dump()
        sys.exit(0)
    make_tetgen_mesh()
    make_small()
    make_formats()
    make_pinning()
