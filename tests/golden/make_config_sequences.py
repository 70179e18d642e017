"""Writes tests/golden/<workload>_sequence.npz: the evaporation sequence of the
plain FP64 path (the reference's relaxation, bit-exact to the oracle on the
GPU parity tests) on a BASELINE workload, with the mixed and learned+mixed
sequences from the same relaxed state beside it.  Needs the GPU (the plain
path runs on it; its first events are checked against the CPU oracle by
tests/test_gpu_configs.py).

    python tests/golden/make_config_sequences.py c1 1000 [OUT.npz]

Events of the plain path: local 3-shell relaxation, one global cycle and a
refresh relaxation to a deviation of 1e-8 V every event (bench.py's plain
options), Cash-Karp flights.  The mixed and learned paths relax to a scaled
FP64 residual of 1e-8 V (multigrid correction solve); the network records
the first 200 events, trains, and predicts the rest.
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import tapsim_modified_amd as tsg  # noqa: E402

RECORD = 200


def log(msg):
    print("[sequences] " + msg, flush=True)


def relaxed_state(tsg, workload):
    """the specimen, its mesh and the relaxed start state shared by every path"""
    w = bench.WORKLOADS[workload]
    xyz, ids, counts = bench.build_specimen(tsg, w)
    mesh = tsg.Mesh(xyz)
    species = tsg.species_table(bench.SPECIES)
    g = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
    g.accel_setup(0.9)
    g.accel_relax(1e-8)
    return xyz, ids, counts, mesh, species, (g.phi(0), g.phi(1)) + g.switches()


def options(method):
    e = bench.event_options(tsg)
    if method:
        e.method, e.accel_tol, e.accel_solver = method, 1e-8, 1
    return e


def run_chunks(g, opts, n, tag):
    out = []
    t = time.time()
    while len(out) < n:
        out += list(g.evaporate(opts, min(100, n - len(out))))
        log("%s: %d events (%.1fs)" % (tag, len(out), time.time() - t))
    return out


def learned_run(mesh, ids, species, state0, n, record=RECORD):
    """n events of the learned path: the first `record` with the mixed
    method while the network records them, then training, then the rest
    from the network's warm start"""
    g = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
    g.set_phi(*state0)
    g.accel_setup(0.9)
    net = tsg.WarmStartNet(max_batch_rows=65536, data_rows=4000 * (record + 8), data_edges=60000 * (record + 8), seed=1)
    g.attach(net, record=True, predict=False)
    ev = run_chunks(g, options(1), min(record, n), "learned (recording)")
    graphs, _, _ = net.dataset()
    net.train(0, graphs - max(1, graphs // 10), epochs=60, batch_graphs=16, lr=2e-3, seed=5)
    g.attach(net, record=False, predict=True)
    if n > record:
        ev += run_chunks(g, options(2), n - record, "learned")
    return g, ev


def main():
    workload = sys.argv[1] if len(sys.argv) > 1 else "c1"
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 1000
    xyz, ids, counts, mesh, species, state0 = relaxed_state(tsg, workload)
    log("%s: %d nodes %s, %d tets" % (workload, len(xyz), counts, mesh.m))
    mixed = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
    mixed.set_phi(*state0)
    mixed.accel_setup(0.9)
    tsg.srand(99)
    me = run_chunks(mixed, options(1), n, "mixed")
    tsg.srand(99)
    _, le = learned_run(mesh, ids, species, state0, n)
    plain = tsg.Grid3D(mesh, ids, species, 0.0, tsg.PLAIN_F64)
    plain.set_phi(*state0)
    tsg.srand(99)
    pe = run_chunks(plain, options(0), n, "plain")
    digest = hashlib.sha256(plain.phi(0).tobytes() + plain.phi(1).tobytes()).hexdigest()
    out = dict(
        plain_node=np.array([e.node for e in pe], np.int32),
        plain_iter=np.array([[e.local_iter, e.global_iter] for e in pe], np.int64),
        plain_pot=np.array([[e.pot_before, e.pot_after] for e in pe], np.float64),
        plain_field=np.array([list(e.field_before) + list(e.field_after) for e in pe], np.float64),
        plain_prob=np.array([e.probability for e in pe], np.float32),
        plain_near_tie=np.array([e.near_tie for e in pe], np.int32),
        plain_phi_sha256=np.frombuffer(bytes.fromhex(digest), np.uint8),
        mixed_node=np.array([e.node for e in me], np.int32),
        learned_node=np.array([e.node for e in le], np.int32),
        learned_kept=np.array([e.warm_start == 1 for e in le], np.int8),
        atoms=np.int64(counts["atoms"]), solute=np.int64(counts["solute"]), nodes=np.int64(len(xyz)))
    path = sys.argv[3] if len(sys.argv) > 3 else os.path.join(ROOT, "tests", "golden", "%s_sequence.npz" % workload)
    np.savez_compressed(path, **out)
    same_m = np.array_equal(out["plain_node"], out["mixed_node"])
    same_l = np.array_equal(out["plain_node"], out["learned_node"])
    log("wrote %s: %d events; mixed equal %s, learned equal %s" % (path, n, same_m, same_l))


if __name__ == "__main__":
    main()
