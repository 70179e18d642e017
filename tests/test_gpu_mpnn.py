"""The warm-start network's device kernels (bf16 MFMA forward and backward,
Adam training) against the CPU oracle of the same network, through the C ABI."""
import sys

import numpy as np
import pytest

import tapsim_modified_amd as tsg
from oracle import oracle as O
from tests.test_mpnn_oracle import random_graph, random_params

pytestmark = pytest.mark.gpu


def log(msg):
    print("[gpu-mpnn] " + msg, flush=True)
    sys.stdout.flush()


def block_graph(sizes, rng):
    """Block-diagonal union of random graphs (one batch of several shells)."""
    ptr, col, a = [0], [], []
    off = 0
    for n in sizes:
        p, c, w = random_graph(n, rng, deg=6)
        ptr += list(p[1:] + ptr[-1])
        col += list(c + off)
        a += list(w)
        off += n
    return np.array(ptr, np.int32), np.array(col, np.int32), np.array(a, np.float32)


@pytest.mark.parametrize("sizes", [[37], [300, 211, 95]])
def test_forward_backward_match_oracle(sizes):
    log("forward/backward sizes %s" % sizes)
    rng = np.random.default_rng(sum(sizes))
    ptr, col, a = block_graph(sizes, rng)
    n = len(ptr) - 1
    x = rng.standard_normal((n, O.MPNN_IN)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    mask = (rng.uniform(size=n) < 0.8).astype(np.float32)
    p = random_params(rng)
    net = tsg.WarmStartNet(max_batch_rows=4096, data_rows=1 << 16, data_edges=1 << 18)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    loss, y, g = net.eval(ptr, col, a, x, t, mask)
    lo, yo, go = O.mpnn(ptr, col, a, x, t, mask, p, emulate=True)
    log("loss gpu %.8g oracle %.8g; max|dy| %.3g of %.3g" % (loss, lo, np.max(np.abs(y - yo)), np.max(np.abs(yo))))
    # bf16 operands: the device accumulates in fp32 where the oracle uses
    # double, so an operand may land on the other side of a bf16 rounding
    # boundary now and then; the tolerances below allow for that and nothing more
    assert np.max(np.abs(y - yo)) <= 2e-4 * np.max(np.abs(yo))
    assert abs(loss - lo) <= 2e-4 * lo
    blocks = [("W_in", O.MPNN_W_IN, O.MPNN_B_IN), ("b_in", O.MPNN_B_IN, O.MPNN_LAYER0)] + \
        [("layer%d" % l, O.MPNN_LAYER0 + l * O.MPNN_LAYER_STRIDE, O.MPNN_LAYER0 + (l + 1) * O.MPNN_LAYER_STRIDE)
         for l in range(3)] + [("out", O.MPNN_W_OUT, O.MPNN_NPARAM)]
    for name, lo_, hi in blocks:
        gg, oo = g[lo_:hi].astype(np.float64), go[lo_:hi]
        rel = np.linalg.norm(gg - oo) / np.linalg.norm(oo)
        log("grad %-7s rel err %.3g" % (name, rel))
        assert rel < 1e-3, name


def test_training_reduces_loss_on_a_learnable_target():
    """Targets that the network can represent (a smoothed function of the
    features over the graph): Adam on the device drives the loss down."""
    log("training")
    rng = np.random.default_rng(5)
    net = tsg.WarmStartNet(max_batch_rows=8192, data_rows=1 << 18, data_edges=1 << 21, seed=3)
    for gidx in range(48):
        n = int(rng.integers(60, 200))
        ptr, col, a = random_graph(n, rng, deg=5)
        x = rng.standard_normal((n, O.MPNN_IN)).astype(np.float32)
        A = np.zeros((n, n))
        for i in range(n):
            A[i, col[ptr[i]:ptr[i + 1]]] = a[ptr[i]:ptr[i + 1]]
        base = x[:, 0] - 0.5 * x[:, 1]
        t = (base + A @ base + 0.5 * A @ (A @ np.maximum(x[:, 2], 0))).astype(np.float32)
        mask = np.ones(n, np.float32)
        net.add_graph(ptr, col, a, x, t, mask)
    g, r, e = net.dataset()
    assert g == 48
    before = net.evaluate(40, 48)
    first, last = net.train(0, 40, epochs=60, batch_graphs=8, lr=2e-3, seed=9)
    after = net.evaluate(40, 48)
    log("train loss %.4g -> %.4g; held-out %.4g -> %.4g" % (first, last, before, after))
    assert last < 0.2 * first
    assert after < 0.3 * before
