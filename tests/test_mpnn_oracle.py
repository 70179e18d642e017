"""The CPU oracle of the warm-start network, pinned by mathematics: its
gradient against central differences of its own forward pass (in double),
its forward pass against closed forms of special weight settings, and its
bfloat16 rounding against torch's."""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def random_graph(n, rng, deg=4):
    """A symmetric random graph with positive symmetric edge weights (CSR)."""
    edges = set()
    for i in range(n):
        for j in rng.choice(n, size=deg, replace=False):
            if i != j:
                edges.add((min(i, j), max(i, j)))
    w = {e: float(rng.uniform(0.05, 0.4)) for e in edges}
    rows = [[] for _ in range(n)]
    for (i, j), v in w.items():
        rows[i].append((j, v))
        rows[j].append((i, v))
    ptr, col, a = [0], [], []
    for r in rows:
        r.sort()
        col += [c for c, _ in r]
        a += [v for _, v in r]
        ptr.append(len(col))
    return np.array(ptr, np.int32), np.array(col, np.int32), np.array(a, np.float32)


def random_params(rng, scale=0.2):
    p = (rng.standard_normal(O.MPNN_NPARAM) * scale).astype(np.float32)
    p[O.MPNN_B_IN:O.MPNN_B_IN + O.MPNN_HID] = rng.uniform(0.05, 0.3, O.MPNN_HID)
    return p


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 5000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -1e-40], np.float32)])
    ref = torch.from_numpy(v).to(torch.bfloat16).float().numpy()
    assert np.array_equal(O.bf16(v).view(np.uint32), ref.view(np.uint32))
    lib = O.load()
    assert all(np.float32(lib.ora_bf16(float(x))).view(np.uint32) == r.view(np.uint32) for x, r in zip(v[:500], ref[:500]))


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(7)
    n = 14
    ptr, col, a = random_graph(n, rng)
    x = rng.standard_normal((n, O.MPNN_IN)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    mask = (rng.uniform(size=n) < 0.7).astype(np.float32)
    p = random_params(rng)
    _, _, g = O.mpnn(ptr, col, a, x, t, mask, p, emulate=False)
    p64 = p.astype(np.float64)

    def loss_at(q):
        # parameters travel as float32: evaluate exactly the rounded point
        return O.mpnn(ptr, col, a, x, t, mask, q.astype(np.float32), emulate=False, grad=False)[0]

    # central differences along random directions, over the whole layout and
    # block by block; the step is small enough not to cross a ReLU kink
    blocks = [(0, O.MPNN_NPARAM), (O.MPNN_W_IN, O.MPNN_B_IN), (O.MPNN_B_IN, O.MPNN_LAYER0),
              (O.MPNN_W_OUT, O.MPNN_NPARAM)] + \
        [(O.MPNN_LAYER0 + l * O.MPNN_LAYER_STRIDE, O.MPNN_LAYER0 + (l + 1) * O.MPNN_LAYER_STRIDE) for l in range(3)]
    for lo, hi in blocks:
        v = np.zeros(O.MPNN_NPARAM)
        v[lo:hi] = rng.standard_normal(hi - lo)
        h = 1e-6
        qp, qm = (p64 + h * v).astype(np.float32), (p64 - h * v).astype(np.float32)
        fd = loss_at(qp) - loss_at(qm)
        an = g @ (qp.astype(np.float64) - qm.astype(np.float64))
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-12, (lo, hi, fd, an)
    # every block of the layout carries gradient
    for lo, hi in [(O.MPNN_W_IN, O.MPNN_B_IN), (O.MPNN_B_IN, O.MPNN_LAYER0), (O.MPNN_W_OUT, O.MPNN_B_OUT)] + \
            [(O.MPNN_LAYER0 + l * O.MPNN_LAYER_STRIDE, O.MPNN_LAYER0 + (l + 1) * O.MPNN_LAYER_STRIDE) for l in range(3)]:
        assert np.abs(g[lo:hi]).max() > 0


def test_forward_closed_forms():
    """W_l = 0: y = relu(X W_in + b) w_out + b_out.  With W_in selecting the
    first features (positive inputs) and only the message half of layer 1
    set to the identity, h_1 = A x + x and y = sum over the selected features."""
    rng = np.random.default_rng(3)
    n = 10
    ptr, col, a = random_graph(n, rng)
    x = rng.standard_normal((n, O.MPNN_IN)).astype(np.float32)
    t = np.zeros(n, np.float32)
    mask = np.ones(n, np.float32)
    p = random_params(rng)
    for l in range(3):
        o = O.MPNN_LAYER0 + l * O.MPNN_LAYER_STRIDE
        p[o:o + O.MPNN_LAYER_STRIDE] = 0
    _, y, _ = O.mpnn(ptr, col, a, x, t, mask, p, emulate=False, grad=False)
    W = p[O.MPNN_W_IN:O.MPNN_B_IN].astype(np.float64).reshape(O.MPNN_IN, O.MPNN_HID)
    h0 = np.maximum(x.astype(np.float64) @ W + p[O.MPNN_B_IN:O.MPNN_LAYER0], 0)
    ref = h0 @ p[O.MPNN_W_OUT:O.MPNN_B_OUT].astype(np.float64) + p[O.MPNN_B_OUT]
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)

    xp = np.abs(x)
    q = np.zeros(O.MPNN_NPARAM, np.float32)
    for k in range(8):
        q[O.MPNN_W_IN + k * O.MPNN_HID + k] = 1.0
    for k in range(8):   # layer 1: message half -> identity on the first 8 channels
        q[O.MPNN_LAYER0 + (O.MPNN_HID + k) * O.MPNN_HID + k] = 1.0
    q[O.MPNN_W_OUT:O.MPNN_W_OUT + 8] = 1.0
    _, y, _ = O.mpnn(ptr, col, a, xp, t, mask, q, emulate=False, grad=False)
    A = np.zeros((n, n))
    for i in range(n):
        for e in range(ptr[i], ptr[i + 1]):
            A[i, col[e]] = a[e]
    ref = ((A @ xp[:, :8].astype(np.float64)) + xp[:, :8]).sum(1)
    assert np.allclose(y, ref, rtol=1e-12)


def test_emulation_stays_close_to_double():
    rng = np.random.default_rng(11)
    n = 40
    ptr, col, a = random_graph(n, rng)
    x = rng.standard_normal((n, O.MPNN_IN)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    mask = np.ones(n, np.float32)
    p = random_params(rng)
    l64, y64, g64 = O.mpnn(ptr, col, a, x, t, mask, p, emulate=False)
    l16, y16, g16 = O.mpnn(ptr, col, a, x, t, mask, p, emulate=True)
    assert np.max(np.abs(y16 - y64)) < 0.05 * np.max(np.abs(y64))
    assert np.dot(g16, g64) / (np.linalg.norm(g16) * np.linalg.norm(g64)) > 0.99
