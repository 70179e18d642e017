"""The oracle's fourth-order Runge-Kutta step (flight integrator 2, an
addition without a reference counterpart) pinned by mathematics, and the
ion-table draw pinned by its definition:

* in a uniform field the step is exact (positions quadratic and velocities
  linear in time) and its error estimate vanishes;
* on the harmonic oscillator x'' = -x the global error at t = 8 falls by
  2^4 when the step halves (fourth order), and the one-step error estimate
  (k4 - k5) / 6 falls by 2^4 as well (the embedded solution is of third
  order) -- a wrong stage offset, weight or estimate breaks one of these;
* the ion draw picks entry k with frequency weight_k / sum (chi-square over
  20 000 events), is a pure function of (seed, event index), and the splitmix64
  mixer matches its published test values."""
import numpy as np
import pytest

from oracle import oracle as O


def test_rk4_uniform_field_is_exact():
    g = np.array([0.0, 1.5, -2.0], np.float32)
    p0 = np.array([0.1, -0.2, 0.3], np.float32)
    v0 = np.array([1.0, 0.5, 0.0], np.float32)
    h, n = 0.125, 16
    t, p, v, ep, ev = O.rk4_analytic(0, g, 1.0, p0, v0, h, n)
    T = h * n
    assert t == pytest.approx(T, rel=1e-6)
    np.testing.assert_allclose(p, p0 + v0 * T + 0.5 * g * T * T, rtol=0, atol=2e-6)
    np.testing.assert_allclose(v, v0 + g * T, rtol=0, atol=2e-6)
    assert np.max(np.abs(ep)) <= 1e-7 and np.max(np.abs(ev)) <= 1e-7


def harmonic(h, steps):
    return O.rk4_analytic(1, [1.0, 1.0, 1.0], 1.0, [1.0, 0.0, 0.0], [0.0, 0.0, 0.0], h, steps)


def test_rk4_is_fourth_order():
    errs = []
    for h in (0.8, 0.4, 0.2):
        t, p, v, _, _ = harmonic(h, int(round(8.0 / h)))
        errs.append(abs(float(p[0]) - np.cos(float(t))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12.0 < r1 < 20.0 and 12.0 < r2 < 20.0, (errs, r1, r2)


def test_rk4_error_estimate_is_third_order_embedded():
    est = []
    for h in (0.4, 0.2, 0.1):
        t, p, v, ep, ev = harmonic(h, 1)
        est.append(abs(float(ep[0])))
        # the estimate is the third-order solution's error: y4 - y3, with y4
        # far closer to the exact value
        assert abs(float(p[0]) - np.cos(h)) < 0.1 * abs(float(ep[0]))
    assert 15.0 < est[0] / est[1] < 17.0 and 15.0 < est[1] / est[2] < 17.0, est


def splitmix64(x):
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_splitmix64_published_values():
    # the generator's published outputs from state 0 (Vigna's reference
    # implementation): the state advances by the golden-ratio increment
    state, out = 0, []
    for _ in range(3):
        out.append(splitmix64(state))
        state = (state + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_ion_pick_matches_definition_and_frequencies():
    w = np.array([0.5, 0.0, 0.3, 0.2], np.float32)
    seed = 12345
    picks = np.array([O.ion_pick(seed, e, w) for e in range(20000)])
    # the definition, restated from the mixer
    for e in range(0, 20000, 997):
        u = (splitmix64((seed << 32) ^ e) >> 11) / float(1 << 53)
        cum = np.cumsum(w.astype(np.float64))
        assert picks[e] == int(np.searchsorted(cum, u * cum[-1], side="right"))
    counts = np.bincount(picks, minlength=4)
    assert counts[1] == 0
    expected = w[[0, 2, 3]] / w.sum() * len(picks)
    chi2 = float(np.sum((counts[[0, 2, 3]] - expected) ** 2 / expected))
    assert chi2 < 13.8  # 2 degrees of freedom, p = 0.001
    assert all(O.ion_pick(seed, e, w) == picks[e] for e in range(50))
    assert any(O.ion_pick(seed + 1, e, w) != picks[e] for e in range(50))
