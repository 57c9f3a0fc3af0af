"""Pins of the oracle's periodic (Ewald) direct sum: independence of the
splitting parameter, periodicity, antisymmetry and the near-field limit."""
import numpy as np
import pytest


@pytest.mark.parametrize("box", [(32.0, 32.0, 32.0), (32.0, 64.0, 16.0)])
def test_ewald_independent_of_alpha(oracle, box):
    rng = np.random.default_rng(1)
    L = np.array(box)
    for _ in range(5):
        x = rng.uniform(-0.5, 0.5, 3) * L
        a1 = oracle.ewald_accel(box, x, alpha=1.5 / L.min())
        a2 = oracle.ewald_accel(box, x, alpha=3.0 / L.min())
        assert np.allclose(a1, a2, rtol=1e-9, atol=1e-13)


def test_ewald_periodic_and_odd(oracle):
    box = (32.0, 48.0, 64.0)
    rng = np.random.default_rng(2)
    for _ in range(5):
        x = rng.uniform(-10, 10, 3)
        a = oracle.ewald_accel(box, x)
        assert np.allclose(oracle.ewald_accel(box, x + [32.0, 0, 0]), a, rtol=1e-9, atol=1e-13)
        assert np.allclose(oracle.ewald_accel(box, x - [0, 48.0, 64.0]), a, rtol=1e-9, atol=1e-13)
        assert np.allclose(oracle.ewald_accel(box, -x), -a, rtol=1e-9, atol=1e-13)


def test_ewald_near_field(oracle):
    # cubic lattice: a(x) = -x/|x|^3 + (4 pi / 3V) x + O(|x|^3 / L^5); the
    # remainder shrinks like |x|^2 relative to the linear term
    L = 32.0
    errs = []
    for s in (1.0, 0.5, 0.25):
        x = np.array([0.6, -0.4, 0.9]) * s
        corr = oracle.ewald_accel((L, L, L), x) + x / np.linalg.norm(x) ** 3
        lin = 4 * np.pi / (3 * L ** 3) * x
        errs.append(np.linalg.norm(corr - lin) / np.linalg.norm(lin))
    assert errs[0] < 1e-2
    assert errs[1] < 0.35 * errs[0] and errs[2] < 0.35 * errs[1]


def test_periodic_sum_two_bodies_is_pair_force(oracle):
    box = (16.0, 16.0, 16.0)
    pos = np.array([[3.0, 4.0, 5.0], [9.5, 1.0, 12.0]], np.float32)
    m = np.array([2.0, 0.5], np.float32)
    a = oracle.direct_sum_periodic(pos, m, 0.25, box)
    d = pos[0].astype(np.float64) - pos[1]
    assert np.allclose(a[0], 0.25 * 0.5 * oracle.ewald_accel(box, d), rtol=1e-12)
    assert np.allclose(a[1], -0.25 * 2.0 * oracle.ewald_accel(box, d), rtol=1e-12)
