"""CPU tests that pin the oracle's gravity timestep criteria (timestep.c:284-344, :525-565) by their
meaning: the timestep is the time in which the acceleration displaces a particle by
ErrTolIntAccuracy x its softening length, 0.5 |a| dt^2 = eta eps (in physical units when comoving)."""
import numpy as np
import pytest


def _params(oracle, soft=(0.01, 0.04)):
    return oracle.make_params(periodic=False, box_size=0.0, soft=soft)


def test_displacement_identity(oracle):
    g = np.random.default_rng(1)
    acc = g.normal(size=(100, 3)) * 10
    st = g.integers(0, 2, 100).astype(np.uint8)
    eps = np.array([0.01, 0.04])[st]
    dt, dd, below = oracle.gravity_timesteps(_params(oracle), acc, np.ones(100), stype=st, eta=0.025, max_dt=1e9)
    a = np.linalg.norm(acc, axis=1)
    np.testing.assert_allclose(0.5 * a * dt ** 2, 0.025 * eps, rtol=1e-12)
    assert below == 0 and dd == 1e9


def test_comoving_physical_units(oracle):
    # physical acceleration a2inv |a|, physical softening a eps, dt in log a via H(a)
    g = np.random.default_rng(2)
    acc = g.normal(size=(50, 3))
    A, H = 0.25, 3.0
    dt, _, _ = oracle.gravity_timesteps(_params(oracle), acc, np.ones(50), eta=0.01, max_dt=1e9, time_a=A, hubble_a=H,
                                        comoving=True)
    a_phys = np.linalg.norm(acc, axis=1) / A ** 2
    np.testing.assert_allclose(0.5 * a_phys * (dt / H) ** 2, 0.01 * A * 0.01, rtol=1e-12)


def test_caps_and_special_cases(oracle):
    acc = np.array([[0.0, 0, 0], [1e6, 0, 0], [1.0, 0, 0], [1.0, 0, 0]])
    mass = np.array([1.0, 1.0, 0.0, 1.0])
    dt, _, below = oracle.gravity_timesteps(_params(oracle), acc, mass, eta=0.025, max_dt=0.5, min_dt=1e-3)
    assert dt[0] == 0.5            # no acceleration: the cap
    assert dt[1] == 1e-3 and below == 1  # below MinSizeTimestep: raised and counted
    assert dt[2] == 0.5            # massless: the cap
    assert dt[3] == pytest.approx(np.sqrt(2 * 0.025 * 0.01), rel=1e-14)


def test_pm_displacement_constraint(oracle):
    g = np.random.default_rng(3)
    acc = g.normal(size=(200, 3)) * 0.1
    apm = g.normal(size=(200, 3)) * 30
    dt, dd, _ = oracle.gravity_timesteps(_params(oracle, soft=(0.01,)), acc, np.ones(200), acc_pm=apm, eta=0.025,
                                         max_dt=1e9)
    # twice the smallest PM timestep (one timebin up), and no particle's dt above it
    assert dd == pytest.approx(2 * np.min(np.sqrt(2 * 0.025 * 0.01 / np.linalg.norm(apm, axis=1))), rel=1e-14)
    assert np.all(dt <= dd) and np.any(dt == dd)
    _, dd2, _ = oracle.gravity_timesteps(_params(oracle, soft=(0.01,)), acc, np.ones(200), acc_pm=apm, eta=0.025,
                                         max_dt=dd / 3)
    assert dd2 == dd / 3
