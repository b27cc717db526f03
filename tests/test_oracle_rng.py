"""Oracle checks for the counter-based RNG, the own log and the
Maxwell-Juttner loader (CPU only).

Philox4x32-10 is pinned by the published Random123 known-answer vectors
(Salmon et al. 2011); the log by math.log; the loader by the closed-form
Maxwell-Juttner mean <gamma> = K1(1/theta)/K2(1/theta) + 3 theta, isotropy
and the lab-frame mean velocity of a boosted distribution (= drift)."""
import math

import numpy as np
import pytest
import scipy.special as sps

import oracle_ctypes as oc


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expect):
    assert tuple(oc.philox(ctr, key)) == expect


def test_log_matches_libm():
    lib = oc.load()
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.random(20000), 10.0 ** rng.uniform(-300, 300, 20000),
                         1.0 + rng.uniform(-1e-6, 1e-6, 2000), [1.0, 2.0, 0.5, 2 ** -53]])
    worst = 0.0
    for x in xs:
        r = lib.orc_log(float(x))
        t = math.log(x)
        if t == 0.0:
            assert r == 0.0
            continue
        worst = max(worst, abs(r - t) / abs(t))
    assert worst < 2e-15


def _mj_sample(theta, beta, n=60000, seed=99):
    us = np.array([oc.sample_mj(theta, beta, seed, i)[1] for i in range(n)])
    return us


@pytest.mark.parametrize("theta", [0.01, 0.05, 0.1, 0.3, 0.6, 1.0])
def test_mj_mean_gamma(theta):
    u = _mj_sample(theta, 0.0)
    g = np.sqrt(1.0 + (u ** 2).sum(1))
    expect = sps.kv(1, 1 / theta) / sps.kv(2, 1 / theta) + 3 * theta
    # statistical: sigma(<gamma>) = std/sqrt(n)
    tol = 5 * g.std() / np.sqrt(len(g))
    assert abs(g.mean() - expect) < tol
    # isotropy
    for c in range(3):
        assert abs(u[:, c].mean()) < 5 * u[:, c].std() / np.sqrt(len(u))
    m2 = (u ** 2).mean(0)
    assert np.allclose(m2, m2.mean(), rtol=0.05)


@pytest.mark.parametrize("beta", [0.3, -0.3])
def test_mj_drift_mean_velocity(beta):
    # number-weighted lab-frame mean of u_x/gamma of a boosted isotropic
    # distribution equals the boost velocity
    u = _mj_sample(0.05, beta)
    g = np.sqrt(1.0 + (u ** 2).sum(1))
    vx = u[:, 0] / g
    assert abs(vx.mean() - beta) < 5 * vx.std() / np.sqrt(len(vx))


def test_loader_layout():
    o = oc.OracleSim([8], [2], capacity=8 * 3 * 2)
    n = o.load_uniform(3, (0.1, 0.1), (0.0, 0.0), 1.0, 43210)
    p = o.particles()
    assert n == 48
    # co-located pairs, electron first (fill_particles, cuda/sim_data.cu:124-161)
    assert np.array_equal(p["x1"][0::2], p["x1"][1::2])
    assert np.all(p["flag"][0::2] >> 29 == 0) and np.all(p["flag"][1::2] >> 29 == 1)
    assert np.all((p["x1"] >= 0) & (p["x1"] < 1))
    assert np.array_equal(p["cell"], np.repeat(np.arange(2, 10), 6).astype(np.uint32))
    assert np.array_equal(p["id"], np.arange(48, dtype=np.uint64))
    assert np.allclose(p["E"], np.sqrt(1 + p["p1"] ** 2 + p["p2"] ** 2 + p["p3"] ** 2))
