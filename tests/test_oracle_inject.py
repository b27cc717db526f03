"""Pins of the oracle's wall injection for the batched 1D ensemble
(orc_inject_left in oracle/rt_oracle.c; the workload's open left wall, no
reference counterpart): 2 particles per pair, e- then e+, in the first
bulk cell with offsets uniform in (0, 1], inward p1, E = sqrt(1 + p^2),
the given weight, consecutive ids; the momenta are Maxwell-Juttner at the
given temperature (<gamma> = K1(1/theta)/K2(1/theta) + 3 theta, p2 and p3
symmetric, folding p1 leaves gamma unchanged); pairs past the capacity
are dropped whole and counted; draws depend on the epoch."""
import numpy as np
import scipy.special as sps

import oracle_ctypes as oc

N, G = 32, 2


def sim(cap=400000, seed=11):
    s = oc.OracleSim([N], [G], size=(1.0, 1.0, 1.0), periodic=(False, False, False),
                     num_species=2, capacity=cap, jexp=30)
    s.seed = seed
    return s


def test_layout_and_bookkeeping():
    s = sim()
    n = s.inject_left(5, 0.2, weight=0.75, epoch=3)
    assert n == 10 and s.overflow == 0
    p = s.particles()
    assert np.all(p["cell"] == G)
    assert np.all((p["x1"] > 0) & (p["x1"] <= 1.0))
    assert np.all(p["x2"] == 0) and np.all(p["x3"] == 0)
    # the pair shares its position; species alternate e-, e+
    np.testing.assert_array_equal(p["x1"][0::2], p["x1"][1::2])
    np.testing.assert_array_equal(p["flag"] >> 29, np.tile([0, 1], 5))
    assert np.all(p["p1"] >= 0)
    np.testing.assert_allclose(p["E"], np.sqrt(1 + p["p1"] ** 2 + p["p2"] ** 2 + p["p3"] ** 2),
                               rtol=1e-15)
    assert np.all(p["w"] == 0.75)
    np.testing.assert_array_equal(p["id"], np.arange(10))
    s.inject_left(2, 0.2, epoch=4)
    q = s.particles()
    np.testing.assert_array_equal(q["id"], np.arange(14))
    np.testing.assert_array_equal(q["x1"][:10], p["x1"])
    assert not np.array_equal(q["x1"][10:12], p["x1"][:2])


def test_capacity_drops_whole_pairs():
    s = sim(cap=7)
    n = s.inject_left(5, 0.1)
    assert n == 6 and s.overflow == 2


def test_maxwell_juttner_statistics():
    theta, pairs = 0.3, 50000
    s = sim()
    s.inject_left(pairs, theta, epoch=0)
    p = s.particles()
    g = p["E"]
    expect = sps.kv(1, 1 / theta) / sps.kv(2, 1 / theta) + 3 * theta
    se = g.std() / np.sqrt(g.size)
    assert abs(g.mean() - expect) < 5 * se
    for k in ("p2", "p3"):
        assert abs(p[k].mean()) < 5 * p[k].std() / np.sqrt(g.size)
    # isotropy before the fold: <p1^2> = <p2^2> = <p3^2>
    m2 = [np.mean(p[k] ** 2) for k in ("p1", "p2", "p3")]
    np.testing.assert_allclose(m2, np.mean(m2), rtol=0.03)
    # offsets uniform: mean 1/2, variance 1/12
    x = p["x1"][0::2]
    assert abs(x.mean() - 0.5) < 0.01 and abs(x.var() - 1 / 12) < 0.005
