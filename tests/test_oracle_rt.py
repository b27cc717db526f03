"""Pins the oracle's pair-creation Monte Carlo and photon transport
(oracle/rt_oracle.c), which restates radiative_transfer.cu:28-301, the
default hooks of user_radiative_functions.cu:13-125 and the photon movers
(ptc_updater.cu:627-741, ptc_updater_logsph.cu:58-132).

  * the count-level known answer of the reference's own test
    (tests/test_particles_dev.cpp:146-165): 1000 electrons, every 10th at
    p1 = 1000, photon_path < 0 -> 100 photons -> 100 pairs -> 1200
    particles, no photons left;
  * energy bookkeeping per event: the parent loses exactly E_ph, each pair
    member has gamma = E_ph / 2 and the photon's direction;
  * the exponential free path has the requested mean (statistical);
  * deterministic compaction: photon k belongs to the k-th emitter;
  * capacity overflow is counted, never written past the store;
  * photons move on straight lines (Cartesian) / radially (log-spherical).
"""
import math

import numpy as np
import pytest

import oracle_ctypes as oc


def kat_state(logsph=False):
    if logsph:
        o = oc.LogSphOracle((32, 32, 1), (3, 3, 0), (math.log(30.0), math.pi, 1.0),
                            capacity=4000)
    else:
        o = oc.OracleSim((32, 32), (2, 2), size=(1.0, 1.0, 1.0), capacity=4000)
    n = 1000
    cell = 10 + 10 * o.dims[0]
    p1 = np.where(np.arange(n) % 10 == 0, 1000.0, 0.0)
    o.set_particles({"cell": np.full(n, cell), "x1": np.zeros(n), "x2": np.zeros(n),
                     "x3": np.zeros(n), "p1": p1, "p2": np.zeros(n), "p3": np.zeros(n),
                     "w": np.ones(n), "flag": np.zeros(n, np.uint32)})
    return o


@pytest.mark.parametrize("logsph", [False, True])
def test_reference_count_kat(logsph):
    o = kat_state(logsph)
    o.set_rt(model=0, gamma_thr=20.0, photon_path=-1.0, E_secondary=4.0, r_cutoff=4.0,
             max_photon=1000)
    assert o.emit_photons() == 100
    assert o.sort_photons() == 100
    assert o.produce_pairs() == 100
    assert o.sort() == 1200
    assert o.sort_photons() == 0


def test_energy_bookkeeping_and_order():
    o = kat_state()
    before = o.particles()
    o.set_rt(model=0, gamma_thr=20.0, photon_path=-1.0, E_secondary=4.0, r_cutoff=4.0,
             max_photon=1000)
    o.emit_photons()
    after = o.particles()
    ph = o.photons()
    emitters = np.nonzero(before["E"] > 20.0)[0]
    assert len(ph["E"]) == len(emitters)
    for k, i in enumerate(emitters):
        Eph = ph["E"][k]
        assert 2.5 <= Eph <= 2.5 + 2 * 3.0
        assert after["E"][i] == pytest.approx(before["E"][i] - Eph, rel=1e-15)
        pm = math.sqrt(after["p1"][i] ** 2 + after["p2"][i] ** 2 + after["p3"][i] ** 2)
        assert pm == pytest.approx(math.sqrt(after["E"][i] ** 2 - 1.0), rel=1e-12)
        assert ph["p1"][k] == pytest.approx(Eph, rel=1e-12)  # along the parent's p
        assert ph["cell"][k] == before["cell"][i]
    n0 = o.number
    o.produce_pairs()
    P = o.particles()
    for k in range(len(emitters)):
        Eph = ph["E"][k]
        for s in range(2):
            j = n0 + 2 * k + s
            assert P["E"][j] == pytest.approx(Eph / 2.0, rel=1e-14)
            assert (P["flag"][j] >> 29) == s and (P["flag"][j] & (1 << 7))
            assert P["p1"][j] > 0.0 and P["p2"][j] == 0.0


def test_exponential_free_path_mean():
    o = oc.OracleSim((64, 64), (2, 2), capacity=40000)
    n = 20000
    rng = np.random.default_rng(2)
    o.set_particles({"cell": rng.integers(2, 66, n) + rng.integers(2, 66, n) * o.dims[0],
                     "x1": rng.random(n), "x2": rng.random(n), "x3": np.zeros(n),
                     "p1": np.full(n, 1e3), "p2": np.zeros(n), "p3": np.zeros(n),
                     "w": np.ones(n), "flag": np.zeros(n, np.uint32)})
    o.set_rt(model=1, gamma_thr=100.0, photon_path=0.2, r_cutoff=1e9, e_frac=0.2,
             max_photon=n)
    assert o.emit_photons() == n
    ph = o.photons()
    assert ph["path"].mean() == pytest.approx(0.2, rel=0.03)
    assert np.all(ph["path"] >= 0.0)
    np.testing.assert_allclose(ph["E"], 0.2 * np.sqrt(1 + 1e6), rtol=1e-15)


def test_capacity_overflow_counted():
    o = kat_state()
    o.set_rt(model=0, gamma_thr=20.0, photon_path=-1.0, r_cutoff=4.0, max_photon=30)
    assert o.emit_photons() == 30
    assert o.overflow == 70
    o.set_rt(model=0, gamma_thr=20.0, photon_path=-1.0, r_cutoff=4.0, max_photon=100,
             max_ptc=1000 + 2 * 40 + 1)
    o.emit_photons()
    assert o.produce_pairs() == 40
    assert o.number == 1080 and o.overflow == 70 + 60


def test_photon_straight_lines():
    o = oc.OracleSim((16, 16), (2, 2), capacity=10)
    o.set_rt(max_photon=4)
    o.ph_number = 1
    o.ph["cell"][0] = 5 + 5 * o.dims[0]
    o.ph["x1"][0], o.ph["x2"][0] = 0.5, 0.5
    o.ph["p1"][0], o.ph["p2"][0], o.ph["p3"][0] = 3.0, 4.0, 0.0
    o.ph["path"][0] = 1.0
    dt = 0.01
    for _ in range(40):
        o.move_photons(dt)
    c = int(o.ph["cell"][0])
    x = ((c % o.dims[0]) - 2 + o.ph["x1"][0]) / 16
    y = ((c // o.dims[0]) - 2 + o.ph["x2"][0]) / 16
    assert x == pytest.approx((3.5 / 16 + 40 * dt * 0.6) % 1.0, abs=1e-12)
    assert y == pytest.approx((3.5 / 16 + 40 * dt * 0.8) % 1.0, abs=1e-12)
    assert o.ph["path"][0] == pytest.approx(1.0 - 40 * dt, abs=1e-12)


def test_photon_radial_logsph():
    o = oc.LogSphOracle((32, 32, 1), (3, 3, 0), (math.log(30.0), math.pi, 1.0), capacity=10)
    o.set_rt(max_photon=4, r_cutoff=20.0)
    o.ph_number = 1
    o.ph["cell"][0] = 8 + 12 * o.dims[0]
    o.ph["x1"][0], o.ph["x2"][0] = 0.5, 0.5
    o.ph["p1"][0] = 5.0
    o.ph["path"][0] = 1.0
    D0 = o.mesh.delta[0]
    r0 = math.exp(D0 * (8 - 3 + 0.5))
    dt = 0.01
    for _ in range(20):
        o.move_photons(dt)
    c = int(o.ph["cell"][0])
    r = math.exp(D0 * ((c % o.dims[0]) - 3 + o.ph["x1"][0]))
    assert r == pytest.approx(r0 + 20 * dt, rel=1e-12)
    assert c // o.dims[0] == 12 and o.ph["x2"][0] == pytest.approx(0.5, abs=1e-12)
    assert o.ph["p1"][0] == pytest.approx(5.0, rel=1e-12)


@pytest.mark.parametrize("logsph", [False, True])
def test_event_counters_kat(logsph):
    """photon_produced / pair_produced (radiative_transfer.cu:48-57,
    120-130) on the known answer: the 100 emitters and the 100 converting
    photons all sit in one cell with unit weight."""
    o = kat_state(logsph)
    o.set_rt(model=0, gamma_thr=20.0, photon_path=-1.0, E_secondary=4.0, r_cutoff=4.0,
             max_photon=1000)
    cell = 10 + 10 * o.dims[0]
    o.emit_photons()
    assert o.photon_produced[cell] == 100.0 and o.photon_produced.sum() == 100.0
    o.produce_pairs()
    assert o.pair_produced[cell] == 100.0 and o.pair_produced.sum() == 100.0


def test_event_counters_count_before_capacity():
    """The reference counts every emitter in its counting pass, before the
    photon slots are assigned: emitters dropped for capacity still count;
    weights add as weights."""
    o = kat_state()
    P = o.particles()
    P["w"] = np.where(np.arange(len(P["w"])) % 20 == 0, 0.25, 1.0)
    o.set_particles(P)
    o.set_rt(model=0, gamma_thr=20.0, photon_path=-1.0, r_cutoff=4.0, max_photon=30)
    assert o.emit_photons() == 30
    # emitters are every 10th particle; every 20th of all carries w = 0.25
    assert o.photon_produced.sum() == 50 * 0.25 + 50 * 1.0
