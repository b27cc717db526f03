"""GPU parity of pair creation and photon transport against the CPU oracle
(radiative_transfer.cu:28-301, user_radiative_functions.cu:13-125,
ptc_updater.cu:627-741, ptc_updater_logsph.cu:58-132).

Bar: the FP64 form is bit-exact in the event set (which particles emit,
which photons convert), the storage order, every particle and photon value
and the ids; the mixed form reproduces the event counts of a case whose
energies are well away from the thresholds."""
import math

import numpy as np
import pytest

import aperture3_amd as apt
import oracle_ctypes as oc
from logsph_cases import LogSphCase

pytestmark = pytest.mark.gpu

PTC_KEYS = ("cell", "x1", "x2", "x3", "p1", "p2", "p3", "E", "w", "flag", "id")
PH_KEYS = ("cell", "x1", "x2", "x3", "p1", "p2", "p3", "E", "w", "path", "id")


def kat_particles(dims):
    n = 1000
    cell = 10 + 10 * dims[0]
    return {"cell": np.full(n, cell, np.uint32), "x1": np.zeros(n), "x2": np.zeros(n),
            "x3": np.zeros(n), "p1": np.where(np.arange(n) % 10 == 0, 1000.0, 0.0),
            "p2": np.zeros(n), "p3": np.zeros(n), "w": np.ones(n),
            "flag": np.zeros(n, np.uint32)}


def cart_pair(precision, **rt):
    o = oc.OracleSim((32, 32), (2, 2), size=(1.0, 1.0, 1.0), capacity=4000)
    p = apt.SimParams(N=[32, 32, 1], guard=[2, 2, 0], periodic_boundary=[True, True, True],
                      max_ptc_number=4000, precision=precision, max_photon_number=2000, **rt)
    env = apt.sim_environment(p)
    d = apt.sim_data(env)
    o.set_rt(max_photon=2000, max_ptc=4000, **{k: v for k, v in rt.items()
                                                if k in ("gamma_thr", "photon_path",
                                                         "E_secondary", "r_cutoff", "e_frac")},
             model=rt.get("rt_model", 0))
    return o, env, d


def assert_state(o, d, ptc_keys=PTC_KEYS, ph_keys=PH_KEYS):
    assert d.num_particles() == o.number
    pg, po = d.get_particles(), o.particles()
    for k in ptc_keys:
        assert np.array_equal(pg[k], po[k]), k
    hg, ho = d.get_photons(), o.photons()
    assert len(hg["cell"]) == len(ho["cell"])
    for k in ph_keys:
        assert np.array_equal(hg[k], ho[k]), ("photon", k)


def test_reference_count_kat_gpu():
    """tests/test_particles_dev.cpp:146-165 on the GPU, FP64 and mixed."""
    for prec in (apt.APT_PREC_FP64, apt.APT_PREC_MIXED):
        o, env, d = cart_pair(prec, gamma_thr=20.0, photon_path=-1.0, r_cutoff=4.0)
        d.set_particles(kat_particles(env.dims))
        rad = apt.radiative_transfer(env)
        assert rad.emit_photons(d) == 100
        d.sort_particles()
        assert d.info()["ph_number"] == 100
        assert rad.produce_pairs(d) == 100
        d.sort_particles()
        assert d.num_particles() == 1200 and d.info()["ph_number"] == 0
        # photon_produced / pair_produced: 100 unit weights in the one cell
        cell = 10 + 10 * env.dims[0]
        for which in (apt.FIELD_PHOTON_PRODUCED, apt.FIELD_PAIR_PRODUCED):
            f = d.get_field(which, 0)
            assert f[cell] == 100.0 and f.sum() == 100.0
        d.close()


@pytest.mark.parametrize("model", [0, 1])
def test_cartesian_pairs_bit_exact(model):
    rt = dict(rt_model=model, gamma_thr=50.0, photon_path=0.02 if model else 0.5,
              E_secondary=4.0, r_cutoff=1e9 if model else 10.0, e_frac=0.2)
    o, env, d = cart_pair(apt.APT_PREC_FP64, **rt)
    rng = np.random.default_rng(4)
    n = 1500
    ptc = {"cell": (rng.integers(2, 34, n) + rng.integers(2, 34, n) * 36).astype(np.uint32),
           "x1": rng.random(n), "x2": rng.random(n), "x3": np.zeros(n),
           "p1": rng.normal(0, 80, n), "p2": rng.normal(0, 80, n), "p3": rng.normal(0, 80, n),
           "w": np.ones(n), "flag": (rng.integers(0, 2, n) << 29).astype(np.uint32)}
    o.set_particles(ptc)
    d.set_particles(ptc)
    pu = apt.ptc_updater(env)
    rad = apt.radiative_transfer(env)
    dt = 0.005
    for step in range(6):
        o.sp.jexp = d.jexp()
        o.ptc_update(dt, step)
        pu.update_particles(d, dt, step)
        k1, k2 = o.emit_photons(), rad.emit_photons(d)
        assert k1 == k2
        k1, k2 = o.produce_pairs(), rad.produce_pairs(d)
        assert k1 == k2
        if step % 3 == 2:
            o.sort()
            o.sort_photons()
            d.sort_particles()
    assert o.rt_epoch == 6
    assert_state(o, d)
    # per-cell event weights: unit weights, so the atomic sums are exact
    assert np.array_equal(d.get_field(apt.FIELD_PHOTON_PRODUCED, 0), o.photon_produced)
    assert np.array_equal(d.get_field(apt.FIELD_PAIR_PRODUCED, 0), o.pair_produced)
    assert o.photon_produced.sum() > 0
    d.close()


def test_logsph_pairs_bit_exact():
    c = LogSphCase(nptc=4000, N=(48, 48))
    o = c.oracle()
    rt = dict(rt_model=1, gamma_thr=3.0, photon_path=0.05, r_cutoff=25.0, e_frac=0.2,
              max_photon_number=20000)
    env, d = c.gpu(precision=apt.APT_PREC_FP64, **rt)
    o.set_rt(model=1, gamma_thr=3.0, photon_path=0.05, r_cutoff=25.0, e_frac=0.2,
             max_photon=20000, max_ptc=c.cap)
    pu, fs = apt.ptc_updater_logsph(env), apt.field_solver_logsph(env)
    rad = apt.radiative_transfer(env)
    for step in range(8):
        o.sp.jexp = d.info()["H"]
        o.ptc_update(c.dt, step)
        pu.update_particles(d, c.dt, step)
        o.field_update(c.dt)
        fs.update_fields(d, c.dt)
        assert o.emit_photons() == rad.emit_photons(d)
        assert o.produce_pairs() == rad.produce_pairs(d)
        if step % 4 == 3:
            o.sort()
            o.sort_photons()
            d.sort_particles()
    assert d.info()["ph_number"] > 0
    assert_state(o, d)
    # event weights summed by atomics in any order, as in the reference:
    # equal to the oracle's slot-order sums up to rounding
    for which, ref in ((apt.FIELD_PHOTON_PRODUCED, o.photon_produced),
                       (apt.FIELD_PAIR_PRODUCED, o.pair_produced)):
        np.testing.assert_allclose(d.get_field(which, 0), ref, rtol=1e-13, atol=0)
    assert np.array_equal(d.get_field(apt.FIELD_EDOTB, 0), o.edotb)
    d.close()


def test_mixed_pair_counts():
    rt = dict(rt_model=1, gamma_thr=50.0, photon_path=0.02, r_cutoff=1e9, e_frac=0.2)
    o, env, d = cart_pair(apt.APT_PREC_MIXED, **rt)
    rng = np.random.default_rng(8)
    n = 2000
    big = rng.random(n) < 0.3
    p1 = np.where(big, 200.0, 5.0) * np.sign(rng.normal(size=n))
    ptc = {"cell": (rng.integers(2, 34, n) + rng.integers(2, 34, n) * 36).astype(np.uint32),
           "x1": rng.random(n), "x2": rng.random(n), "x3": np.zeros(n), "p1": p1,
           "p2": np.zeros(n), "p3": np.zeros(n), "w": np.ones(n),
           "flag": np.zeros(n, np.uint32)}
    o.set_particles(ptc)
    d.set_particles(ptc)
    rad = apt.radiative_transfer(env)
    assert o.emit_photons() == rad.emit_photons(d) == int(big.sum())
    ho, hg = o.photons(), d.get_photons()
    np.testing.assert_allclose(hg["p1"], ho["p1"], rtol=1e-6)
    np.testing.assert_allclose(hg["path"], ho["path"], rtol=1e-5, atol=1e-9)
    d.close()
