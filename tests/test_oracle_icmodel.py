"""Pins of the analytic emission models in the oracle (ic_oracle.c:
InverseComptonPL1D, inverse_compton_power_law.cuh:11-190, and
InverseComptonDummy, inverse_compton_dummy.cuh:7-40).  f_inv1 / f_inv2 are
inverse CDFs: on a dense grid of u their outputs are monotone and
continuous, and the density du/de has the model's power laws -- e^1 below
er = 2 gamma e_min, e^-(alpha+1) up to e_s and e^-(alpha+2) beyond (f_inv1),
e^1 below et = er / (2 er + 1), e^-1 up to er and e^-(alpha+2) beyond
(f_inv2); draw_photon_ep's density is e^1 on its first branch and
e^-(alpha+1) on its second, never below e1p; the emitted fraction is the
emission rate; free paths are exponential; photon energies carry the sign
of p; the dummy emits E_ph at photon_path above E_cutoff."""
import numpy as np
import pytest

import oracle_ctypes as oc

F32 = lambda x: float(np.float32(x))  # noqa: E731
ALPHA, ES, EMIN = 2.0, F32(0.2), F32(1e-3)


def model(**kw):
    args = dict(alpha=ALPHA, e_s=ES, e_min=EMIN, photon_path=1.0, ic_rate=0.01)
    args.update(kw)
    return oc.ICModelOracle(1, **args)


def log_slope(f, lo, hi, n=200001):
    """d ln(du/de) / d ln e of the inverse CDF f over e in (lo, hi)."""
    u = ((np.arange(n) + 0.5) / n).astype(np.float32)
    e = np.array([f(float(x)) for x in u])
    assert np.all(np.diff(e) >= 0)
    dens = np.diff(u.astype(np.float64)) / np.diff(e)
    em = np.sqrt(e[1:] * e[:-1])
    sel = (em > lo) & (em < hi) & np.isfinite(dens) & (dens > 0)
    x, y = np.log(em[sel]), np.log(dens[sel])
    return np.polyfit(x, y, 1)[0], e


def test_f_inv1_power_laws():
    o = model()
    g = 20.0  # below e_s / (2 e_min) = 100: the f_inv1 regime
    er = 2 * g * EMIN
    s1, _ = log_slope(lambda u: o.f_inv1(u, g), er * 0.05, er * 0.8)
    s2, _ = log_slope(lambda u: o.f_inv1(u, g), er * 1.3, ES * 0.7)
    s3, e = log_slope(lambda u: o.f_inv1(u, g), ES * 1.5, ES * 6)
    assert abs(s1 - 1) < 0.02 and abs(s2 + ALPHA + 1) < 0.02 and abs(s3 + ALPHA + 2) < 0.05
    # continuous through both branch points (no jump between grid neighbours)
    mid = (e > er * 0.5) & (e < ES * 2)
    assert np.max(e[1:][mid[1:]] / e[:-1][mid[1:]]) < 1.001


def test_f_inv2_power_laws():
    o = model()
    g = 400.0
    er = 2 * g * EMIN
    et = er / (2 * er + 1)
    s1, _ = log_slope(lambda u: o.f_inv2(u, g), et * 0.05, et * 0.8)
    s2, _ = log_slope(lambda u: o.f_inv2(u, g), et * 1.2, er * 0.8)
    s3, e = log_slope(lambda u: o.f_inv2(u, g), er * 1.5, er * 6)
    assert abs(s1 - 1) < 0.02 and abs(s2 + 1) < 0.02 and abs(s3 + ALPHA + 2) < 0.05
    mid = (e > et * 0.5) & (e < er * 2)
    assert np.max(e[1:][mid[1:]] / e[:-1][mid[1:]]) < 1.001


def test_draw_ep_branches():
    o = model()
    g = 20.0
    gemin2 = 2 * g * EMIN
    # first branch (e1p / (1 - 2 e1p) <= 2 gamma e_min): ep^2 uniform on
    # [e1p^2, e_lim^2], e_lim = e1p / (1 - 2 e1p)
    e1p = 0.01
    e_lim = e1p / (1 - 2 * e1p)
    s, ep = log_slope(lambda u: o.draw_ep(e1p, g, u), e1p * 1.001, e_lim * 0.999, 20001)
    assert abs(s - 1) < 0.02
    np.testing.assert_allclose([ep[0], ep[-1]], [e1p, e_lim], rtol=1e-3)
    # second branch (e1p > 2 gamma e_min): density ep^-(alpha+1) above e1p
    e1p = 0.3
    assert e1p > gemin2
    s, ep = log_slope(lambda u: o.draw_ep(e1p, g, u), e1p * 1.2, e1p * 5, 20001)
    assert abs(s + ALPHA + 1) < 0.03 and ep.min() >= e1p * (1 - 1e-12)


def test_emission_rate_and_free_path():
    n = 200000
    for gamma, rate in ((50.0, 0.3), (400.0, 0.3 * np.log(2 * F32(400 * EMIN) + F32(0.90517))
                                      / F32(400 * EMIN))):
        o = model(ic_rate=0.3)
        e, eph, path = o.emit(np.full(n, gamma), np.full(n, 0.5 * gamma), seed=3)
        frac = e.mean()
        assert abs(frac - rate) < 5 * np.sqrt(rate * (1 - rate) / n) + 1e-6, (gamma, frac, rate)
        on = e == 1
        assert np.all(eph[on] > 0) and np.all(eph[~on] == 0) and np.all(path[~on] == 0)
        # path * rate(E) / photon_path is Exp(1)
        E = np.abs(eph[on])
        r = np.where(E * EMIN < 2, (E * EMIN / 2) ** ALPHA, 2 / (E * EMIN))
        x = path[on] * r
        assert abs(x.mean() - 1) < 0.02 and abs(x.var() - 1) < 0.05
    o = model(ic_rate=1.0)
    p = np.linspace(-30, 30, 1001)
    e, eph, _ = o.emit(np.sqrt(1 + p * p), p, seed=9)
    on = e == 1
    assert np.all(np.sign(eph[on]) == np.sign(p[on]))


def test_dummy():
    o = oc.ICModelOracle(0, photon_path=0.7, E_cutoff=10.0, E_ph=4.0)
    g = np.array([1.0, 9.99, 10.0, 10.01, 1e3])
    e, eph, path = o.emit(g, -g)
    np.testing.assert_array_equal(e, [0, 0, 0, 1, 1])
    np.testing.assert_array_equal(eph, [0, 0, 0, 4.0, 4.0])
    np.testing.assert_array_equal(path, [0, 0, 0, 0.7, 0.7])
