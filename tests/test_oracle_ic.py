"""Pins of the inverse-Compton oracle (oracle/ic_oracle.c, the reference's
inverse_compton of src/cuda/radiation/rt_ic.cu:22-547) against what the
physics fixes: the Klein-Nishina cross section's Thomson expansion and
ultra-relativistic asymptote, the Breit-Wheeler cross section's threshold
expansion, the Thomson-limit scattering rate n0 sigma_T (to the 100/99 of
the reference's closed mu quadrature), the KN-limit rate from the
asymptote, cumulative spectra that rise to 1, sampled energies that follow
those spectra to within one grid bin, and rate interpolation that is exact
at the grid points and linear in ln(gamma) between them."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
from oracle_ctypes import ICOracle  # noqa: E402

from ic_cases import SPECTRA, THOMSON  # noqa: E402

NG, NEP = 40, 40


@pytest.fixture(scope="module")
def soft():
    o = ICOracle(NG, NEP)
    kind, p, emin, emax, n0 = SPECTRA["soft_pl"]
    o.init(kind, p, emin, emax, n0)
    return o


def test_grids():
    o = ICOracle(NG, NEP)
    assert o.gammas[0] == float(np.float32(1.01))  # MIN_GAMMA is 1.01f
    np.testing.assert_allclose(o.gammas[-1], 1e15, rtol=1e-13)
    r = o.gammas[1:] / o.gammas[:-1]
    np.testing.assert_allclose(r, r[0], rtol=1e-12)
    np.testing.assert_allclose(o.ep, np.arange(NEP) / (NEP - 1), rtol=0, atol=3e-16)
    np.testing.assert_allclose(o.log_ep, np.logspace(-10, 0, NEP), rtol=1e-13)


def test_sigma_ic_limits():
    o = ICOracle(4, 4)
    # Thomson expansion 1 - 2x + 26/5 x^2 - 133/10 x^3: below x = 1e-3 the
    # reference keeps the quadratic; above, the closed form loses ~eps/x^3
    for x in (1e-5, 5e-4):
        assert abs(o.sigma_ic(x) - (1 - 2 * x + 5.2 * x * x)) < 1e-15, x
        assert abs(o.sigma_ic(x) - (1 - 2 * x + 5.2 * x * x - 13.3 * x ** 3)) < 14 * x ** 3
    for x in (2e-3, 1e-2, 3e-2):
        series = 1 - 2 * x + 5.2 * x * x - 13.3 * x ** 3
        assert abs(o.sigma_ic(x) - series) < 40 * x ** 4 + 1e-15 / x ** 3, x
    # ultra-relativistic: (3 / 8x) (ln 2x + 1/2)
    for x in (1e5, 1e8, 1e11):
        np.testing.assert_allclose(o.sigma_ic(x), 3 / (8 * x) * (np.log(2 * x) + 0.5),
                                   rtol=5 * np.log(x) / x)
    # monotone decreasing
    xs = np.logspace(-6, 8, 200)
    v = np.array([o.sigma_ic(x) for x in xs])
    assert np.all(np.diff(v) < 0)


def test_sigma_gg_threshold_and_peak():
    o = ICOracle(4, 4)
    # near threshold: (1 - b^2)[(3 - b^4) ln((1+b)/(1-b)) - 2b(2 - b^2)]
    # = 2b + 2b^3 - 24/5 b^5 + O(b^7)
    for b in (1e-4, 1e-3, 1e-2, 5e-2):
        np.testing.assert_allclose(o.sigma_gg(b), 2 * b + 2 * b ** 3 - 4.8 * b ** 5,
                                   rtol=10 * b ** 6, atol=1e-15)
    # vanishes as b -> 1 and peaks (3/16 of it is ~0.26 sigma_T) near s ~ 2
    assert o.sigma_gg(1 - 1e-12) < 1e-9
    bs = np.linspace(0.01, 0.99, 197)
    v = np.array([o.sigma_gg(b) for b in bs])
    peak = 3 / 16 * v.max()
    assert 0.25 < peak < 0.27
    s_peak = 1 / (1 - bs[v.argmax()] ** 2)
    assert 1.6 < s_peak < 2.4


def test_ic_rate_thomson_limit(soft):
    """At gamma = 1.01 and e <= 1e-4 every x < 1e-3: the rate is
    n0 sigma_T times the mu quadrature's 100/99 (closed endpoints) times
    the e quadrature of the normalised spectrum (1 - emin/emax = 0.99 for
    alpha = 1, to the rectangle rule's edge error)."""
    ratio = soft.ic_rate[0] / THOMSON
    np.testing.assert_allclose(ratio, 100 / 99 * 0.99, rtol=0.01)
    # flat across the Thomson regime, falling in the Klein-Nishina regime
    np.testing.assert_allclose(soft.ic_rate[:4] / soft.ic_rate[0], 1.0, rtol=1e-2)
    assert np.all(np.diff(soft.ic_rate[-15:]) < 0)


def test_ic_rate_klein_nishina_asymptote():
    """Mono-energetic photons at e0 = 1e-5 and gamma ~ 1e15: x ~ 1e10 and
    (1 - mu) sigma(x) -> (3 / 8 gamma e)(ln(2 gamma e (1 - mu)) + 1/2)."""
    o = ICOracle(NG, NEP)
    kind, p, emin, emax, n0 = SPECTRA["mono"]
    o.init(kind, p, emin, emax, n0)
    g, e0 = o.gammas[-1], p[0]
    mu = np.arange(100) * (2 / 99) - 1
    mu = mu[1 - mu > 1e-9]
    expect = 0.5 * (2 / 99) * np.sum(3 / (8 * g * e0) * (np.log(2 * g * e0 * (1 - mu)) + 0.5))
    np.testing.assert_allclose(o.ic_rate[-1] / THOMSON, expect, rtol=0.03)
    # and the Thomson end of the same table
    np.testing.assert_allclose(o.ic_rate[0] / THOMSON, 100 / 99, rtol=0.03)


def test_gg_rate_threshold():
    """No pairs below eph = 2 or where eph * emax <= 1 (s <= 1 at every
    mu); positive once eph * emax > 1."""
    o = ICOracle(NG, NEP)
    kind, p, emin, emax, n0 = SPECTRA["soft_pl"]
    o.init(kind, p, emin, emax, n0)
    below = o.gammas * emax <= 1.0
    assert np.all(o.gg_rate[below] == 0.0)
    assert np.all(o.gg_rate[~below][1:] > 0.0)


@pytest.mark.parametrize("name", sorted(SPECTRA))
def test_cumulative_spectra(name):
    o = ICOracle(NG, NEP)
    kind, p, emin, emax, n0 = SPECTRA[name]
    o.init(kind, p, emin, emax, n0)
    for t in (o.dnde, o.dnde_th):
        rows = ~np.isnan(t).any(axis=1)
        assert rows.sum() >= NG // 2
        assert np.all(t[rows, -1] == 1.0)
        assert np.all(np.diff(t[rows], axis=1) >= 0)
        assert np.all(t[rows] >= 0)
    # e1 = 0 never scatters: the linear table starts at 0
    rows = ~np.isnan(o.dnde).any(axis=1)
    assert np.all(o.dnde[rows, 0] == 0.0)


@pytest.mark.parametrize("gn", [10, 20, 30])
def test_samples_follow_table_within_one_bin(soft, gn):
    """Above the two lowest bins the sampler inverts the linear-grid CDF:
    P(e1 <= ep[i]) lies between the table's CDF one bin either side."""
    g = soft.gammas[gn] * (1 + 1e-12)
    n = 200000
    e1 = soft.gen_photon_e(np.full(n, g), seed=5) / g
    assert np.all(e1 >= 0) and np.all(e1 * g <= g - 1.01)
    cdf = soft.dnde[gn]
    emp = np.array([(e1 <= x).mean() for x in soft.ep])
    tol = 4 / np.sqrt(n)
    for i in range(3, NEP - 1):
        assert cdf[i - 1] - tol <= emp[i] <= cdf[i + 1] + tol, (i, emp[i], cdf[i - 1:i + 2])


def test_samples_deterministic_and_clamped(soft):
    g = np.geomspace(0.5, 1e16, 300)
    a = soft.gen_photon_e(g, seed=11, epoch=3)
    b = soft.gen_photon_e(g, seed=11, epoch=3)
    c = soft.gen_photon_e(g, seed=11, epoch=4)
    np.testing.assert_array_equal(a, b)
    assert np.any(a != c)
    assert np.all(a >= 0) and np.all(a <= np.maximum(g - 1.01, 0))


def test_accept_reject(soft):
    """generate_photon_energies: with the rate scaled to ~0.5 per sample
    (n0 = 0.5 / sigma_T) about half the samples scatter; with n0 = 1 none."""
    g = np.full(20000, 2.0)
    assert np.all(soft.gen_photon_e(g, seed=3, accept=True) == 0.0)
    o = ICOracle(NG, NEP)
    kind, p, emin, emax, _ = SPECTRA["soft_pl"]
    o.init(kind, p, emin, emax, 0.5 / THOMSON)
    rate = o.rates(g[:1])[0][0]
    frac = (o.gen_photon_e(g, seed=3, accept=True) > 0).mean()
    assert abs(frac - rate) < 0.02 and 0.3 < rate < 0.7


def test_rate_interpolation(soft):
    gn = np.arange(1, NG - 1)
    at = soft.gammas[gn] * (1 + 1e-13)
    ic, gg = soft.rates(at)
    np.testing.assert_allclose(ic, soft.ic_rate[gn], rtol=1e-9)
    np.testing.assert_allclose(gg, soft.gg_rate[gn], rtol=1e-9, atol=1e-11 * soft.gg_rate.max())
    mid = np.sqrt(soft.gammas[gn] * soft.gammas[gn + 1])
    ic, gg = soft.rates(mid)
    np.testing.assert_allclose(ic, 0.5 * (soft.ic_rate[gn] + soft.ic_rate[gn + 1]), rtol=1e-9)
    # outside the table: constant above 1e15, no pairs below eph = 2
    ic, gg = soft.rates(np.array([2e15, 1e20, 1.5, 0.5]))
    assert ic[0] == ic[1] == soft.ic_rate[-1] and gg[0] == soft.gg_rate[-1]
    assert gg[2] == 0.0 and gg[3] == 0.0


def test_random_gamma():
    o = ICOracle(4, 4)
    g = o.random_gamma(100000, seed=2)
    assert np.all(g >= 1.01 * (1 - 1e-12)) and np.all(g <= 1e10 * (1 + 1e-12))
    lg = (np.log(g) - np.log(1.01)) / (np.log(1e10) - np.log(1.01))
    assert abs(lg.mean() - 0.5) < 0.01 and abs(lg.var() - 1 / 12) < 0.005


@pytest.fixture(scope="module")
def mono_fine():
    o = ICOracle(NG, 200)
    kind, p, emin, emax, n0 = SPECTRA["mono"]
    o.init(kind, p, emin, emax, n0)
    return o, p[0]


def test_thomson_table_mean_energy(mono_fine):
    """Thomson regime (Gamma = 4 gamma e0 << 1): the isotropic
    Blumenthal-Gould kernel 2q ln q + (1 + 2q)(1 - q) has mean q = 1/3,
    so the mean scattered energy is (4/3) gamma^2 e0, i.e. e1 = (4/3) gamma
    e0.  The log-grid table holds the number CDF at its points."""
    o, e0 = mono_fine
    for gn in (4, 5):
        g = o.gammas[gn]
        assert 4 * g * e0 < 5e-3
        dc = np.diff(np.concatenate([[0.0], o.dnde_th[gn]]))
        mean = np.sum(dc * o.log_ep)
        np.testing.assert_allclose(mean / (4 / 3 * g * e0), 1.0, rtol=0.01)


def test_klein_nishina_table_limit(mono_fine):
    """Gamma -> infinity at fixed e1: q = e1 / (Gamma (1 - e1)) -> 0 and
    Gamma q -> e1 / (1 - e1), so the kernel tends to 1 + e1^2 / (2 (1 - e1))
    on 0 < e1 < 1; the linear-grid CDF is its running sum."""
    o, _ = mono_fine
    e1 = o.ep
    dens = np.zeros_like(e1)
    m = (e1 > 0) & (e1 < 1)
    dens[m] = 1 + e1[m] ** 2 / (2 * (1 - e1[m]))
    limit = np.cumsum(dens) / dens.sum()
    np.testing.assert_allclose(o.dnde[-1], limit, rtol=1e-7, atol=1e-12)
