"""Pins of the triplet pair-production oracle (oracle/ic_oracle.c, the
reference's triplet_pairs, src/cuda/radiation/rt_tpp.cu:21-294): the
cross-section fit starts at the s = 4 threshold, is continuous across its
pieces and tends to the Bethe-Heitler (28/9) ln 2s - 218/27; the mean pair
energy fraction is 1/3 at threshold (three equal particles) and its two
fits join at s = 100; at threshold E_plus_min = E_plus_max = gamma/3; the
rate grows as (28/9) ln gamma with the Thomson-like weight the IC table
measures on the same grid; the pair-energy CDF rises by exp(-1.75 dg) per
grid step inside the kinematic window (the E^-1.75 spectrum); samples
follow the CDF to within one bin."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
from oracle_ctypes import ICOracle, TPPOracle  # noqa: E402

from ic_cases import SPECTRA, THOMSON  # noqa: E402

NG = 40
RE2_137 = 7.91402e-26 / 137.0


@pytest.fixture(scope="module")
def mono():
    o = TPPOracle(NG)
    kind, p, emin, emax, n0 = SPECTRA["mono"]
    o.init(kind, p, emin, emax, n0)
    return o


def test_cross_section_fit():
    o = TPPOracle(4)
    assert o.total_f(3.99) == 0.0 and o.total_f(4.0) == 0.0
    assert 0 < o.total_f(4.01) < 1e-5
    for s in (4.6, 6.0, 14.0):
        a, b = o.total_f(s * (1 - 1e-12)), o.total_f(s)
        assert abs(a - b) < 1e-4 * max(b, 1e-3), s
    for s in (1e6, 1e8, 1e10):
        bh = 28 / 9 * np.log(2 * s) - 218 / 27
        np.testing.assert_allclose(o.total_f(s), bh, rtol=np.log(2 * s) ** 3 / s)
    s = np.geomspace(4.001, 1e6, 400)
    assert np.all(np.diff([o.total_f(x) for x in s]) > 0)


def test_mean_pair_energy_fit():
    o = TPPOracle(4)
    assert o.approx_em(4.01) == 0.0
    np.testing.assert_allclose(o.approx_em(4.01 * (1 + 1e-12)), 1 / 3, rtol=1e-3)
    np.testing.assert_allclose(o.approx_em(100.0), o.approx_em(100.0 * (1 + 1e-12)), rtol=1e-5)
    for s in (1e3, 1e6):
        np.testing.assert_allclose(o.approx_em(s), 0.187 * np.log(2 * s) ** 2 / s, rtol=1e-14)


def test_pair_energy_bounds():
    o = TPPOracle(4)
    g = 100.0
    lo, hi = o.e_plus(g, 4.0)
    assert lo == hi
    np.testing.assert_allclose(lo, g / 3, rtol=1e-15)
    lo, hi = o.e_plus(g, 3.0)
    assert np.isnan(lo) and np.isnan(hi)
    for s in (5.0, 50.0, 5e4):
        lo, hi = o.e_plus(g, s)
        assert 0 < lo < g / 3 < hi < g
        # the two roots of (1 + 2s) E^2 - 2 gamma (s - 1) E + c: their sum
        np.testing.assert_allclose(lo + hi, 2 * g * (s - 1) / (1 + 2 * s), rtol=1e-14)


def test_rate_grows_as_bethe_heitler(mono):
    """d rate / d ln gamma -> n0 r_e^2 / 137 (28/9) W with W the weight
    sum_mu 0.5 dmu (1 - beta mu) sum_e n(e) e de, which the IC table's
    Thomson end measures as ic_rate[0] / (n0 sigma_T)."""
    ic = ICOracle(NG, 4)
    kind, p, emin, emax, n0 = SPECTRA["mono"]
    ic.init(kind, p, emin, emax, n0)
    w = ic.ic_rate[0] / (n0 * THOMSON)
    slope = np.diff(mono.rate[-6:]) / mono.dg
    np.testing.assert_allclose(slope / (n0 * RE2_137 * 28 / 9), w, rtol=2e-3)
    # zero below threshold: s <= gamma e (1 + beta) < 4 for every photon
    smax = mono.gammas * (p[0] + p[1]) * 2.0
    assert np.all(mono.rate[smax < 3.9] == 0.0) and np.all(mono.em[smax < 3.9] == 0.0)
    on = mono.rate > 0
    assert np.all(mono.em[on] > 0) and np.all(mono.em[on] < 1 / 3)
    assert np.all(np.diff(mono.em[on]) < 0)


def test_pair_spectrum_power_law(mono):
    fin = ~np.isnan(mono.dnde).any(axis=1)
    # finite exactly where every photon of the grid has s = gamma e >= 4
    emin = SPECTRA["mono"][2]
    np.testing.assert_array_equal(fin, mono.gammas * emin >= 4.0)
    for n in np.flatnonzero(fin):
        row = mono.dnde[n]
        assert row[-1] == 1.0 and np.all(np.diff(row) >= 0)
        d = np.diff(np.concatenate([[0.0], row]))
        assert np.all(d[mono.gammas > mono.gammas[n]] == 0)
        # away from the window's edges (the two photon energies of the mono
        # line have slightly different windows) and above the CDF's rounding
        inner = np.flatnonzero(d > 1e-8)[1:-2]
        if inner.size >= 2:
            np.testing.assert_allclose(d[inner[1:]] / d[inner[:-1]], np.exp(-1.75 * mono.dg),
                                       rtol=1e-5)


def test_samples_follow_cdf(mono):
    n_s = 100000
    for gn in (25, 35):
        g = mono.gammas[gn] * (1 + 1e-12)
        e = mono.gen_ep(np.full(n_s, g), seed=7)
        assert np.all(e <= g - 1.01) and np.all(e > 0)
        cdf = mono.dnde[gn]
        tol = 4 / np.sqrt(n_s)
        for i in range(1, NG - 1):
            emp = (e <= mono.gammas[i]).mean()
            assert cdf[i - 1] - tol <= emp <= cdf[i + 1] + tol, (gn, i)
    a = mono.gen_ep(np.geomspace(1, 1e16, 100), seed=1, epoch=2)
    np.testing.assert_array_equal(a, mono.gen_ep(np.geomspace(1, 1e16, 100), seed=1, epoch=2))


def test_rate_interpolation(mono):
    gn = np.arange(1, NG - 1)
    r, m = mono.rates(mono.gammas[gn] * (1 + 1e-13))
    np.testing.assert_allclose(r, mono.rate[gn], rtol=1e-9, atol=1e-11 * mono.rate.max())
    np.testing.assert_allclose(m, mono.em[gn], rtol=1e-9, atol=1e-11)
    r, m = mono.rates(np.array([2e15]))
    assert r[0] == mono.rate[-1] and m[0] == mono.em[-1]
