"""GPU parity of the inverse-Compton tables and sampling (apt_ic_*, the
reference's inverse_compton, src/cuda/radiation/rt_ic.cu:22-547) against
the oracle: FP64 bit-exact in the grids, the IC and gamma-gamma rates, both
cumulative spectra, the sampled photon energies (with and without the
accept step), the interpolated rates and the random Lorentz factors; the
reference's own test case (tests/test_ic.cu:24-69: broken power law, gamma
= 10) gives energies in [0, gamma - 1.01].  The triplet tables
(apt_tpp_*, rt_tpp.cu:21-294) are bit-exact in the rate, the mean pair
energy, the pair-energy CDF, the sampled pair energies and the
interpolated rates."""
import ctypes as C

import numpy as np
import pytest
import torch

import aperture3_amd as apt
import oracle_ctypes as oc
from ic_cases import SPECTRA, THOMSON

pytestmark = pytest.mark.gpu


def pair(name, ng=40, nep=40, n0=None):
    kind, p, emin, emax, n0_ = SPECTRA[name]
    n0 = n0_ if n0 is None else n0
    ic = apt.inverse_compton(apt.SimParams(n_gamma=ng, n_ep=nep))
    spec = [apt.Spectra.power_law_hard, apt.Spectra.power_law_soft,
            apt.Spectra.broken_power_law, apt.Spectra.black_body,
            apt.Spectra.mono_energetic][kind](*p)
    ic.init(spec, emin, emax, n0)
    o = oc.ICOracle(ng, nep)
    o.init(kind, p, emin, emax, n0)
    return ic, o


def same(a, b, what):
    assert a.shape == b.shape, what
    assert np.array_equal(a, b, equal_nan=True), (what, np.nanmax(np.abs(a - b)))


@pytest.mark.parametrize("name", sorted(SPECTRA))
def test_tables_bit_exact(name):
    print(f"ic tables: {name}", flush=True)
    ic, o = pair(name)
    same(ic.gammas(), o.gammas, "gammas")
    same(ic.ep(), o.ep, "ep")
    same(ic.log_ep(), o.log_ep, "log_ep")
    same(ic.ic_rate(), o.ic_rate, "ic_rate")
    same(ic.gg_rate(), o.gg_rate, "gg_rate")
    same(ic.dNde(), o.dnde, "dNde")
    same(ic.dNde_thomson(), o.dnde_th, "dNde_thomson")
    assert np.isfinite(o.ic_rate).all() and (o.ic_rate > 0).any()


def test_ragged_grid_and_sampling_bit_exact():
    print("ic: 23 x 37 grid, samples and rates", flush=True)
    ic, o = pair("soft_pl", ng=23, nep=37)
    same(ic.dNde(), o.dnde, "dNde")
    same(ic.dNde_thomson(), o.dnde_th, "dNde_thomson")
    rng = np.random.default_rng(4)
    g = np.concatenate([np.exp(rng.uniform(np.log(0.5), np.log(1e16), 5000)),
                        ic.gammas(), [1.01, 2.0, 1e15, 2e15, 1.0]])
    same(ic.gen_photon_e(g, seed=77, epoch=5), o.gen_photon_e(g, seed=77, epoch=5), "eph")
    a, b = ic.rates(g), o.rates(g)
    same(a[0], b[0], "ic rate")
    same(a[1], b[1], "gg rate")


def test_device_entry_points_bit_exact():
    """generate_photon_energies (accept step), gen_photon_e and
    generate_random_gamma on device tensors, n0 chosen so that about half
    of the samples scatter."""
    print("ic: device tensors", flush=True)
    ic, o = pair("soft_pl", n0=0.5 / THOMSON)
    n = 30001
    gam = torch.empty(n, dtype=torch.float64, device="cuda")
    ic.generate_random_gamma(gam, seed=9, epoch=1)
    gh = gam.cpu().numpy()
    same(gh, o.random_gamma(n, seed=9, epoch=1), "random gamma")
    eph = torch.empty_like(gam)
    ic.generate_photon_energies(eph, gam, seed=9, epoch=2)
    expect = o.gen_photon_e(gh, seed=9, epoch=2, accept=True)
    same(eph.cpu().numpy(), expect, "generate_photon_energies")
    assert 0.05 < (expect > 0).mean() < 0.95
    ic.gen_photon_e_device(eph, gam, seed=9, epoch=3)
    same(eph.cpu().numpy(), o.gen_photon_e(gh, seed=9, epoch=3), "gen_photon_e device")


def test_reference_case_gamma_10():
    """tests/test_ic.cu:24-69 (which prints the energies): 100 electrons at
    gamma = 10 scattering the broken power law; the reference tables are
    600 x 600 (sim_params.h:87-88)."""
    print("ic: reference case 600 x 600", flush=True)
    kind, p, emin, emax, n0 = SPECTRA["test_ic_cu"]
    ic = apt.inverse_compton(apt.SimParams())
    assert (ic.n_gamma, ic.n_ep) == (600, 600)
    ic.init(apt.Spectra.broken_power_law(*p), emin, emax, n0)
    e = ic.gen_photon_e(np.full(100, 10.0), seed=1234)
    assert np.all(np.isfinite(e)) and np.all(e >= 0) and np.all(e <= 10.0 - 1.01)
    cdf = ic.dNde()
    rows = ~np.isnan(cdf).any(axis=1)
    assert np.all(cdf[rows, -1] == 1.0)


def test_errors():
    lib = apt.load_library()
    h = C.c_void_p()
    assert lib.apt_ic_create(2, 40, C.byref(h)) == 1
    ic = apt.inverse_compton(apt.SimParams(n_gamma=8, n_ep=8))
    with pytest.raises(apt.AptError, match="apt_ic_init has not run"):
        ic.ic_rate()
    with pytest.raises(apt.AptError, match="apt_ic_init has not run"):
        ic.gen_photon_e(np.ones(4))
    with pytest.raises(apt.AptError, match="bad spectrum"):
        ic.init(apt.Spectra.mono_energetic(1e-5, 1e-6), 1e-3, 1e-7)


def tpp_pair(name, ng=40):
    kind, p, emin, emax, n0 = SPECTRA[name]
    tp = apt.triplet_pairs(apt.SimParams(n_gamma=ng))
    spec = [apt.Spectra.power_law_hard, apt.Spectra.power_law_soft,
            apt.Spectra.broken_power_law, apt.Spectra.black_body,
            apt.Spectra.mono_energetic][kind](*p)
    tp.init(spec, emin, emax, n0)
    o = oc.TPPOracle(ng)
    o.init(kind, p, emin, emax, n0)
    return tp, o


@pytest.mark.parametrize("name", sorted(SPECTRA))
def test_triplet_tables_bit_exact(name):
    """rt_tpp.cu:188-275 (NaN rows where a soft photon has s < 4 included)."""
    print(f"tpp tables: {name}", flush=True)
    tp, o = tpp_pair(name)
    same(tp.gammas(), o.gammas, "gammas")
    same(tp.rate(), o.rate, "rate")
    same(tp.Em(), o.em, "Em")
    same(tp.dNde(), o.dnde, "dNde")


def test_triplet_sampling_and_rates_bit_exact():
    print("tpp: 29-point grid, samples and rates", flush=True)
    tp, o = tpp_pair("mono", ng=29)
    same(tp.dNde(), o.dnde, "dNde")
    rng = np.random.default_rng(8)
    g = np.concatenate([np.exp(rng.uniform(np.log(0.5), np.log(1e16), 5000)),
                        tp.gammas(), [1.01, 1e15, 2e15]])
    same(tp.gen_tpp_Ep(g, seed=21, epoch=6), o.gen_ep(g, seed=21, epoch=6), "E_p")
    a, b = tp.rates(g), o.rates(g)
    same(a[0], b[0], "rate")
    same(a[1], b[1], "Em")
    tp2 = apt.triplet_pairs(apt.SimParams(n_gamma=8))
    with pytest.raises(apt.AptError, match="apt_tpp_init has not run"):
        tp2.rate()


def test_cpp_ic_driver():
    """The C++ host mirror (ic_driver: tests/test_ic.cu's scenario through
    inverse_compton / triplet_pairs / Spectra) prints the same photon
    energies and triplet rates as the oracle, bit for bit."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(apt.LIB_PATH), "ic_driver")
    print("ic_driver 40 40 1234", flush=True)
    out = subprocess.run([exe, "40", "40", "1234"], capture_output=True, text=True,
                         timeout=120, check=True).stdout.split()
    vals = np.array([float(v) for v in out])
    assert vals.size == 140
    kind, p, emin, emax, n0 = SPECTRA["test_ic_cu"]
    o = oc.ICOracle(40, 40)
    o.init(kind, p, emin, emax, n0)
    same(vals[:100], o.gen_photon_e(np.full(100, 10.0), seed=1234), "ic_driver energies")
    t = oc.TPPOracle(40)
    t.init(kind, p, emin, emax, n0)
    same(vals[100:], t.rate, "ic_driver triplet rates")


@pytest.mark.parametrize("kind", ["power_law", "dummy"])
def test_emission_models_bit_exact(kind):
    """InverseComptonPL1D / InverseComptonDummy (apt_ic_model_emit) against
    the oracle: emission flags, signed photon energies and free paths bit
    for bit over Lorentz factors spanning both f_inv regimes, both signs
    of p, two epochs."""
    print(f"ic model: {kind}", flush=True)
    p = apt.SimParams(delta_t=0.3)
    m = (apt.InverseComptonPL1D if kind == "power_law" else apt.InverseComptonDummy)(p)
    o = oc.ICModelOracle(0 if kind == "dummy" else 1, alpha=p.spectral_alpha, e_s=p.e_s,
                         e_min=p.e_min, photon_path=p.photon_path,
                         ic_rate=0.3 / p.ic_path, E_cutoff=p.E_cutoff, E_ph=p.E_ph)
    rng = np.random.default_rng(12)
    n = 50000
    pp = np.exp(rng.uniform(np.log(0.01), np.log(1e5), n)) * rng.choice([-1.0, 1.0], n)
    g = np.sqrt(1 + pp * pp)
    gd = torch.tensor(g, device="cuda")
    pd = torch.tensor(pp, device="cuda")
    for epoch in (0, 7):
        e, eph, path = m.emit(gd, pd, seed=5, epoch=epoch)
        oe, oeph, opath = o.emit(g, pp, seed=5, epoch=epoch)
        same(e.cpu().numpy(), oe, "emit")
        same(eph.cpu().numpy(), oeph, "eph")
        same(path.cpu().numpy(), opath, "path")
        assert 0 < oe.sum() < n
