"""GPU parity of the 1D GR gap (APT_COORD_1DGR; Grid_1dGR::compute_coef,
grid_1dgr.cpp:21-175; ptc_updater_1dgr.cu:16-503; field_solver_1dgr.cu:
14-58) against the oracle: FP64 bit-exact in the coefficient arrays, the
initial condition, and every particle, photon, field, current and charge
value after steps with smoothing, tracked output steps and photons."""
import numpy as np
import pytest

import aperture3_amd as apt
import oracle_ctypes as oc
from gr1d_cases import GUARD, LOWER, N, SIZE, profile

pytestmark = pytest.mark.gpu

MULT = 4
CAP = N * 2 * MULT + 256


def make(smoothing=1):
    kw = dict(a=0.99, omega=0.2, B0=10.0)
    p = apt.SimParams(N=[N, 1, 1], guard=[GUARD, 0, 0], lower=[LOWER, 0.0, 0.0],
                      size=[SIZE, 1.0, 1.0], coord_system="1DGR", precision=apt.APT_PREC_FP64,
                      max_ptc_number=CAP, max_photon_number=64, track_percent=0.3,
                      current_smoothing=smoothing, data_interval=3, q_e=1.0, **kw)
    env = apt.sim_environment(p)
    d = apt.sim_data(env)
    apt.Grid_1dGR.compute_coef(d, profile())
    o = oc.GR1dOracle(N, GUARD, LOWER, SIZE, profile(), track=0.3, capacity=CAP,
                      data_interval=3, smoothing=smoothing, seed=p.random_seed, **kw)
    return env, d, o


def test_coefficients_and_initial_condition():
    print("1dgr: coefficients + initial condition", flush=True)
    env, d, o = make()
    for k, name in enumerate(apt.GR1D_COEF_NAMES):
        assert np.array_equal(apt.Grid_1dGR.coefficient(d, name), o.gcoef[k]), name
    apt.ptc_updater_1dgr(env).prepare_initial_condition(d, MULT)
    assert o.initial(MULT) == d.num_particles()
    pg, po = d.get_particles(), o.particles()
    for k in ("cell", "x1", "p1", "E", "w", "flag", "id"):
        assert np.array_equal(pg[k], po[k]), k
    assert np.any(pg["flag"] & 2)
    d.close()


@pytest.mark.parametrize("smoothing", [0, 2])
def test_steps_bit_exact(smoothing):
    print("1dgr: 9 steps, smoothing", smoothing, flush=True)
    env, d, o = make(smoothing)
    apt.ptc_updater_1dgr(env).prepare_initial_condition(d, MULT)
    o.initial(MULT)
    rng = np.random.default_rng(5)
    pg = d.get_particles()
    live = pg["cell"] != apt.MAX_CELL
    p1 = np.where(live, rng.normal(0, 300.0, len(live)), 0.0)
    pg["p1"] = p1
    o.ptc["p1"][: o.number] = p1
    d.set_particles(pg)
    e0 = rng.normal(0, 1e-3, env.ncell())
    d.set_field(apt.FIELD_E, 0, e0)
    o.f["E"][0][:] = e0
    m = 12
    c = rng.integers(GUARD + 5, N + GUARD - 5, m)
    ph = {"cell": c.astype(np.uint32), "x1": rng.random(m), "x2": np.zeros(m),
          "x3": np.zeros(m), "p1": rng.normal(0, 20.0, m), "p2": np.zeros(m),
          "p3": rng.normal(0, 0.5, m), "w": np.ones(m), "path": np.ones(m),
          "flag": np.full(m, 2, np.uint32)}
    d.set_photons(ph)
    o.set_rt(max_photon=64)
    for k in ("cell", "x1", "x2", "x3", "p1", "p2", "p3", "w", "path", "flag"):
        o.ph[k][:m] = ph[k]
    o.ph_number = m
    o.sp.jexp = d.jexp()
    pu, fs = apt.ptc_updater_1dgr(env), apt.field_solver_1dgr(env)
    dt = 1e-3
    for step in range(9):
        pu.update_particles(d, dt, step)
        fs.update_fields(d, dt)
        o.ptc_update(dt, step)
        o.field_update(dt)
    pg, po = d.get_particles(), o.particles()
    for k in ("cell", "x1", "p1", "p2", "E", "w", "flag", "id"):
        assert np.array_equal(pg[k], po[k]), k
    assert np.any(pg["p2"] != 0)  # u_0 at infinity of tracked particles
    hg, ho = d.get_photons(), o.photons()
    for k in ("cell", "x1", "p1", "p2", "E"):
        assert np.array_equal(hg[k], ho[k]), ("photon", k)
    for comp in range(3):
        assert np.array_equal(d.get_field(apt.FIELD_E, comp), o.f["E"][comp]), ("E", comp)
    assert np.array_equal(d.get_field(apt.FIELD_J, 0), o.f["J"][0])
    for s in range(2):
        assert np.array_equal(d.get_field(apt.FIELD_RHO, s), o.rho[s]), ("rho", s)
    assert np.abs(o.f["J"][0]).max() > 0
    d.close()
