"""Step-level invariants of the oracle's full Cartesian step (push, deposit,
Yee update; ptc_updater.cu:748-841 + field_solver.cu:334-358).

Gauss's law: with J in deposit units (charge x cells moved, no dt or
Delta factor) and E -= dt J, div E - (dt/Delta) rho is conserved exactly by
the discrete scheme after the step-0 half step (which applies 1.5 dt J
once).  Its drift is the round-off of the run.  div B stays at round-off."""
import numpy as np
import pytest

from cases import Case


def bulk(c, a):
    v = a.reshape(c.dims[2], c.dims[1], c.dims[0])
    g = c.guard
    return v[g[2]:c.dims[2] - g[2], g[1]:c.dims[1] - g[1], g[0]:c.dims[0] - g[0]]


@pytest.mark.parametrize("N,ppc", [((256,), 20), ((24, 24), 4), ((8, 8, 8), 2)])
def test_gauss_residual_drift(N, ppc):
    c = Case(N, ppc, theta=(0.1, 0.1), data_interval=1,
             Bbg=(1.0, 0.0, 0.0) if len(N) == 1 else (0.0, 0.0, 1.0))
    o = c.oracle()
    kappa = c.dt / c.delta[0]
    res = []
    for step in range(30):
        o.ptc_update(c.dt, step)
        o.field_update(c.dt, step)
        rho = o.rho[0] + o.rho[1]
        res.append(bulk(c, o.divE - kappa * rho).copy())
    scale = np.abs(bulk(c, o.divE)).max()
    drift = max(np.abs(r - res[1]).max() for r in res[1:])
    assert drift < 1e-12 * max(scale, 1.0), drift
    assert np.abs(bulk(c, o.divB)).max() < 1e-10


def test_c1_energy_sane():
    """The 1D gap case for 50 steps: particles stay, energy bounded."""
    c = Case((512,), 10)
    o = c.oracle()
    n0 = o.number
    g0 = o.particles()["E"].sum()
    for step in range(50):
        o.ptc_update(c.dt, step)
        o.field_update(c.dt, step)
    p = o.particles()
    assert o.number == n0
    assert np.all((p["x1"] >= 0) & (p["x1"] < 1))
    assert np.all((p["cell"] >= 2) & (p["cell"] < 514))
    assert abs(p["E"].sum() / g0 - 1) < 1e-2
