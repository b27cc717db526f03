"""GPU parity of the per-cell moments (apt_collect_diagnostics;
AdditionalDiagnostics::collect_diagnostics, cuda/core/additional_diagnostics.cu:
14-100) against the oracle.  Summed weights are exact (dyadic weights);
the summed w gamma is added by device atomics in any order, as in the
reference, so the mean gamma matches the oracle's slot-order sum to 1e-13
(FP64) and to FP32 rounding (mixed)."""
import numpy as np
import pytest

import aperture3_amd as apt
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [apt.APT_PREC_FP64, apt.APT_PREC_MIXED])
def test_moments_match_oracle(prec):
    print("diagnostics: 2D 24x20, 3 species, photons, precision", prec, flush=True)
    N, g = (24, 20), 2
    o = oc.OracleSim(N, (g, g), capacity=20000, num_species=3)
    p = apt.SimParams(N=[N[0], N[1], 1], guard=[g, g, 0], max_ptc_number=20000,
                      num_species=3, precision=prec, max_photon_number=4000)
    env = apt.sim_environment(p)
    d = apt.sim_data(env)
    rng = np.random.default_rng(11)
    n = 12000
    cell = (rng.integers(g, N[0] + g, n) + rng.integers(g, N[1] + g, n) * o.dims[0])
    cell[rng.random(n) < 0.05] = apt.MAX_CELL
    ptc = {"cell": cell.astype(np.uint32), "x1": rng.random(n), "x2": rng.random(n),
           "x3": np.zeros(n), "p1": rng.normal(0, 3, n), "p2": rng.normal(0, 3, n),
           "p3": rng.normal(0, 3, n), "w": rng.choice([0.5, 1.0, 2.0], n),
           "flag": (rng.integers(0, 3, n) << 29).astype(np.uint32)}
    if prec == apt.APT_PREC_MIXED:  # values the FP32 store holds exactly
        for k in ("x1", "x2", "p1", "p2", "p3"):
            ptc[k] = ptc[k].astype(np.float32).astype(np.float64)
    o.set_particles(ptc)
    d.set_particles(ptc)
    m = 3000
    ph = {"cell": (rng.integers(g, N[0] + g, m) + rng.integers(g, N[1] + g, m)
                   * o.dims[0]).astype(np.uint32),
          "x1": rng.random(m), "x2": rng.random(m), "x3": np.zeros(m),
          "p1": rng.normal(0, 5, m), "p2": np.zeros(m), "p3": np.zeros(m),
          "w": rng.choice([0.25, 1.0], m), "path": rng.random(m)}
    d.set_photons(ph)
    o.set_rt(max_photon=4000)
    for k in ("cell", "x1", "x2", "x3", "p1", "p2", "p3", "w", "path"):
        o.ph[k][:m] = ph[k]
    o.ph_number = m
    apt.AdditionalDiagnostics(env).collect_diagnostics(d)
    num, gam, phn = o.collect_diagnostics()
    rtol = 1e-13 if prec == apt.APT_PREC_FP64 else 2e-6
    for s in range(3):
        assert np.array_equal(d.get_field(apt.FIELD_PTC_NUM, s), num[s]), s
        np.testing.assert_allclose(d.get_field(apt.FIELD_GAMMA, s), gam[s], rtol=rtol,
                                   atol=0, err_msg=str(s))
    assert np.array_equal(d.get_field(apt.FIELD_PHOTON_NUM, 0), phn)
    assert num[0].sum() > 0 and phn.sum() == ph["w"].sum()
    d.close()
