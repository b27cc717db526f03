"""Pins the oracle's 3D log-spherical force-free / vacuum solver
(oracle/ffe_oracle.c, restating ffe_solver_logsph.cu):

  * SSP-RK3 composition: away from the boundaries the three stages equal
    the cubic stability polynomial (1 + L + L^2/2 + L^3/6) u of the
    single-stage operator L (dE, dB of rk_push), built from independent
    rk_push calls -- pins the stage coefficients and the en/e/de algebra;
  * a vacuum static dipole (E = 0, B = B0) is left exactly unchanged;
  * the absorbing layer's profile 1 - coef (i/L)^2;
  * clean_epar leaves E.B = 0 and is idempotent for uniform fields;
  * check_eGTb enforces |E| <= |B| and leaves |E| <= |B| fields alone.
"""
import math

import numpy as np
import pytest

import oracle_ctypes as oc

SIZE = (math.log(30.0), math.pi, 2.0 * math.pi)


def make(N=(20, 18, 16), **kw):
    return oc.FFEOracle(N, (3, 3, 3), SIZE, **kw)


def interior(o, a, pad=6):
    v = a.reshape(o.dims[2], o.dims[1], o.dims[0])
    return v[pad:-pad, pad:-pad, pad:-pad]


def test_rk3_is_the_cubic_polynomial():
    o = make()
    rng = np.random.default_rng(1)
    n = o.ncell
    B0 = [rng.normal(size=n) for _ in range(3)]
    E = [rng.normal(size=n) for _ in range(3)]
    dB0 = [rng.normal(size=n) for _ in range(3)]
    for c in range(3):
        o.f["Bbg"][c][:] = B0[c]

    def L(e, b):  # state (E, B - B0) -> (dE, dB)
        for c in range(3):
            o.f["E"][c][:] = e[c]
            o.f["B"][c][:] = B0[c] + b[c]
        return o.rk_push(0.01)

    u = (E, dB0)
    l1 = L(*u)
    l2 = L(*l1)
    l3 = L(*l2)
    expect_E = [u[0][c] + l1[0][c] + l2[0][c] / 2 + l3[0][c] / 6 for c in range(3)]
    expect_b = [u[1][c] + l1[1][c] + l2[1][c] / 2 + l3[1][c] / 6 for c in range(3)]
    for c in range(3):
        o.f["E"][c][:] = E[c]
        o.f["B"][c][:] = B0[c] + dB0[c]
    o.update(0.01, omega=0.0)
    for c in range(3):
        np.testing.assert_allclose(interior(o, o.f["E"][c]), interior(o, expect_E[c]),
                                   rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(interior(o, o.f["B"][c] - B0[c]), interior(o, expect_b[c]),
                                   rtol=1e-12, atol=1e-12)


def test_static_dipole_unchanged():
    """E = 0, B = B0 = the star's dipole: dE = dB = 0, so the RK3 stages only
    re-round B (c1 Bn + c2 B is B to an ulp) and E stays zero."""
    o = make()
    d0, d1, d2 = o.dims
    g = 3
    for c, st in ((0, 0b001), (1, 0b010)):
        x1 = o.mesh.delta[0] * (np.arange(d0) - g + (1.0 if st & 1 else 0.5))
        x2 = o.mesh.delta[1] * (np.arange(d1) - g + (1.0 if st & 2 else 0.5))
        T, R = np.meshgrid(x2, np.exp(x1), indexing="ij")
        f = 2e3 * np.cos(T) / R ** 3 if c == 0 else 1e3 * np.sin(T) / R ** 3
        v = np.broadcast_to(f, (d2, d1, d0)).ravel()
        o.f["B"][c][:] = v
        o.f["Bbg"][c][:] = v
    B = [b.copy() for b in o.f["B"]]
    for _ in range(3):
        o.update(0.005, omega=0.0)
    for c in range(3):
        np.testing.assert_allclose(interior(o, o.f["B"][c], 4), interior(o, B[c], 4),
                                   rtol=1e-14, atol=0)
        assert np.abs(interior(o, o.f["E"][c], 4)).max() <= 1e-12


def test_damping_profile():
    o = make(damping_length=5, damping_coef=0.1)
    for c in range(3):
        o.f["E"][c][:] = 1.0
    o.boundary(0.0)
    v = o.f["E"][0].reshape(o.dims[2], o.dims[1], o.dims[0])[5, 8, -5:]
    np.testing.assert_allclose(v, 1.0 - 0.1 * (np.arange(5) / 5.0) ** 2, rtol=1e-15)


def test_clean_epar_uniform():
    o = make()
    e, b = np.array([0.3, -0.2, 0.5]), np.array([1.0, 2.0, -0.5])
    for c in range(3):
        o.f["E"][c][:] = e[c]
        o.f["B"][c][:] = b[c]
    o.clean_epar()
    o.check_egtb()  # |E| < |B|: a pure copy of the cleaned field back to E
    got = np.array([interior(o, o.f["E"][c])[0, 0, 0] for c in range(3)])
    expect = e - e.dot(b) * b / (b.dot(b) + float(np.float32(1e-7)))
    np.testing.assert_allclose(got, expect, rtol=1e-14)
    # what remains of E.B is the TINY regulariser's share: (E.B) TINY/(B^2+TINY)
    assert abs(got.dot(b)) <= 1e-7 * abs(e.dot(b))
    o.clean_epar()
    o.check_egtb()
    again = np.array([interior(o, o.f["E"][c])[0, 0, 0] for c in range(3)])
    np.testing.assert_allclose(again, got, rtol=1e-7)


def test_check_egtb_uniform():
    o = make()
    e, b = np.array([3.0, 4.0, 0.0]), np.array([0.0, 0.0, 2.0])
    for c in range(3):
        o.f["E"][c][:] = e[c]
        o.f["B"][c][:] = b[c]
    o.dE[:] = np.concatenate([np.full(o.ncell, e[c]) for c in range(3)])
    o.check_egtb()
    got = np.array([interior(o, o.f["E"][c])[0, 0, 0] for c in range(3)])
    tiny = float(np.float32(1e-7))
    scale = math.sqrt((4.0 + tiny) / (25.0 + tiny))
    np.testing.assert_allclose(got, e * scale, rtol=1e-14)
    assert np.linalg.norm(got) <= 2.0 * (1.0 + 1e-7)  # |B| up to the TINY regulariser
