"""Oracle checks of the Vay push (vay_push.cuh:8-30) through a one-particle
step in uniform fields, pinned by properties of the published scheme
(Vay 2008): pure E is an exact impulse, pure B conserves |u| and rotates by
2 atan(|tau|/gamma), and crossed fields at the E x B drift leave u
unchanged.  A sign error, a dropped term or a transposed cross product
fails at least one of these."""
import numpy as np
import pytest

import oracle_ctypes as oc


def one_step(u, E, B, dt=1e-3, q=1.0, species=1):
    """One particle in a uniform field on a small periodic 1D mesh."""
    o = oc.OracleSim([16], [2], q_e=q, capacity=1)
    for c in range(3):
        o.f["E"][c][:] = E[c]
        o.f["B"][c][:] = B[c]
        o.f["Bbg"][c][:] = B[c]
    o.set_particles({"cell": np.array([8], np.uint32), "x1": [0.3], "x2": [0.0], "x3": [0.0],
                     "p1": [u[0]], "p2": [u[1]], "p3": [u[2]], "w": [1.0],
                     "flag": np.array([species << 29], np.uint32)})
    o.ptc_update(dt, 1)
    p = o.particles()
    return np.array([p["p1"][0], p["p2"][0], p["p3"][0]]), p["E"][0]


def test_pure_electric_impulse():
    u0 = np.array([0.3, -0.2, 0.1])
    E = np.array([0.7, -1.1, 0.4])
    dt = 1e-3
    u1, g = one_step(u0, E, [0, 0, 0], dt=dt)
    # q/m = +1 for positrons: du = dt q E / m exactly
    assert np.allclose(u1, u0 + dt * E, rtol=0, atol=1e-15)
    assert abs(g - np.sqrt(1 + u1 @ u1)) < 1e-14


@pytest.mark.parametrize("B", [(0.0, 0.0, 2.0), (1.0, -0.5, 0.25), (0.0, 3.0, 0.0)])
def test_pure_magnetic_rotation(B):
    B = np.array(B)
    u0 = np.array([0.6, 0.2, -0.3])
    dt = 0.05
    u1, g = one_step(u0, [0, 0, 0], B, dt=dt)
    assert abs(np.linalg.norm(u1) - np.linalg.norm(u0)) < 1e-14
    g0 = np.sqrt(1 + u0 @ u0)
    tau = 0.5 * dt * B
    b = B / np.linalg.norm(B)
    # rotation angle about b, measured on the perpendicular components
    up0 = u0 - (u0 @ b) * b
    up1 = u1 - (u1 @ b) * b
    ang = np.arctan2(np.linalg.norm(np.cross(up0, up1)), up0 @ up1)
    assert abs(ang - 2 * np.arctan(np.linalg.norm(tau) / g0)) < 1e-12
    # sense of rotation: positive charge gyrates clockwise about B
    assert np.cross(up0, up1) @ b < 0
    assert abs(u1 @ b - u0 @ b) < 1e-14


def test_exb_drift_is_exact():
    B = np.array([0.0, 0.0, 1.0])
    v = np.array([0.4, -0.25, 0.0])
    E = -np.cross(v, B)
    gam = 1 / np.sqrt(1 - v @ v)
    u0 = gam * v
    u1, _ = one_step(u0, E, B, dt=0.1)
    assert np.allclose(u1, u0, rtol=0, atol=5e-15)


def test_electron_opposite_sign():
    u0 = np.array([0.0, 0.0, 0.0])
    E = np.array([1.0, 0.0, 0.0])
    u1, _ = one_step(u0, E, [0, 0, 0], dt=0.01, species=0)
    assert np.allclose(u1, [-0.01, 0, 0], atol=1e-16)


def test_ignore_em_flag_keeps_momentum():
    o = oc.OracleSim([16], [2], capacity=1)
    o.f["E"][0][:] = 5.0
    o.set_particles({"cell": np.array([8], np.uint32), "x1": [0.5], "x2": [0.0], "x3": [0.0],
                     "p1": [0.1], "p2": [0.0], "p3": [0.0], "w": [1.0],
                     "flag": np.array([(1 << 29) | (1 << 4)], np.uint32)})
    o.ptc_update(0.01, 0)
    assert o.particles()["p1"][0] == 0.1
