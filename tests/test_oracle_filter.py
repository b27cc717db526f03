"""Pinning of the oracle's current filters against the binomial filter they
restate (filter_current_cart_{1,2,3}d, ptc_updater.cu:501-625;
filter_current_logsph, ptc_updater_logsph.cu:301-331).

The binomial filter is the separable [1, 2, 1] / 4 kernel along every axis:
its response to a unit impulse is the outer product of (1/4, 1/2, 1/4)
over the dimensions, and it maps a constant to itself.  The reference's 3D
kernel writes 0.03152f for the 1/32 weight of the eight cells that are
off-axis in one plane of z and in-plane in x or y; the oracle keeps that
weight unless asked to fix it.  The log-spherical filter is the 2D binomial
of J A divided by A, with a reflecting edge at the bulk boundaries."""
import math

import numpy as np
import pytest

import oracle_ctypes as oc

K1 = np.array([0.25, 0.5, 0.25])


def impulse_response(dim, fix3d):
    N = [8] * dim
    o = oc.OracleSim(N, [2] * dim, periodic=(True, True, True))
    d = o.dims
    j = np.zeros(o.ncell)
    c = [2 + 4 if a < dim else 0 for a in range(3)]
    j[c[0] + c[1] * d[0] + c[2] * d[0] * d[1]] = 1.0
    o.filter_current(j, fix3d=fix3d)
    J = j.reshape(d[2], d[1], d[0])
    if dim == 1:
        return J[0, 0, c[0] - 2: c[0] + 3]
    if dim == 2:
        return J[0, c[1] - 2: c[1] + 3, c[0] - 2: c[0] + 3]
    return J[c[2] - 2: c[2] + 3, c[1] - 2: c[1] + 3, c[0] - 2: c[0] + 3]


def binomial(dim):
    k = np.pad(K1, 1)
    out = k
    for _ in range(dim - 1):
        out = np.multiply.outer(k, out)
    return out


@pytest.mark.parametrize("dim", [1, 2])
def test_cartesian_filter_is_binomial(dim):
    assert np.array_equal(impulse_response(dim, False), binomial(dim))


def test_cartesian_filter_3d_binomial_and_reference_weight():
    exact = binomial(3)
    assert np.array_equal(impulse_response(3, True), exact)
    ref = impulse_response(3, False)
    diff = np.argwhere(ref != exact)
    # exactly the eight cells at dz = +-1 with one of dx, dy = +-1 and the
    # other 0, which carry 0.03152f instead of 1/32
    assert len(diff) == 8
    for kz, jy, ix in diff:
        dz, dy, dx = kz - 2, jy - 2, ix - 2
        assert abs(dz) == 1 and abs(dx) + abs(dy) == 1
        assert ref[kz, jy, ix] == float(np.float32(0.03152))


def test_cartesian_filter_keeps_constant():
    o = oc.OracleSim([12, 10], [2, 2], periodic=(True, True, True))
    j = np.full(o.ncell, 0.7)
    o.filter_current(j)
    assert np.allclose(j, 0.7, rtol=0, atol=1e-15)


def logsph_setup():
    o = oc.LogSphOracle((16, 12, 1), (2, 2, 0), (math.log(30.0), math.pi, 1.0))
    A = o.coefficient("A2_e")
    return o, A


def test_logsph_filter_keeps_j_times_area_constant():
    """J = c / A is a fixed point: the binomial of a constant is the
    constant, in the interior and at the reflecting bulk edges."""
    o, A = logsph_setup()
    d = o.dims
    bulk = np.zeros((d[1], d[0]), bool)
    bulk[2:d[1] - 2, 2:d[0] - 2] = True
    bulk = bulk.ravel() & (A != 0)
    j = np.zeros(o.ncell)
    j[bulk] = 3.0 / A[bulk]
    before = j.copy()
    o.filter(j, np.where(A != 0, A, 1.0))
    assert np.allclose(j[bulk], before[bulk], rtol=1e-13)


def test_logsph_filter_impulse():
    """A unit of J A at one interior cell spreads as the 2D binomial
    kernel, divided by the area of each receiving cell."""
    o, A = logsph_setup()
    d = o.dims
    A = np.where(A != 0, A, 1.0)
    c1, c2 = 8, 6
    p = c1 + c2 * d[0]
    j = np.zeros(o.ncell)
    j[p] = 1.0 / A[p]
    o.filter(j, A)
    JA = (j * A).reshape(d[1], d[0])[c2 - 2: c2 + 3, c1 - 2: c1 + 3]
    assert np.allclose(JA, binomial(2), rtol=1e-14, atol=1e-16)
