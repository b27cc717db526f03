"""Oracle checks of the move + Esirkepov deposit (ptc_updater.cu:220-499).

Pinned by: discrete continuity at round-off for random particles in 1D, 2D
and 3D, including every crossing pattern dc in {-1,0,1}^d (rho after minus
rho before equals minus the discrete divergence of J); the total flux
sum_faces J1 = q w dx/Delta of a single particle; the out-of-plane current
sum J = q w v; and periodic wrap of cells and current."""
import itertools

import numpy as np
import pytest

import oracle_ctypes as oc


def make(dim, N=12):
    return oc.OracleSim([N] * dim, [2] * dim, capacity=4096, data_interval=1)


def random_particles(o, n, rng, umax=3.0):
    m = o.mesh
    dim = m.dim
    g = [m.guard[d] for d in range(3)]
    N = [m.dims[d] - 2 * g[d] for d in range(3)]
    c = [rng.integers(g[d], g[d] + N[d], n) for d in range(3)]
    cell = (c[0] + c[1] * m.dims[0] + c[2] * m.dims[0] * m.dims[1]).astype(np.uint32)
    ptc = {"cell": cell,
           "x1": rng.random(n), "x2": rng.random(n) if dim > 1 else rng.normal(size=n),
           "x3": rng.random(n) if dim > 2 else rng.normal(size=n),
           "p1": rng.uniform(-umax, umax, n), "p2": rng.uniform(-umax, umax, n),
           "p3": rng.uniform(-umax, umax, n), "w": rng.uniform(0.5, 2.0, n),
           "flag": (rng.integers(0, 2, n).astype(np.uint32) << 29)}
    return ptc


def bulk(o, a):
    m = o.mesh
    v = a.reshape(m.dims[2], m.dims[1], m.dims[0])
    return v[m.guard[2]:m.dims[2] - m.guard[2], m.guard[1]:m.dims[1] - m.guard[1],
             m.guard[0]:m.dims[0] - m.guard[0]]


def div_j(o):
    """Minus the discrete divergence of J in deposit units, on the nodes:
    node k gains J(k) - J(k+1) along each in-plane axis."""
    m = o.mesh
    out = 0.0
    for d in range(m.dim):
        J = bulk(o, o.f["J"][d])  # periodic: the bulk wraps onto itself
        out = out + (J - np.roll(J, -1, axis=2 - d))
    return out


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_continuity_random(dim):
    rng = np.random.default_rng(11 + dim)
    o = make(dim)
    ptc = random_particles(o, 600, rng)
    o.set_particles(ptc)
    # rho at the starting positions: a step with dt = 0 moves nothing
    o.ptc_update(0.0, 0, deposit_rho=True)
    rho0 = [bulk(o, r.copy()) for r in o.rho[:2]]
    o.ptc_update(0.4 * o.mesh.delta[0], 1, deposit_rho=True)
    rho1 = [bulk(o, r.copy()) for r in o.rho[:2]]
    drho = (rho1[0] + rho1[1]) - (rho0[0] + rho0[1])
    dj = div_j(o)
    scale = np.abs(drho).max()
    assert scale > 0.1
    assert np.abs(drho - dj).max() < 1e-13 * max(scale, 1.0)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_every_crossing_pattern(dim):
    """One particle per dc pattern, positions chosen to cross (or not) in
    each direction; continuity must hold for each alone."""
    for dc in itertools.product((-1, 0, 1), repeat=dim):
        o = make(dim, N=8)
        x = [0.5] * 3
        u = [0.0] * 3
        for d in range(dim):
            x[d] = {-1: 0.1, 0: 0.5, 1: 0.9}[dc[d]]
            u[d] = {-1: -0.6, 0: 0.1, 1: 0.6}[dc[d]]
        cell = 5 + 5 * o.mesh.dims[0] * (dim > 1) + 5 * o.mesh.dims[0] * o.mesh.dims[1] * (dim > 2)
        o.set_particles({"cell": np.array([cell], np.uint32), "x1": [x[0]], "x2": [x[1]],
                         "x3": [x[2]], "p1": [u[0]], "p2": [u[1]], "p3": [u[2]], "w": [1.0],
                         "flag": np.array([1 << 29], np.uint32)})
        o.ptc_update(0.0, 0, deposit_rho=True)
        r0 = bulk(o, o.rho[1].copy())
        dt = 0.5 * o.mesh.delta[0]
        o.ptc_update(dt, 1, deposit_rho=True)
        r1 = bulk(o, o.rho[1].copy())
        assert abs(r0.sum() - 1.0) < 1e-14 and abs(r1.sum() - 1.0) < 1e-14
        assert np.abs((r1 - r0) - div_j(o)).max() < 1e-14, dc
        # the particle really crossed as intended
        c = int(o.particles()["cell"][0])
        idx = [c % o.mesh.dims[0], (c // o.mesh.dims[0]) % o.mesh.dims[1],
               c // (o.mesh.dims[0] * o.mesh.dims[1])]
        for d in range(dim):
            assert idx[d] - 5 == dc[d]


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_total_flux_single_particle(dim):
    o = make(dim, N=8)
    q, w, u = 1.0, 1.7, np.array([0.5, -0.3, 0.2])
    g = np.sqrt(1 + u @ u)
    cell = 4 + 4 * o.mesh.dims[0] * (dim > 1) + 4 * o.mesh.dims[0] * o.mesh.dims[1] * (dim > 2)
    o.set_particles({"cell": np.array([cell], np.uint32), "x1": [0.3], "x2": [0.6],
                     "x3": [0.45], "p1": [u[0]], "p2": [u[1]], "p3": [u[2]], "w": [w],
                     "flag": np.array([1 << 29], np.uint32)})
    dt = 0.4 * o.mesh.delta[0]
    o.ptc_update(dt, 0)
    v = u / g
    for d in range(3):
        total = o.f["J"][d].sum()
        if d < dim:
            expect = q * w * v[d] * dt / o.mesh.delta[d]
        else:
            expect = q * w * v[d]  # out-of-plane current: q w v (no dx/dt)
        assert abs(total - expect) < 1e-14, d


def test_periodic_wrap_of_cell_and_current():
    o = make(1, N=8)
    # particle in the last bulk cell moving right wraps to the first
    o.set_particles({"cell": np.array([9], np.uint32), "x1": [0.95], "x2": [0.0], "x3": [0.0],
                     "p1": [2.0], "p2": [0.0], "p3": [0.0], "w": [1.0],
                     "flag": np.array([1 << 29], np.uint32)})
    o.ptc_update(0.5 * o.mesh.delta[0], 0)
    p = o.particles()
    assert p["cell"][0] == 2
    # guard cells of J stay empty: everything was folded into the bulk
    for d in range(3):
        J = o.f["J"][d]
        assert np.all(J[:2] == 0) and np.all(J[-2:] == 0)


def test_ignore_current_flag():
    o = make(2, N=8)
    o.set_particles({"cell": np.array([4 + 4 * 12], np.uint32), "x1": [0.5], "x2": [0.5],
                     "x3": [0.0], "p1": [0.5], "p2": [0.5], "p3": [0.5], "w": [1.0],
                     "flag": np.array([(1 << 29) | (1 << 3)], np.uint32)})
    o.ptc_update(0.01, 0)
    for d in range(3):
        assert not o.f["J"][d].any()


def test_fixed_point_split_roundtrip():
    lib = oc.load()
    import ctypes as C
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.normal(size=2000), rng.normal(size=200) * 1e-9,
                           [0.0, 1.0, -1.0, 2 ** -40, -(2 ** -30)]])
    hi = C.c_int64()
    lo = C.c_int64()
    for v in vals:
        lib.orc_fx_split(C.c_double(v), C.c_int(22), C.byref(hi), C.byref(lo))
        back = lib.orc_fx_value(hi.value, lo.value, 22)
        assert abs(back - v) <= 2.0 ** -63
