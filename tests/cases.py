"""Shared set-ups for the parity tests: the 1D Cartesian gap case (4096
cells, e-/e+ pairs, B_x = 1, E_x = 1e-3 sin(2 pi 4 x), dt = dx/2) and
smaller 2D/3D boxes, built identically for the oracle and the HIP library.
"""
from __future__ import annotations

import numpy as np

import oracle_ctypes as oc


def stagger_positions(dims, guard, delta, lower, st, dim):
    """Quadmesh::pos (quadmesh.h:119-150) along each axis for stagger st."""
    out = []
    for d in range(3):
        s = (st >> d) & 1
        in_cell = s * 0.5 + 0.5
        n = np.arange(dims[d])
        if d < dim:
            out.append(lower[d] + delta[d] * (n - guard[d] + in_cell))
        else:
            out.append(np.full(dims[d], in_cell * delta[d]))
    return out


def field_from(fn, dims, guard, delta, lower, st, dim):
    x1, x2, x3 = stagger_positions(dims, guard, delta, lower, st, dim)
    Z, Y, X = np.meshgrid(x3, x2, x1, indexing="ij")
    return np.ascontiguousarray(np.broadcast_to(fn(X, Y, Z), X.shape), dtype=np.float64).ravel()


class Case:
    """A periodic Cartesian case: grid, initial fields, loading."""

    def __init__(self, N, ppc, theta=(0.1, 0.1), drift=(0.0, 0.0), seed=43210,
                 E1=None, Bbg=(1.0, 0.0, 0.0), q_e=1.0, cfl=0.5, guard=2,
                 data_interval=100):
        self.N = list(N) + [1] * (3 - len(N))
        self.dim = len(N)
        self.guard = [guard if d < self.dim else 0 for d in range(3)]
        self.size = [1.0, 1.0, 1.0]
        self.lower = [0.0, 0.0, 0.0]
        self.dims = [self.N[d] + 2 * self.guard[d] for d in range(3)]
        self.delta = [self.size[d] / self.N[d] for d in range(3)]
        self.ncell = self.dims[0] * self.dims[1] * self.dims[2]
        self.ppc = ppc
        self.theta = theta
        self.drift = drift
        self.seed = seed
        self.q_e = q_e
        self.dt = cfl * self.delta[0]
        self.data_interval = data_interval
        self.Bbg = Bbg
        self.E1 = E1 if E1 is not None else (lambda x, y, z: 1.0e-3 * np.sin(2 * np.pi * 4 * x))
        nbulk = self.N[0] * self.N[1] * self.N[2]
        self.nptc = nbulk * ppc * 2

    def fields(self):
        st_e = (0b110, 0b101, 0b011)
        st_b = (0b001, 0b010, 0b100)
        args = (self.dims, self.guard, self.delta, self.lower)
        E = [field_from(self.E1, *args, st_e[0], self.dim), np.zeros(self.ncell), np.zeros(self.ncell)]
        B = [field_from(lambda x, y, z, v=v: v + 0 * x, *args, st_b[c], self.dim)
             for c, v in enumerate(self.Bbg)]
        return E, B

    # ---- oracle -------------------------------------------------------------
    def oracle(self, load=True, jexp=22):
        o = oc.OracleSim(self.N[: self.dim], self.guard[: self.dim], size=self.size,
                         periodic=(True, True, True), q_e=self.q_e,
                         capacity=self.nptc, jexp=jexp,
                         data_interval=self.data_interval)
        E, B = self.fields()
        for c in range(3):
            o.f["E"][c][:] = E[c]
            o.f["B"][c][:] = B[c]
            o.f["Bbg"][c][:] = B[c]
        if load:
            o.load_uniform(self.ppc, self.theta, self.drift, 1.0, self.seed)
        return o

    # ---- HIP library --------------------------------------------------------
    def gpu(self, precision=0, load=True, capacity=None):
        import aperture3_amd as apt
        p = apt.SimParams(N=self.N, guard=self.guard, size=self.size, lower=self.lower,
                          periodic_boundary=[True, True, True], q_e=self.q_e,
                          max_ptc_number=capacity or self.nptc + 64,
                          precision=precision, data_interval=self.data_interval,
                          random_seed=self.seed)
        env = apt.sim_environment(p)
        data = apt.sim_data(env)
        E, B = self.fields()
        for c in range(3):
            data.set_field(apt.FIELD_E, c, E[c])
            data.set_field(apt.FIELD_B, c, B[c])
            data.set_field(apt.FIELD_BBG, c, B[c])
        if load:
            data.fill_multiplicity(1.0, self.ppc, self.theta, self.drift, sort=False)
        return env, data
