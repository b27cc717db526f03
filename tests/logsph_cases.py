"""Shared set-up of the 2D log-spherical parity cases: a (ln r, theta) grid
over r in [1, 30], theta in [0, pi], a dipole background field, random
particles and injection, built identically for the oracle and the HIP
library (the pattern of labs/pulsar/src/main.cpp:45-62, 113-127)."""
from __future__ import annotations

import math

import numpy as np

import oracle_ctypes as oc

LNR = math.log(30.0)
FLAG_IGNORE_EM = 1 << 4


def pos(dims_guard_delta, d, n, x):
    _, guard, delta = dims_guard_delta
    return delta[d] * ((n - guard[d]) + x)


class LogSphCase:
    def __init__(self, N=(64, 64), guard=3, mu=1.0e2, nptc=20000, seed=9, cap=None,
                 damping_length=12, damping_coef=0.01, data_interval=3,
                 current_smoothing=1, omega=0.1, gravity=0.0, rad_cooling=0,
                 q_e=1.0, dt=None):
        self.N = list(N)
        self.guard = [guard, guard, 0]
        self.size = [LNR, math.pi, 1.0]
        self.dims = [N[0] + 2 * guard, N[1] + 2 * guard, 1]
        self.delta = [self.size[0] / N[0], self.size[1] / N[1], 1.0]
        self.ncell = self.dims[0] * self.dims[1]
        self.mu = mu
        self.nptc = nptc
        self.seed = seed
        self.cap = cap or (nptc + 40 * N[1] + 1024)
        self.lp = dict(compactness=0.0, omega=omega, gravity_on=1 if gravity else 0,
                       gravity=gravity, rad_cooling_on=rad_cooling,
                       rad_cooling_coef=0.1, B0=mu, damping_length=damping_length,
                       damping_coef=damping_coef, inject_ions=1, track_percent=0.05,
                       current_smoothing=current_smoothing)
        self.data_interval = data_interval
        self.q_e = q_e
        self.dt = dt if dt is not None else 0.3 * self.delta[0]

    def coord(self, d, n, st):
        return self.delta[d] * ((n - self.guard[d]) + (1.0 if st else 0.5))

    def dipole(self):
        """B1 = 2 mu cos(theta)/r^3, B2 = mu sin(theta)/r^3 at their staggers."""
        out = []
        for c, st in ((0, 0b001), (1, 0b010)):
            x1 = self.coord(0, np.arange(self.dims[0]), st & 1)
            x2 = self.coord(1, np.arange(self.dims[1]), (st >> 1) & 1)
            T, R = np.meshgrid(x2, np.exp(x1), indexing="ij")
            f = 2 * self.mu * np.cos(T) / R ** 3 if c == 0 else self.mu * np.sin(T) / R ** 3
            out.append(np.ascontiguousarray(f).ravel())
        out.append(np.zeros(self.ncell))
        return out

    def perturbation(self):
        rng = np.random.default_rng(self.seed + 1)
        return [rng.normal(0, 1e-3, self.ncell) for _ in range(3)]

    def particles(self):
        rng = np.random.default_rng(self.seed)
        n, g = self.nptc, self.guard
        c1 = rng.integers(g[0], self.dims[0] - g[0] - 4, n)
        c2 = rng.integers(g[1], self.dims[1] - g[1], n)
        sp = rng.integers(0, 3, n).astype(np.uint32)
        flag = (sp << 29) | np.where(rng.random(n) < 0.05, FLAG_IGNORE_EM, 0).astype(np.uint32)
        return {"cell": (c1 + c2 * self.dims[0]).astype(np.uint32),
                "x1": rng.random(n), "x2": rng.random(n), "x3": rng.uniform(-3, 3, n),
                "p1": rng.normal(0, 2.0, n), "p2": rng.normal(0, 2.0, n),
                "p3": rng.normal(0, 2.0, n), "w": rng.uniform(0.2, 1.0, n),
                "flag": flag.astype(np.uint32), "id": np.arange(n, dtype=np.uint64)}

    # ---- oracle -------------------------------------------------------------
    def oracle(self, jexp=22):
        o = oc.LogSphOracle((self.N[0], self.N[1], 1), tuple(self.guard), tuple(self.size),
                            capacity=self.cap, jexp=jexp, num_species=3, q_e=self.q_e,
                            data_interval=self.data_interval, **self.lp)
        B = self.dipole()
        dB = self.perturbation()
        for c in range(3):
            o.f["Bbg"][c][:] = B[c]
            o.f["B"][c][:] = B[c] + dB[c]
        o.set_particles(self.particles())
        return o

    # ---- HIP library --------------------------------------------------------
    def gpu(self, precision=0, **rt):
        import aperture3_amd as apt
        p = apt.SimParams(N=[self.N[0], self.N[1], 1], guard=self.guard, size=self.size,
                          periodic_boundary=[False, False, False], q_e=self.q_e,
                          max_ptc_number=self.cap, precision=precision, num_species=3,
                          data_interval=self.data_interval, random_seed=43210,
                          coord_system="LogSpherical",
                          current_smoothing=self.lp["current_smoothing"],
                          **{k: v for k, v in self.lp.items() if k != "current_smoothing"},
                          **rt)
        env = apt.sim_environment(p)
        data = apt.sim_data(env)
        B = self.dipole()
        dB = self.perturbation()
        for c in range(3):
            data.set_field(apt.FIELD_BBG, c, B[c])
            data.set_field(apt.FIELD_B, c, B[c] + dB[c])
        data.set_particles(self.particles())
        return env, data
