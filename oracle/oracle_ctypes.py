"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
as the checker, and by bench.py's cpu_baseline leg.  The product
(aperture3_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAX_CELL = 0xFFFFFFFF


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_log.restype = C.c_double
    lib.orc_log.argtypes = [C.c_double]
    lib.orc_fx_value.restype = C.c_double
    lib.orc_fx_value.argtypes = [C.c_int64, C.c_int64, C.c_int]
    lib.orc_fx_split.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64)]
    lib.orc_load_uniform.restype = C.c_uint64
    lib.orc_sort_by_cell.restype = C.c_uint64
    lib.orc_sample_mj.restype = C.c_int
    lib.orc_sample_mj.argtypes = [C.c_double, C.c_double, C.c_uint64, C.c_uint64,
                                  C.POINTER(C.c_double)]
    return lib


class OrcMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("dims", C.c_int * 3), ("guard", C.c_int * 3),
                ("periodic", C.c_int * 3), ("delta", C.c_double * 3),
                ("inv_delta", C.c_double * 3), ("lower", C.c_double * 3)]


_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)


class OrcPtc(C.Structure):
    _fields_ = [("cell", C.POINTER(C.c_uint32)), ("x1", _PD), ("x2", _PD), ("x3", _PD),
                ("p1", _PD), ("p2", _PD), ("p3", _PD), ("E", _PD), ("w", _PD),
                ("flag", C.POINTER(C.c_uint32)), ("id", C.POINTER(C.c_uint64)),
                ("number", C.c_uint64)]


class OrcFields(C.Structure):
    _fields_ = [("E", _PD * 3), ("B", _PD * 3), ("Ebg", _PD * 3), ("Bbg", _PD * 3),
                ("J", _PD * 3), ("Rho", _PD * 4), ("divE", _PD), ("divB", _PD),
                ("Jacc", (_PI64 * 2) * 3), ("Racc", (_PI64 * 2) * 4)]


class OrcSpecies(C.Structure):
    _fields_ = [("q", C.c_double * 8), ("m", C.c_double * 8), ("num_species", C.c_int),
                ("jexp", C.c_int), ("deposit_rho", C.c_int)]


def _d(a):
    return a.ctypes.data_as(_PD)


class OracleSim:
    """The whole Cartesian state as numpy arrays plus the oracle's step
    functions.  Mesh conventions follow Quadmesh (quadmesh.h)."""

    PTC_FIELDS = (("cell", np.uint32), ("x1", np.float64), ("x2", np.float64),
                  ("x3", np.float64), ("p1", np.float64), ("p2", np.float64),
                  ("p3", np.float64), ("E", np.float64), ("w", np.float64),
                  ("flag", np.uint32), ("id", np.uint64))

    def __init__(self, N, guard, size=(1.0, 1.0, 1.0), lower=(0.0, 0.0, 0.0),
                 periodic=(True, True, True), q_e=1.0, ion_mass=1.0,
                 num_species=2, capacity=0, jexp=22, data_interval=100):
        self.lib = load()
        N = list(N) + [1] * (3 - len(N))
        guard = list(guard) + [0] * (3 - len(guard))
        self.mesh = OrcMesh()
        dim = 1
        for d in range(3):
            self.mesh.dims[d] = N[d] + 2 * guard[d]
            self.mesh.guard[d] = guard[d]
            self.mesh.periodic[d] = 1 if (periodic[d] if d < len(periodic) else False) else 0
            self.mesh.delta[d] = size[d] / N[d]
            self.mesh.inv_delta[d] = 1.0 / self.mesh.delta[d]
            self.mesh.lower[d] = lower[d]
            if N[d] > 1:
                dim = d + 1
        self.mesh.dim = dim
        self.dims = [self.mesh.dims[d] for d in range(3)]
        self.ncell = self.dims[0] * self.dims[1] * self.dims[2]
        self.data_interval = data_interval
        nc = self.ncell
        self.f = {k: [np.zeros(nc) for _ in range(3)] for k in ("E", "B", "Ebg", "Bbg", "J")}
        self.rho = [np.zeros(nc) for _ in range(4)]
        self.divE = np.zeros(nc)
        self.divB = np.zeros(nc)
        self.jacc = [[np.zeros(nc, np.int64) for _ in range(2)] for _ in range(3)]
        self.racc = [[np.zeros(nc, np.int64) for _ in range(2)] for _ in range(4)]
        self.sp = OrcSpecies()
        for i in range(8):
            self.sp.q[i] = q_e
            self.sp.m[i] = q_e
        self.sp.q[0] = -q_e
        self.sp.m[2] = q_e * ion_mass
        self.sp.num_species = num_species
        self.sp.jexp = jexp
        self.capacity = capacity
        self.ptc = {k: np.zeros(capacity, dt) for k, dt in self.PTC_FIELDS}
        self.ptc["cell"][:] = MAX_CELL
        self.number = 0
        self._bind()

    # ---- structs -------------------------------------------------------------
    def _bind(self):
        F = OrcFields()
        for c in range(3):
            F.E[c] = _d(self.f["E"][c])
            F.B[c] = _d(self.f["B"][c])
            F.Ebg[c] = _d(self.f["Ebg"][c])
            F.Bbg[c] = _d(self.f["Bbg"][c])
            F.J[c] = _d(self.f["J"][c])
            for l in range(2):
                F.Jacc[c][l] = self.jacc[c][l].ctypes.data_as(_PI64)
        for s in range(4):
            F.Rho[s] = _d(self.rho[s])
            for l in range(2):
                F.Racc[s][l] = self.racc[s][l].ctypes.data_as(_PI64)
        F.divE = _d(self.divE)
        F.divB = _d(self.divB)
        self.F = F

    def _ptc_struct(self):
        P = OrcPtc()
        P.cell = self.ptc["cell"].ctypes.data_as(C.POINTER(C.c_uint32))
        for k in ("x1", "x2", "x3", "p1", "p2", "p3", "E", "w"):
            setattr(P, k, _d(self.ptc[k]))
        P.flag = self.ptc["flag"].ctypes.data_as(C.POINTER(C.c_uint32))
        P.id = self.ptc["id"].ctypes.data_as(C.POINTER(C.c_uint64))
        P.number = self.number
        return P

    # ---- particles -----------------------------------------------------------
    def set_particles(self, ptc: dict):
        n = len(ptc["cell"])
        if n > self.capacity:
            self.capacity = n
            self.ptc = {k: np.zeros(n, dt) for k, dt in self.PTC_FIELDS}
        for k, dt in self.PTC_FIELDS:
            if k in ptc:
                self.ptc[k][:n] = np.asarray(ptc[k], dtype=dt)
            elif k == "E":
                p1, p2, p3 = (np.asarray(ptc[c], np.float64) for c in ("p1", "p2", "p3"))
                self.ptc["E"][:n] = np.sqrt(1.0 + p1 * p1 + p2 * p2 + p3 * p3)
            elif k == "id":
                self.ptc["id"][:n] = np.arange(n, dtype=np.uint64)
        self.ptc["cell"][n:] = MAX_CELL
        self.number = n

    def particles(self):
        return {k: self.ptc[k][: self.number].copy() for k, _ in self.PTC_FIELDS}

    def load_uniform(self, ppc, theta=(0.0, 0.0), drift=(0.0, 0.0), weight=1.0, seed=43210):
        m = self.mesh
        bulk = 1
        for d in range(3):
            bulk *= m.dims[d] - 2 * m.guard[d]
        n = bulk * ppc * 2
        if n > self.capacity:
            self.capacity = n
            self.ptc = {k: np.zeros(n, dt) for k, dt in self.PTC_FIELDS}
        P = self._ptc_struct()
        got = self.lib.orc_load_uniform(C.byref(m), C.byref(P), C.c_int(ppc),
                                        C.c_double(theta[0]), C.c_double(theta[1]),
                                        C.c_double(drift[0]), C.c_double(drift[1]),
                                        C.c_double(weight), C.c_uint64(seed))
        self.number = int(got)
        return self.number

    # ---- step ----------------------------------------------------------------
    def ptc_update(self, dt, step, deposit_rho=None):
        if deposit_rho is None:
            deposit_rho = ((step + 1) % self.data_interval) == 0
        self.sp.deposit_rho = 1 if deposit_rho else 0
        P = self._ptc_struct()
        self.lib.orc_ptc_update(C.byref(self.mesh), C.byref(self.F), C.byref(P),
                                C.byref(self.sp), C.c_double(dt), C.c_uint32(step))
        if not deposit_rho:
            for s in range(4):
                self.rho[s][:] = 0.0

    def field_update(self, dt, step):
        self.lib.orc_field_update(C.byref(self.mesh), C.byref(self.F), C.c_double(dt),
                                  C.c_uint32(step))

    def update_e(self, dt):
        self.lib.orc_update_e(C.byref(self.mesh), C.byref(self.F), C.c_double(dt))

    def update_b(self, dt):
        self.lib.orc_update_b(C.byref(self.mesh), C.byref(self.F), C.c_double(dt))

    def compute_divs(self):
        self.lib.orc_compute_divs(C.byref(self.mesh), C.byref(self.F))

    def guard_copy(self, a):
        self.lib.orc_guard_copy(C.byref(self.mesh), _d(a))

    def filter_current(self, a, fix3d=False):
        tmp = np.zeros_like(a)
        self.lib.orc_filter_current(C.byref(self.mesh), _d(a), _d(tmp), C.c_int(1 if fix3d else 0))

    def sort(self):
        scratch = np.zeros(max(self.number, 1), np.uint64)
        P = self._ptc_struct()
        kept = self.lib.orc_sort_by_cell(C.byref(self.mesh), C.byref(P),
                                         scratch.ctypes.data_as(C.POINTER(C.c_uint64)))
        self.ptc["cell"][int(kept): self.number] = MAX_CELL
        self.number = int(kept)
        return self.number

    def clear_guard_cells(self):
        P = self._ptc_struct()
        self.lib.orc_clear_guard_cells(C.byref(self.mesh), C.byref(P))

    def energy_histogram(self, nbins, E_max, flag_bit=-1):
        h = np.zeros(nbins, np.uint32)
        P = self._ptc_struct()
        self.lib.orc_energy_histogram(C.byref(P), h.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      C.c_int(nbins), C.c_double(E_max), C.c_int(flag_bit))
        return h


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def sample_mj(theta, beta, seed, pid):
    lib = load()
    u = (C.c_double * 3)()
    ok = lib.orc_sample_mj(theta, beta, seed, pid, u)
    return ok, (u[0], u[1], u[2])
