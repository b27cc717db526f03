"""Oracle checks of sort and compaction (particle_base_impl.hpp:250-305,
kernels.cu:25-34) and of the energy histogram (kernels.cu:36-69).

The count known answer mirrors the reference's own sort test
(tests/test_particles_dev.cpp:115-144, mesh of bin/test_particles.toml:
400x400 with guard 3): 2000 particles in cells (200..300)^2 with every
tenth empty, 100 more empty and 900 in guard cells leave 1800 after
clear_guard_cells + sort_by_cell.  Stability and permutation invariants are
checked on random data."""
import numpy as np

import oracle_ctypes as oc


def test_sort_count_known_answer():
    o = oc.OracleSim([400, 400], [3, 3], periodic=(False, False), capacity=3000)
    rng = np.random.default_rng(0)
    d0 = o.dims[0]
    cells = []
    for i in range(2000):
        if i % 10 == 0:
            cells.append(oc.MAX_CELL)
        else:
            cells.append(int(rng.integers(200, 301)) + int(rng.integers(200, 301)) * d0)
    cells += [oc.MAX_CELL] * 100
    cells += [1 + int(rng.integers(200, 301)) * d0 for _ in range(900)]
    n = len(cells)
    o.set_particles({"cell": np.array(cells, np.uint32), "x1": rng.random(n), "x2": rng.random(n),
                     "x3": np.zeros(n), "p1": np.zeros(n), "p2": np.zeros(n), "p3": np.zeros(n),
                     "w": np.ones(n), "flag": np.zeros(n, np.uint32)})
    assert o.number == 3000
    o.clear_guard_cells()
    assert o.sort() == 1800


def test_sort_stable_permutation():
    o = oc.OracleSim([16, 16], [2, 2], capacity=5000)
    rng = np.random.default_rng(1)
    n = 5000
    cells = rng.integers(0, o.ncell, n).astype(np.uint32)
    cells[rng.random(n) < 0.1] = oc.MAX_CELL
    ids = np.arange(n, dtype=np.uint64)
    x1 = rng.random(n)
    o.set_particles({"cell": cells, "x1": x1, "x2": np.zeros(n), "x3": np.zeros(n),
                     "p1": np.zeros(n), "p2": np.zeros(n), "p3": np.zeros(n),
                     "w": np.ones(n), "flag": np.zeros(n, np.uint32), "id": ids})
    kept = o.sort()
    p = o.particles()
    assert kept == int((cells != oc.MAX_CELL).sum())
    assert np.all(np.diff(p["cell"].astype(np.int64)) >= 0)
    # stable: ids increase within each cell; payload travels with the id
    same = np.diff(p["cell"].astype(np.int64)) == 0
    assert np.all(np.diff(p["id"].astype(np.int64))[same] > 0)
    assert np.array_equal(p["x1"], x1[p["id"].astype(np.int64)])
    expect = np.sort(ids[cells != oc.MAX_CELL])
    assert np.array_equal(np.sort(p["id"]), expect)


def test_energy_histogram_bins():
    o = oc.OracleSim([8], [2], capacity=8)
    g = np.array([1.0, 1.5, 2.0, 10.0, 99.0, 1e6, 1.0 + 1e-12, 3.0])
    u = np.sqrt(g * g - 1)
    o.set_particles({"cell": np.full(8, 4, np.uint32), "x1": np.zeros(8), "x2": np.zeros(8),
                     "x3": np.zeros(8), "p1": u, "p2": np.zeros(8), "p3": np.zeros(8),
                     "w": np.ones(8), "flag": np.zeros(8, np.uint32)})
    o.ptc["E"][:8] = g
    h = o.energy_histogram(10, 100.0)
    dl = np.log(100.0) / 10
    idx = np.clip(np.floor(np.log(g) / dl).astype(int), 0, 9)
    assert np.array_equal(h, np.bincount(idx, minlength=10).astype(np.uint32))
