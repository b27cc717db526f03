"""Pins the oracle's per-cell moments (orc_collect_diagnostics), which
restate collect_ptc_diagnostics / collect_photon_diagnostics and the
divideBy with Op_DivAssign (cuda/core/additional_diagnostics.cu:14-100,
core/detail/op.hpp:42-49): closed forms on hand-placed particles."""
import numpy as np

import oracle_ctypes as oc


def state():
    o = oc.OracleSim((8, 6), (2, 2), capacity=64, num_species=3)
    d0 = o.dims[0]
    c1, c2 = 3 + 2 * d0, 5 + 4 * d0
    cells = [c1, c1, c1, c2, c2, oc.MAX_CELL, c2]
    species = [0, 0, 1, 1, 2, 0, 0]
    w = [0.5, 2.0, 1.0, 3.0, 1e-6, 7.0, 1.0]
    p1 = [0.0, 1.0, 2.0, 0.5, 3.0, 1.0, 0.0]
    n = len(cells)
    o.set_particles({"cell": np.array(cells, np.uint32), "x1": np.full(n, 0.5),
                     "x2": np.full(n, 0.5), "x3": np.zeros(n), "p1": np.array(p1),
                     "p2": np.zeros(n), "p3": np.zeros(n), "w": np.array(w),
                     "flag": (np.array(species, np.uint32) << 29)})
    return o, c1, c2, w, p1


def test_moments_closed_form():
    o, c1, c2, w, p1 = state()
    num, gam, phn = o.collect_diagnostics()
    g = [np.sqrt(1.0 + v * v) for v in p1]
    assert num[0][c1] == 2.5 and num[1][c1] == 1.0 and num[1][c2] == 3.0
    assert num[0][c2] == 1.0  # the empty slot (weight 7) is skipped
    assert gam[0][c1] == (0.5 * g[0] + 2.0 * g[1]) / 2.5
    assert gam[1][c2] == g[3] and gam[0][c2] == g[6]
    # summed weight 1e-6 < 1e-5: gamma is set to 0 (Op_DivAssign)
    assert num[2][c2] == 1e-6 and gam[2][c2] == 0.0
    for s in range(3):
        others = np.ones(o.ncell, bool)
        others[[c1, c2]] = False
        assert not np.any(num[s][others]) and not np.any(gam[s][others])
    assert not np.any(phn)


def test_photon_weights():
    o, c1, c2, w, p1 = state()
    o.set_rt(max_photon=8)
    o.ph["cell"][:3] = [c1, c2, c2]
    o.ph["w"][:3] = [0.25, 1.5, 2.0]
    o.ph_number = 3
    _, _, phn = o.collect_diagnostics()
    assert phn[c1] == 0.25 and phn[c2] == 3.5 and phn.sum() == 3.75
