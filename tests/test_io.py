"""CPU checks of the HDF5 output library (include/aperture_io.h): it loads and
exports every declared entry point; the downsampling of
multi_array::downsample (cuda/core/multi_array.cu:18-68, interpolation.cuh:
213-262) keeps constants and shifts a linear ramp by the reference's
neighbour offsets; grid.h5 holds Quadmesh::pos at the downsampled nodes
(data_exporter.cpp:120-216); data.xmf has the reference's XDMF layout
(:227-388) and a restart keeps its outputs up to the restart time
(:595-636).  None of these touch a GPU."""
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import aperture3_amd as apt
import h5read
from aperture3_amd.exporter import IO_LIB_PATH, downsample, load_io_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(h5read.HDF5_LIB is None or not os.path.exists(IO_LIB_PATH),
                                reason="HDF5 library or libaperture_io.so not available")


def test_io_library_exports_header():
    text = open(os.path.join(ROOT, "include", "aperture_io.h")).read()
    names = set(re.findall(r"\b(apt_io_\w+)\s*\(", text))
    assert len(names) == 9
    lib = load_io_library()
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_downsample_constant(dim, d):
    N = [12, 9, 6][:dim] + [1] * (3 - dim)
    guard = [2] * dim + [0] * (3 - dim)
    dims = [N[a] + 2 * guard[a] for a in range(3)]
    f = np.full(dims[0] * dims[1] * dims[2], -1.75)
    for st in range(8):
        out = downsample(f, dims, guard, dim, st, d)
        n = [N[a] // d if (a < dim and N[a] > d) else N[a] for a in range(3)]
        assert out.size == n[0] * n[1] * n[2]
        assert np.all(out == np.float32(-1.75))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_downsample_linear_ramp(dim):
    """f = i + 10 j + 100 k: along an axis with stagger bit 1 the node
    value is the sampled index; with bit 0 it is the mean of the indices -2
    and -1 from it (interpolate*, out stagger 0b111), i.e. index - 1.5."""
    N = [8, 6, 4][:dim] + [1] * (3 - dim)
    guard = [2] * dim + [0] * (3 - dim)
    dims = [N[a] + 2 * guard[a] for a in range(3)]
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]),
                          indexing="ij")
    f = (i + 10.0 * j + 100.0 * k).ravel()
    d = 2
    for st in range(8):
        out = downsample(f, dims, guard, dim, st, d)
        n = [N[a] // d if a < dim else 1 for a in range(3)]
        expect = np.zeros((n[2], n[1], n[0]))
        kk, jj, ii = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]),
                                 indexing="ij")
        for a, (idx, scale) in enumerate(((ii, 1.0), (jj, 10.0), (kk, 100.0))):
            if a >= dim:
                continue
            node = idx * d + guard[a]
            shift = 0.0 if (st >> a) & 1 else -1.5
            expect = expect + scale * (node + shift)
        np.testing.assert_array_equal(out.reshape(n[2], n[1], n[0]), expect.astype(np.float32))


def test_downsample_rejects_small_guard():
    dims, guard = [10, 1, 1], [1, 0, 0]
    with pytest.raises(apt.AptError):
        downsample(np.zeros(10), dims, guard, 1, 0, 1)


def grid_file(tmp_path, **kw):
    p = apt.SimParams(data_dir=str(tmp_path), **kw)
    env = apt.sim_environment(p)
    apt.data_exporter(env)
    with h5read.File(str(tmp_path / "grid.h5")) as f:
        return env, {n: f.read(n) for n in f.names()}


def test_grid_2d_cartesian(tmp_path):
    env, g = grid_file(tmp_path, N=[8, 6, 1], guard=[2, 2, 0], lower=[0.5, -1.0, 0.0],
                       size=[2.0, 3.0, 1.0], downsample=2)
    assert set(g) == {"x1", "x2"}
    assert g["x1"].shape == (3, 4) and g["x1"].dtype == np.float32
    # Quadmesh::pos(i, n, false): lower + delta (n - guard + 0.5), n = i d + guard
    x = 0.5 + 0.25 * (2 * np.arange(4) + 0.5)
    y = -1.0 + 0.5 * (2 * np.arange(3) + 0.5)
    np.testing.assert_array_equal(g["x1"], np.tile(x, (3, 1)).astype(np.float32))
    np.testing.assert_array_equal(g["x2"], np.tile(y[:, None], (1, 4)).astype(np.float32))


def test_grid_logsph(tmp_path):
    env, g = grid_file(tmp_path, N=[8, 4, 1], guard=[3, 3, 0], lower=[0.0, 0.0, 0.0],
                       size=[np.log(20.0), np.pi, 1.0], coord_system="LogSpherical",
                       precision=apt.APT_PREC_FP64)
    lr = (np.arange(8) + 0.5) * np.log(20.0) / 8
    th = ((np.arange(4) + 0.5) * np.pi / 4).astype(np.float32)
    r = np.exp(lr).astype(np.float32)
    np.testing.assert_allclose(g["x1"], np.outer(np.sin(th), r), rtol=2e-7)
    np.testing.assert_allclose(g["x2"], np.outer(np.cos(th), r), rtol=2e-7, atol=1e-6)
    # r^2 = x1^2 + x2^2 at every node
    np.testing.assert_allclose(np.hypot(g["x1"], g["x2"]), np.tile(r, (4, 1)), rtol=1e-6)


def test_grid_1d_and_3d(tmp_path):
    env, g = grid_file(tmp_path, N=[10, 1, 1], guard=[2, 0, 0], size=[5.0, 1.0, 1.0])
    assert set(g) == {"x1"} and g["x1"].shape == (10,)
    np.testing.assert_array_equal(g["x1"], (0.5 * (np.arange(10) + 0.5)).astype(np.float32))
    env, g = grid_file(tmp_path, N=[4, 3, 2], guard=[2, 2, 2], size=[4.0, 3.0, 2.0])
    assert g["x3"].shape == (2, 3, 4)
    # every plane holds its own coordinates
    np.testing.assert_array_equal(g["x3"][:, 0, 0], np.array([0.5, 1.5], np.float32))
    np.testing.assert_array_equal(g["x1"][1, 2], np.array([0.5, 1.5, 2.5, 3.5], np.float32))


def xmf_grids(path):
    root = ET.parse(path).getroot()
    coll = root.find("Domain").find("Grid")
    assert coll.get("GridType") == "Collection" and coll.get("CollectionType") == "Temporal"
    return coll.findall("Grid")


def test_xmf_layout(tmp_path):
    p = apt.SimParams(data_dir=str(tmp_path), N=[16, 8, 1], guard=[2, 2, 0], downsample=2,
                      num_species=3)
    ex = apt.data_exporter(apt.sim_environment(p))
    ex._times, ex._nums = [0.0, 0.125, 1e-05], [0, 1, 7]
    ex.write_xmf()
    text = open(tmp_path / "data.xmf").read()
    assert text.startswith('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n')
    assert text.endswith("</Grid>\n</Domain>\n</Xdmf>\n")
    grids = xmf_grids(tmp_path / "data.xmf")
    assert [g.find("Time").get("Value") for g in grids] == ["0", "0.125", "1e-05"]
    g = grids[2]
    assert g.find("Topology").get("Type") == "2DSMesh"
    assert g.find("Topology").get("NumberOfElements") == "4 8"
    items = [e.text.strip() for e in g.find("Geometry").findall("DataItem")]
    assert items == ["grid.h5:x1", "grid.h5:x2"]
    attrs = [a.get("Name") for a in g.findall("Attribute")]
    assert attrs == ["E1", "E2", "E3", "B1", "B2", "B3", "J1", "J2", "J3", "Rho_e", "Rho_p",
                     "Rho_i", "divE", "divB", "photon_produced", "pair_produced"]
    assert g.findall("Attribute")[3].find("DataItem").text.strip() == "fld.00007.h5:B1"


def test_xmf_1d(tmp_path):
    p = apt.SimParams(data_dir=str(tmp_path), N=[32, 1, 1], guard=[2, 0, 0])
    ex = apt.data_exporter(apt.sim_environment(p))
    ex._times, ex._nums = [2.5], [0]
    ex.write_xmf()
    g = xmf_grids(tmp_path / "data.xmf")[0]
    assert g.find("Topology").get("NumberOfElements") == "32 1"
    assert [e.text.strip() for e in g.find("Geometry").findall("DataItem")] == ["grid.h5:x1"]


def test_xmf_restart_keeps_outputs_to_restart_time(tmp_path):
    p = apt.SimParams(data_dir=str(tmp_path), N=[8, 8, 1], guard=[2, 2, 0])
    env = apt.sim_environment(p)
    ex = apt.data_exporter(env)
    ex._times, ex._nums = [0.0, 0.5, 1.0, 1.5], [0, 1, 2, 3]
    ex.write_xmf()
    ex2 = apt.data_exporter(env)
    ex2.prepare_xmf_restart(10, 1.0)
    assert ex2._times == [0.0, 0.5, 1.0] and ex2._nums == [0, 1, 2]
    assert len(xmf_grids(tmp_path / "data.xmf")) == 3


def test_read_dataset_and_coef_profile(tmp_path):
    """apt_io_read_dataset reads what the HDF5 C library wrote (any stored
    float type, converted to FP64); read_coef_h5 returns the seven profile
    datasets of coef.h5."""
    from aperture3_amd.exporter import read_coef_h5, read_dataset
    path = str(tmp_path / "coef.h5")
    rng = np.random.default_rng(1)
    data = {k: rng.random(17) for k in apt.GR1D_PROFILE}
    data["th"] = data["th"].astype(np.float32)
    h5read.write(path, data)
    np.testing.assert_array_equal(read_dataset(path, "r"), data["r"])
    prof = read_coef_h5(path)
    for k in apt.GR1D_PROFILE:
        np.testing.assert_array_equal(prof[k], data[k].astype(np.float64))
    with pytest.raises(apt.AptError):
        read_dataset(path, "missing")
