"""data_exporter -- HDF5 output and snapshots (src/utils/data_exporter.h:16-90,
src/utils/data_exporter.cpp) over the C ABI of ``libaperture_io.so``
(include/aperture_io.h).

The files written are the reference's: ``grid.h5``, ``fld.%05d.h5`` with
the default field set, ``ptc.%05d.h5`` with the tracked particles,
``data.xmf`` and the snapshot file.  The library is loaded when the first
exporter is made; if it is missing that raises (the HDF5 work has no
Python fallback either).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import shutil

import numpy as np

from . import (APT_OK, COORD_LOGSPH, AptError, AptGrid, load_library, sim_data,
               sim_environment)

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(_HERE, "libaperture_io.so")
MAX_TRACKED = 1000000  # constant_defs.h:34

_P = C.c_void_p
IO_SIGNATURES = {
    "apt_io_last_error": [],
    "apt_io_save_snapshot": [_P, C.c_char_p, C.c_uint32, C.c_double, C.c_int, C.c_int,
                             C.c_uint64],
    "apt_io_load_snapshot": [_P, C.POINTER(AptGrid), C.c_char_p, C.POINTER(C.c_uint32),
                             C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "apt_io_write_grid": [C.POINTER(AptGrid), C.c_int, C.c_char_p, C.c_int],
    "apt_io_write_fields": [_P, C.POINTER(AptGrid), C.c_int, C.c_char_p, C.c_int, C.c_int],
    "apt_io_write_tracked": [_P, C.POINTER(AptGrid), C.c_int, C.c_char_p, C.c_int,
                             C.c_uint64, C.c_uint64],
    "apt_io_write_xmf": [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                         C.POINTER(C.c_int), C.c_int, C.c_int],
    "apt_io_read_dataset": [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_uint64,
                            C.POINTER(C.c_uint64)],
    "apt_io_downsample": [C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                          C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)],
}

_io = None


def load_io_library(path: str = IO_LIB_PATH):
    """Load libaperture_io.so (and libaperture_amd.so, which it links)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m aperture3_amd.build` "
            "(needs the HDF5 C headers)")
    load_library()
    lib = C.CDLL(path)
    for name, args in IO_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name.endswith("last_error") else C.c_int
    _io = lib
    return lib


def _check(lib, st, what):
    if st != APT_OK:
        msg = lib.apt_io_last_error()
        raise AptError(f"{what}: status {st}: {msg.decode() if msg else ''}")


def downsample(f, dims, guard, dim, stagger, d):
    """multi_array::downsample (cuda/core/multi_array.cu:18-68, 296-326) of
    one host array with guard cells: every d-th node from the guard,
    interpolated from `stagger` to the node stagger 0b111.  Returns the
    FP32 array of N/d values per axis, x fastest."""
    lib = load_io_library()
    f = np.ascontiguousarray(f, np.float64).ravel()
    ext = [(dims[a] - 2 * guard[a]) if a < dim else 1 for a in range(3)]
    ext = [e // d if e > d else e for e in ext]
    out = np.empty(ext[0] * ext[1] * ext[2], np.float32)
    cd = (C.c_int * 3)(*dims)
    cg = (C.c_int * 3)(*guard)
    _check(lib, lib.apt_io_downsample(f.ctypes.data_as(C.POINTER(C.c_double)), cd, cg, dim,
                                      stagger, d, out.ctypes.data_as(C.POINTER(C.c_float))),
           "apt_io_downsample")
    return out


def read_dataset(filename: str, name: str):
    """A whole HDF5 dataset as FP64 (H5File::read_vector)."""
    lib = load_io_library()
    n = C.c_uint64()
    _check(lib, lib.apt_io_read_dataset(filename.encode(), name.encode(), None, 0,
                                        C.byref(n)), "apt_io_read_dataset")
    out = np.empty(n.value, np.float64)
    _check(lib, lib.apt_io_read_dataset(filename.encode(), name.encode(),
                                        out.ctypes.data_as(C.POINTER(C.c_double)), out.size,
                                        C.byref(n)), "apt_io_read_dataset")
    return out


def read_coef_h5(filename: str):
    """The field-line profile of the 1D GR gap from coef.h5 (the datasets
    Grid_1dGR::compute_coef reads, grid_1dgr.cpp:60-66), for
    Grid_1dGR.compute_coef."""
    from . import GR1D_PROFILE
    return {k: read_dataset(filename, k) for k in GR1D_PROFILE}


class data_exporter:
    """data_exporter(sim_environment&, uint32_t& timestep)
    (data_exporter.cpp:80-117): makes the output directory, copies the
    configuration file and writes grid.h5."""

    def __init__(self, env: sim_environment, timestep: int = 0):
        self.lib = load_io_library()
        self.env = env
        p = env.params()
        self.params = p
        self.outputDirectory = p.data_dir if p.data_dir.endswith("/") else p.data_dir + "/"
        os.makedirs(self.outputDirectory, exist_ok=True)
        d = p.downsample
        # output extent: N / downsample along the mesh dimensions
        self.out_ext = [p.N[a] // d if a < env.dim else 1 for a in range(3)]
        self.m_output_num = 0
        self._times: list[float] = []
        self._nums: list[int] = []
        self._grid = self._grid_struct()
        self.copy_config_file()
        self.write_grid()

    def _grid_struct(self):
        p = self.params
        g = AptGrid()
        for a in range(3):
            g.N[a] = p.N[a]
            g.guard[a] = p.guard[a]
            g.lower[a] = p.lower[a]
            g.size[a] = p.size[a]
            g.periodic[a] = 1 if p.periodic_boundary[a] else 0
        return g

    def _dir(self):
        return self.outputDirectory.rstrip("/").encode()

    # ---- data_exporter.cpp:120-224 ------------------------------------------
    def write_grid(self):
        coord = COORD_LOGSPH if self.params.coord_system == "LogSpherical" else 0
        _check(self.lib, self.lib.apt_io_write_grid(C.byref(self._grid), coord, self._dir(),
                                                    self.params.downsample),
               "apt_io_write_grid")

    def copy_config_file(self):
        """Copies params.conf_file to <data_dir>/config.toml when one is set."""
        src = self.params.conf_file
        if src and os.path.exists(src):
            shutil.copyfile(src, self.outputDirectory + "config.toml")

    # ---- data_exporter.cpp:639-665 ------------------------------------------
    def write_output(self, data: sim_data, timestep: int, time: float):
        self.write_field_output(data, timestep, time)
        self._times.append(float(time))
        self._nums.append(self.m_output_num)
        self.write_xmf()
        self.write_ptc_output(data, timestep, time)
        self.m_output_num += 1

    def write_field_output(self, data: sim_data, timestep: int, time: float):
        _check(self.lib, self.lib.apt_io_write_fields(
            data.ctx, C.byref(self._grid), self.params.num_species, self._dir(),
            self.m_output_num, self.params.downsample), "apt_io_write_fields")

    def write_ptc_output(self, data: sim_data, timestep: int, time: float):
        p = self.params
        cap = min(p.max_ptc_number, MAX_TRACKED)
        cap_ph = min(max(p.max_photon_number, 2), MAX_TRACKED)
        _check(self.lib, self.lib.apt_io_write_tracked(
            data.ctx, C.byref(self._grid), p.num_species, self._dir(), self.m_output_num,
            max(cap, 2), cap_ph), "apt_io_write_tracked")

    def write_xmf(self):
        n = len(self._times)
        ext = (C.c_int * 3)(*self.out_ext)
        t = (C.c_double * max(n, 1))(*self._times)
        nums = (C.c_int * max(n, 1))(*self._nums)
        _check(self.lib, self.lib.apt_io_write_xmf(self._dir(), self.env.dim, ext, t, nums, n,
                                                   self.params.num_species),
               "apt_io_write_xmf")

    # ---- data_exporter.cpp:399-593 ------------------------------------------
    def save_snapshot(self, filename: str, data: sim_data, step: int, time: float):
        p = self.params
        _check(self.lib, self.lib.apt_io_save_snapshot(
            data.ctx, filename.encode(), step, time, p.data_interval, self.m_output_num,
            p.random_seed), "apt_io_save_snapshot")

    def load_snapshot(self, filename: str, data: sim_data):
        """Returns (step, time); (None, None) when the file does not exist,
        as the reference proceeds without it (data_exporter.cpp:493-499)."""
        if not os.path.exists(filename):
            print("Can't find restart file, proceeding without loading it!")
            return None, None
        step = C.c_uint32()
        time = C.c_double()
        num = C.c_int()
        _check(self.lib, self.lib.apt_io_load_snapshot(
            data.ctx, C.byref(self._grid), filename.encode(), C.byref(step), C.byref(time),
            C.byref(num)),
            "apt_io_load_snapshot")
        self.m_output_num = num.value
        self.prepare_xmf_restart(step.value, time.value)
        return step.value, time.value

    def prepare_xmf_restart(self, restart_step: int, time: float):
        """Keeps the outputs of data.xmf up to the one at `time`
        (data_exporter.cpp:595-636)."""
        path = self.outputDirectory + "data.xmf"
        self._times, self._nums = [], []
        if os.path.exists(path):
            text = open(path).read()
            for block in re.findall(r'<Grid Name="quadmesh" Type="Uniform">(.*?)</Grid>', text,
                                    re.S):
                tm = re.search(r'<Time Type="Single" Value="([^"]+)"', block)
                fm = re.search(r"fld\.(\d{5})\.h5:", block)
                if not tm or not fm:
                    continue
                t = float(tm.group(1))
                self._times.append(t)
                self._nums.append(int(fm.group(1)))
                if abs(np.float32(t) - np.float32(time)) < 1.0e-4:
                    break
        self.write_xmf()
