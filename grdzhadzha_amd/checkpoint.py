"""KITPCHK1 checkpoints of a level
(Examples/GRHDLevelDataBinaryPNKITPDisk/Main_GRHDLevelDataBinaryPNKITPDisk.cpp
:1191-1592).

File layout (version 1, native little-endian), as the reference writes it:

    "KITPCHK1"                                  8 bytes
    u32 version (1), i32 num_components (NUM_VARS)
    i32 num_cells, i32 num_z_cells, i32 num_ghosts
    i32 steps, i32 frame_index
    f64 time, f64 next_output_time, f64 next_checkpoint_time
    u64 recovery totals: num_cells, failed, floored, resets
    Diagnostics: 7 x f64 (mass, disk_mass, tau, max_rho, max_pressure,
                 max_velocity, max_lorentz_factor), 3 x u64 (failed,
                 floored, resets)
    u64 signature length, signature bytes (config_signature)
    u64 value count, value count x f64 payload

The payload is the ProblemDomain grown by num_ghosts, point major with the
NUM_VARS components of a cell adjacent.  On this side the level lives in
HBM: `grhd_level_checkpoint_pack` / `_unpack` transpose between the level's
FABs and the payload on the device, and one transfer moves the payload
across PCIe; the header is written and validated here, with the
reference's error messages.  A checkpoint written with one box layout
restores into any other.
"""
import ctypes
import os
import struct
from dataclasses import dataclass, field, fields

import numpy as np
import torch

from .operator import _ptr, _raise, _stream

MAGIC = b"KITPCHK1"
VERSION = 1
MAX_SIGNATURE = 16384


@dataclass
class KITPConfig:
    """The SimulationConfig fields that a checkpoint records or validates
    (Main_GRHDLevelDataBinaryPNKITPDisk.cpp:36-83), with its defaults, in
    the order config_signature prints them (:1202-1224)."""
    num_cells: int = 192
    num_z_cells: int = 25
    num_ghosts: int = 2
    domain_length: float = 96.0
    output_interval: float = 0.5
    cfl: float = 0.14
    max_dt: float = 0.02
    mass_1: float = 0.5
    mass_2: float = 0.5
    separation: float = 8.0
    binary_phase: float = 0.0
    binary_time: float = 0.0
    binary_angular_velocity: float = 0.0
    softening_radius: float = 0.5
    metric_time_derivative_step: float = 1.0e-3
    mach: float = 10.0
    aspect_ratio: float = 0.1
    cavity_radius_factor: float = 2.5
    cavity_delta: float = 1.0e-5
    outer_cutoff_radius: float = 1.0e10
    rho_peak: float = 4.0e-5
    rho_floor: float = 1.0e-10
    pressure_floor: float = 8.0e-14
    max_velocity_squared: float = 0.72
    limiter_theta: float = 1.3
    use_reconstruction: bool = True
    use_static_metric_sources: bool = True
    use_metric_volume_time_sources: bool = True
    alpha_viscosity: float = 0.03
    viscosity_beta: float = 0.0
    sink_size: float = 1.0
    sink_rate: float = 1.0e4
    sink_shape: float = 4.0

    def domain(self):
        """make_domain_box (:227-233): [0, n-1]^2 x [0, nz-1]."""
        return ((0, 0, 0),
                (self.num_cells - 1, self.num_cells - 1,
                 self.num_z_cells - 1))

    def value_count(self, num_vars):
        """checkpoint_value_count over checkpoint_storage_box (:1360-1375)."""
        g = self.num_ghosts
        return num_vars * (self.num_cells + 2 * g) ** 2 * \
            (self.num_z_cells + 2 * g)


def format_value(v):
    """One value as `std::ostream << std::setprecision(17)` prints it:
    ints and bools as integers, doubles as printf's %.17g."""
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, int):
        return "%d" % v
    return "%.17g" % v


def config_signature(config: KITPConfig):
    """config_signature (:1202-1224): the fields joined by single spaces."""
    return " ".join(format_value(getattr(config, f.name))
                    for f in fields(config))


@dataclass
class Diagnostics:
    """The example's Diagnostics record (:109-121) as a checkpoint stores
    it (write_checkpoint_diagnostics, :1242-1261)."""
    mass: float = 0.0
    disk_mass: float = 0.0
    tau: float = 0.0
    max_rho: float = 0.0
    max_pressure: float = 0.0
    max_velocity: float = 0.0
    max_lorentz_factor: float = 1.0
    failed_recoveries: int = 0
    floored_primitives: int = 0
    conserved_resets: int = 0


@dataclass
class RestartData:
    """RestartData (:1191-1200)."""
    time: float = 0.0
    next_output_time: float = 0.0
    next_checkpoint_time: float = 0.0
    steps: int = 0
    frame_index: int = 0
    initial_diagnostics: Diagnostics = field(default_factory=Diagnostics)
    recovery_totals: dict = field(default_factory=lambda: {
        "num_cells": 0, "num_failed_recoveries": 0,
        "num_floored_primitives": 0, "num_conserved_resets": 0})


_TOTALS = ("num_cells", "num_failed_recoveries", "num_floored_primitives",
           "num_conserved_resets")


def checkpoint_path(checkpoint_dir, steps):
    """checkpoint_path (:1377-1383)."""
    return os.path.join(checkpoint_dir, "checkpoint_step_%08d.bin" % steps)


def _header(config, restart, num_vars, value_count):
    d = restart.initial_diagnostics
    sig = config_signature(config).encode()
    parts = [MAGIC,
             struct.pack("<Ii", VERSION, num_vars),
             struct.pack("<iii", config.num_cells, config.num_z_cells,
                         config.num_ghosts),
             struct.pack("<ii", restart.steps, restart.frame_index),
             struct.pack("<ddd", restart.time, restart.next_output_time,
                         restart.next_checkpoint_time),
             struct.pack("<4Q", *[int(restart.recovery_totals[k])
                                  for k in _TOTALS]),
             struct.pack("<7d", d.mass, d.disk_mass, d.tau, d.max_rho,
                         d.max_pressure, d.max_velocity,
                         d.max_lorentz_factor),
             struct.pack("<3Q", d.failed_recoveries, d.floored_primitives,
                         d.conserved_resets),
             struct.pack("<Q", len(sig)), sig,
             struct.pack("<Q", value_count)]
    return b"".join(parts)


def write_checkpoint_file(path, config, restart, values, num_vars):
    """Header + host payload, written to `path`.tmp and renamed (:1437-1495)."""
    values = np.ascontiguousarray(values, dtype="<f8")
    if values.size != config.value_count(num_vars):
        raise ValueError("payload does not match the checkpoint grid")
    tmp = path + ".tmp"
    try:
        with open(tmp, "wb") as out:
            out.write(_header(config, restart, num_vars, values.size))
            out.write(values.tobytes())
    except OSError as e:
        raise RuntimeError("failed while writing checkpoint " + tmp) from e
    try:
        os.replace(tmp, path)
    except OSError as e:
        raise RuntimeError("could not finalize checkpoint " + path) from e


def _read(f, fmt):
    size = struct.calcsize(fmt)
    b = f.read(size)
    if len(b) != size:
        raise RuntimeError("checkpoint ended unexpectedly")
    return struct.unpack(fmt, b)


def read_checkpoint_header(f, config, num_vars):
    """read_checkpoint_header (:1289-1346): validates magic, version,
    component count, grid and configuration signature."""
    if f.read(8) != MAGIC:
        raise RuntimeError("not a KITP checkpoint")
    version, ncomp = _read(f, "<Ii")
    cells, zcells, ghosts = _read(f, "<iii")
    if version != VERSION or ncomp != num_vars:
        raise RuntimeError("unsupported KITP checkpoint format")
    if (cells, zcells, ghosts) != (config.num_cells, config.num_z_cells,
                                   config.num_ghosts):
        raise RuntimeError("checkpoint grid does not match this run")
    r = RestartData()
    r.steps, r.frame_index = _read(f, "<ii")
    r.time, r.next_output_time, r.next_checkpoint_time = _read(f, "<ddd")
    r.recovery_totals = dict(zip(_TOTALS, _read(f, "<4Q")))
    dd = _read(f, "<7d")
    di = _read(f, "<3Q")
    r.initial_diagnostics = Diagnostics(*dd, *di)
    (n,) = _read(f, "<Q")
    if n > MAX_SIGNATURE:
        raise RuntimeError("invalid checkpoint configuration signature")
    sig = f.read(n)
    if len(sig) != n or sig != config_signature(config).encode():
        raise RuntimeError("checkpoint configuration does not match this run")
    (count,) = _read(f, "<Q")
    return r, count


def read_checkpoint_metadata(path, config, num_vars):
    """read_checkpoint_metadata (:1507-1514)."""
    try:
        f = open(path, "rb")
    except OSError as e:
        raise RuntimeError("could not open checkpoint " + path) from e
    with f:
        return read_checkpoint_header(f, config, num_vars)[0]


def read_checkpoint_file(path, config, num_vars):
    """Header and host payload (:1527-1555)."""
    try:
        f = open(path, "rb")
    except OSError as e:
        raise RuntimeError("could not reopen checkpoint " + path) from e
    with f:
        restart, count = read_checkpoint_header(f, config, num_vars)
        expected = config.value_count(num_vars)
        if count != expected:
            raise RuntimeError("checkpoint state size does not match the grid")
        raw = f.read(8 * expected)
        if len(raw) != 8 * expected:
            raise RuntimeError("checkpoint state data is incomplete")
    return restart, np.frombuffer(raw, dtype="<f8")


def _storage_inside(level, config):
    lo, hi = config.domain()
    g = config.num_ghosts
    for b in level.boxes:
        for d in range(3):
            if b.glo[d] < lo[d] - g or b.ghi[d] > hi[d] + g:
                raise RuntimeError("state ghost box exceeds checkpoint storage")


def _device_call(level, name, src, dst, config):
    lo, hi = config.domain()
    I3 = ctypes.c_int * 3
    rc = getattr(level.lib, name)(
        level.ctx.handle, level._cboxes, len(level.boxes), _ptr(src),
        level.ncomp, I3(*lo), I3(*hi), config.num_ghosts, _ptr(dst),
        _stream(None))
    _raise(rc, None, name)


def pack_level(level, pool, config):
    """Device payload (float64 tensor) of a level pool."""
    level._check(pool)
    _storage_inside(level, config)
    values = torch.empty(config.value_count(level.ncomp), dtype=torch.float64,
                         device=pool.device)
    _device_call(level, "grhd_level_checkpoint_pack", pool, values, config)
    return values


def unpack_level(level, values, pool, config):
    """Every storage cell of the level pool from a device payload."""
    level._check(pool)
    _storage_inside(level, config)
    if values.numel() != config.value_count(level.ncomp) or \
            values.dtype != torch.float64 or not values.is_contiguous():
        raise ValueError("payload does not match the checkpoint grid")
    _device_call(level, "grhd_level_checkpoint_unpack", values, pool, config)


def write_checkpoint(path, level, pool, config, restart: RestartData):
    """write_checkpoint (:1385-1505) for a level in HBM."""
    values = pack_level(level, pool, config)
    write_checkpoint_file(path, config, restart, values.cpu().numpy(),
                          level.ncomp)


def read_checkpoint_state(path, level, pool, config):
    """read_checkpoint_state (:1516-1592) into a level in HBM; returns the
    RestartData of the header."""
    restart, host = read_checkpoint_file(path, config, level.ncomp)
    values = torch.from_numpy(host.copy()).to(pool.device)
    unpack_level(level, values, pool, config)
    return restart


__all__ = ["KITPConfig", "Diagnostics", "RestartData", "config_signature",
           "checkpoint_path", "write_checkpoint_file", "read_checkpoint_file",
           "read_checkpoint_header", "read_checkpoint_metadata",
           "write_checkpoint", "read_checkpoint_state", "pack_level",
           "unpack_level"]
