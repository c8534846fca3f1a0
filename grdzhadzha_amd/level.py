"""A level of boxes in one HBM pool, driven through the level C-ABI.

`LevelOperator` mirrors FixedBGLevelDataFiniteVolumeOperator /
TimeDependentFixedBGLevelDataFiniteVolumeOperator
(Source/Matter/GRHD/FixedBGFArrayBoxFiniteVolume.hpp:446-614, :677-830) for a
LevelData of many boxes: every call launches once per kernel for the whole
level (device box table), exchanges internal ghosts between boxes that share
a ProblemDomain, and reduces diagnostics and max 1/dt over the level.

The pool is one float64 device tensor; box b's FArrayBox (ncomp components
over its storage box, component-major, i fastest) starts at
`offsets[b]`.  `pack` / `unpack` move per-box numpy FABs in and out.
"""
import ctypes

import numpy as np
import torch

from . import abi
from .operator import Context, _ptr, _raise, _stream
from .options import (BoxSpec, CircularBinaryPNParams, KerrSchildParams,
                      OperatorOptions)


class LevelOperator:
    def __init__(self, background, boxes, options: OperatorOptions = None,
                 ncomp: int = 11, comp_base: int = 0, time_dependent=False,
                 context: Context = None):
        if not isinstance(background, (KerrSchildParams,
                                       CircularBinaryPNParams)):
            raise TypeError("background must be KerrSchildParams or "
                            "CircularBinaryPNParams")
        self.background = background
        self.boxes = list(boxes)
        self.ncomp = ncomp
        self.comp_base = comp_base
        self.options = options or OperatorOptions()
        if not time_dependent:
            o = dict(self.options.__dict__)
            o["use_metric_volume_time_sources"] = False
            self.options = OperatorOptions(**o)
        self.ctx = context or Context()
        self.lib = self.ctx.lib
        sizes = [ncomp * int(np.prod(b.storage_shape)) for b in self.boxes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(
            np.int64)
        self.pool_size = int(sum(sizes))
        self.interior = [b.num_interior for b in self.boxes]
        self.num_interior = int(sum(self.interior))
        arr = (abi.GrhdLevelBox * len(self.boxes))()
        for n, b in enumerate(self.boxes):
            arr[n].box = b.to_c()
            arr[n].state_offset = int(self.offsets[n])
        self._cboxes = arr
        self._cparams = self.options.to_c()
        self._cbg = background.to_c()

    # ------------------------------------------------------------ pool
    def new_pool(self, device="cuda"):
        return torch.zeros(self.pool_size, dtype=torch.float64, device=device)

    def pack(self, fabs, device="cuda"):
        host = np.empty(self.pool_size)
        for n, (b, f) in enumerate(zip(self.boxes, fabs)):
            o = int(self.offsets[n])
            host[o:o + f.size] = np.ascontiguousarray(f).ravel()
        return torch.from_numpy(host).to(device)

    def unpack(self, pool):
        host = pool.detach().cpu().numpy()
        out = []
        for n, b in enumerate(self.boxes):
            o = int(self.offsets[n])
            size = self.ncomp * int(np.prod(b.storage_shape))
            out.append(host[o:o + size].reshape((self.ncomp,) +
                                                b.storage_shape))
        return out

    def split_interior(self, flat, ncomp=None):
        """Per-box arrays of an interior pool (rhs or mask)."""
        host = flat.detach().cpu().numpy() if torch.is_tensor(flat) else flat
        out, o = [], 0
        for b in self.boxes:
            shape = ((ncomp,) if ncomp else ()) + b.interior_shape
            size = int(np.prod(shape))
            out.append(host[o:o + size].reshape(shape))
            o += size
        return out

    def _check(self, pool):
        if not (pool.is_cuda and pool.dtype == torch.float64 and
                pool.is_contiguous() and pool.numel() == self.pool_size):
            raise ValueError("pool must be a contiguous float64 HIP tensor "
                             "of %d doubles" % self.pool_size)

    def _head(self, time):
        return (self.ctx.handle, self._cboxes, len(self.boxes),
                ctypes.byref(self._cparams), ctypes.byref(self._cbg),
                ctypes.c_double(time))

    # ------------------------------------------------------------ ops
    def compute_rhs(self, rhs, pool, time=0.0, source_mask=None,
                    stream=None):
        self._check(pool)
        if rhs.numel() != self.ncomp * self.num_interior:
            raise ValueError("rhs pool must hold ncomp x interior cells")
        err = abi.GrhdError()
        rc = self.lib.grhd_level_compute_rhs(
            *self._head(time), _ptr(pool), self.ncomp, self.comp_base,
            _ptr(rhs), _ptr(source_mask), ctypes.byref(err),
            _stream(stream))
        _raise(rc, err, "level compute_rhs")

    def add_conserved_rhs_to(self, rhs, pool, time=0.0, stream=None):
        """add_conserved_rhs_to (FixedBGFArrayBoxFiniteVolume.hpp:484-496):
        rhs[D, S_i, tau] += this level's rhs; other components untouched."""
        self._check(pool)
        if rhs.numel() != self.ncomp * self.num_interior:
            raise ValueError("rhs pool must hold ncomp x interior cells")
        err = abi.GrhdError()
        rc = self.lib.grhd_level_add_conserved_rhs(
            *self._head(time), _ptr(pool), self.ncomp, self.comp_base,
            _ptr(rhs), ctypes.byref(err), _stream(stream))
        _raise(rc, err, "level add_conserved_rhs_to")

    def recover_primitives(self, pool, time=0.0, mask=None, stream=None):
        self._check(pool)
        diag = abi.GrhdDiagnostics()
        rc = self.lib.grhd_level_recover_primitives(
            *self._head(time), _ptr(pool), self.ncomp, self.comp_base,
            _ptr(mask), ctypes.byref(diag), _stream(stream))
        _raise(rc, None, "level recover_primitives")
        return diag.as_dict()

    def compute_max_inverse_dt(self, pool, time=0.0, stream=None):
        self._check(pool)
        out = ctypes.c_double()
        err = abi.GrhdError()
        rc = self.lib.grhd_level_compute_max_inverse_dt(
            *self._head(time), _ptr(pool), self.ncomp, self.comp_base,
            ctypes.byref(out), ctypes.byref(err), _stream(stream))
        _raise(rc, err, "level compute_max_inverse_dt")
        return out.value

    def compute_stable_dt(self, pool, cfl, time=0.0, stream=None):
        if cfl <= 0.0:
            raise ValueError("GRHD CFL number must be positive")
        return cfl / max(self.compute_max_inverse_dt(pool, time, stream),
                         1.0e-14)

    def advance_ssprk2(self, pool, dt, time=0.0, mask=None, stream=None):
        self._check(pool)
        diag = abi.GrhdDiagnostics()
        err = abi.GrhdError()
        rc = self.lib.grhd_level_advance_ssprk2(
            *self._head(time), ctypes.c_double(dt), _ptr(pool), self.ncomp,
            self.comp_base, _ptr(mask), ctypes.byref(diag),
            ctypes.byref(err), _stream(stream))
        _raise(rc, err, "level advance_ssprk2")
        return diag.as_dict()

    def advance_rk4(self, pool, dt, nsteps=1, time=0.0, mask=None,
                    stream=None):
        self._check(pool)
        diag = abi.GrhdDiagnostics()
        err = abi.GrhdError()
        rc = self.lib.grhd_level_advance_rk4(
            *self._head(time), ctypes.c_double(dt), int(nsteps), _ptr(pool),
            self.ncomp, self.comp_base, _ptr(mask), ctypes.byref(diag),
            ctypes.byref(err), _stream(stream))
        _raise(rc, err, "level advance_rk4")
        return diag.as_dict()


    # ------------------------------------------------------- diagnostics
    def conserved_sum(self, pool, stream=None):
        """Level sums of D, S_1, S_2, S_3, tau over interior cells
        (compute_leveldata_conserved_sum, LevelDataDriver.hpp:22-57), as a
        5-tuple of floats."""
        self._check(pool)
        out = (ctypes.c_double * 5)()
        rc = self.lib.grhd_level_conserved_sum(
            self.ctx.handle, self._cboxes, len(self.boxes), _ptr(pool),
            self.ncomp, self.comp_base, out, _stream(stream))
        _raise(rc, None, "level conserved_sum")
        return tuple(out)

    def conserved_state_is_finite(self, pool, stream=None):
        """leveldata_conserved_state_is_finite (LevelDataDriver.hpp:59-92)."""
        self._check(pool)
        out = ctypes.c_int()
        rc = self.lib.grhd_level_conserved_state_is_finite(
            self.ctx.handle, self._cboxes, len(self.boxes), _ptr(pool),
            self.ncomp, self.comp_base, ctypes.byref(out), _stream(stream))
        _raise(rc, None, "level conserved_state_is_finite")
        return bool(out.value)


def compute_leveldata_conserved_sum(level: "LevelOperator", pool):
    """Reference-named free function (LevelDataDriver.hpp:22-57)."""
    return level.conserved_sum(pool)


def leveldata_conserved_state_is_finite(level: "LevelOperator", pool):
    """Reference-named free function (LevelDataDriver.hpp:67-92)."""
    return level.conserved_state_is_finite(pool)


def split_box(box: BoxSpec, axis: int, at: int):
    """Two boxes splitting `box` along `axis` (cells < at, cells >= at),
    sharing its ProblemDomain -- the reference's x-split layouts
    (GRHDLevelDataMichelAccretion.cpp:375-401)."""
    lo1, hi1 = list(box.lo), list(box.hi)
    lo2, hi2 = list(box.lo), list(box.hi)
    hi1[axis] = at - 1
    lo2[axis] = at
    mk = lambda lo, hi: BoxSpec(tuple(lo), tuple(hi), box.ghosts, box.dx,  # noqa: E731
                                box.center, box.domain_lo, box.domain_hi,
                                box.periodic)
    return mk(lo1, hi1), mk(lo2, hi2)


def sub_fab(full_fab, full_box: BoxSpec, sub: BoxSpec):
    """Copy of the part of a FAB on `full_box` covering `sub`'s storage."""
    sl = tuple(slice(sub.glo[d] - full_box.glo[d],
                     sub.ghi[d] - full_box.glo[d] + 1) for d in (2, 1, 0))
    return np.ascontiguousarray(full_fab[(slice(None),) + sl]).copy()
