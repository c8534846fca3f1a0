"""Host-side mirror of the reference's level operators over the C-ABI.

`FixedBGLevelDataFiniteVolumeOperator` and
`TimeDependentFixedBGLevelDataFiniteVolumeOperator` keep the method names,
argument meaning and error behaviour of
Source/Matter/GRHD/FixedBGFArrayBoxFiniteVolume.hpp:446-614 and :677-830,
for one box of a level held in HBM as a torch tensor in FArrayBox layout:
shape (ncomp, nz, ny, nx), float64, contiguous, on a HIP device.  Errors the
reference throws as std::runtime_error raise RuntimeError with the
reference's message; std::invalid_argument raises ValueError.

PyTorch is used here only to own device memory and name streams; every
computation runs in the HIP library through include/grhd_abi.h.
"""
import ctypes

import torch

from . import abi
from .options import (BoxSpec, CircularBinaryPNParams, KerrSchildParams,
                      OperatorOptions)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _raise(rc, err, what):
    if rc == abi.GRHD_OK:
        return
    if rc == abi.GRHD_ERR_INVALID_ARGUMENT:
        raise ValueError("GRHD %s: invalid argument" % what)
    iv = "(%d,%d,%d)" % tuple(err.iv) if err is not None else "?"
    res = err.residual if err is not None else float("nan")
    if rc == abi.GRHD_ERR_RECOVERY_GHOST_FILL:
        raise RuntimeError(
            "GRHD primitive recovery failed at outflow source %s while "
            "filling fixed-background boundary, residual %g" % (iv, res))
    if rc == abi.GRHD_ERR_RECOVERY_RHS:
        raise RuntimeError("GRHD primitive recovery failed at %s with "
                           "residual %g on fixed background" % (iv, res))
    if rc == abi.GRHD_ERR_RECOVERY_SOURCES:
        raise RuntimeError(
            "GRHD primitive recovery failed at %s while computing "
            "fixed-background metric sources, residual %g" % (iv, res))
    if rc == abi.GRHD_ERR_RECOVERY_CFL:
        raise RuntimeError(
            "GRHD primitive recovery failed at %s while computing "
            "fixed-background CFL speed, residual %g" % (iv, res))
    raise RuntimeError("GRHD %s: HIP device error (code %d)" % (what, rc))


class Context:
    """Owns one grhd_ctx (device scratch) on the current HIP device."""

    def __init__(self):
        self.lib = abi.load_library()
        h = ctypes.c_void_p()
        rc = self.lib.grhd_ctx_create(ctypes.byref(h))
        if rc != abi.GRHD_OK:
            raise RuntimeError("grhd_ctx_create failed (%d)" % rc)
        self.handle = h

    def set_timing(self, enable=True):
        """Per-kernel-class HIP-event timing of later calls (resets)."""
        self.lib.grhd_ctx_set_timing(self.handle, int(bool(enable)))

    def timing(self):
        """{class: (total_ms, launches)} accumulated since set_timing."""
        ms = (ctypes.c_double * len(abi.KERNEL_CLASSES))()
        n = (ctypes.c_longlong * len(abi.KERNEL_CLASSES))()
        self.lib.grhd_ctx_get_timing(self.handle, ms, n)
        return {name: (ms[i], n[i])
                for i, name in enumerate(abi.KERNEL_CLASSES)}

    def set_autotune(self, enable=True):
        """Time every face launch (walk length 1/2/4/8, fused or not) on
        the first call of each new level and precision, and keep the
        fastest; off, size rules choose.  Results do not depend on it."""
        rc = self.lib.grhd_ctx_set_autotune(self.handle, int(bool(enable)))
        if rc != abi.GRHD_OK:
            raise RuntimeError("grhd_ctx_set_autotune failed (%d)" % rc)

    def set_face_launch(self, walk, fused=False, mixed=False):
        """Pin the face launch of later calls for FP64 (mixed=False) or
        FP32 faces: walk 1, 2, 4 or 8 faces per thread, or 0 for the 8 x 8
        tiles, fused into one launch or on streams; walk=None unpins.  A
        pin comes before tuning; results do not depend on it."""
        w = -1 if walk is None else int(walk)
        rc = self.lib.grhd_ctx_set_face_launch(self.handle, int(bool(mixed)),
                                               w, int(bool(fused)))
        if rc != abi.GRHD_OK:
            raise ValueError("grhd_ctx_set_face_launch: walk %r fused %r "
                             "(%d)" % (walk, fused, rc))

    def face_launch(self, mixed=False):
        """{'walk', 'fused', 'tuned', 'ms'} of the current level's face
        launch for FP64 (mixed=False) or FP32 faces; walk 0 means the y/z
        faces are formed in 8 x 8 tiles rather than by walkers."""
        walk, fused, tuned = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        rc = self.lib.grhd_ctx_get_face_launch(
            self.handle, int(bool(mixed)), ctypes.byref(walk),
            ctypes.byref(fused), ctypes.byref(tuned), ctypes.byref(ms))
        if rc != abi.GRHD_OK:
            raise RuntimeError("grhd_ctx_get_face_launch: no level call yet "
                               "(%d)" % rc)
        return {"walk": walk.value, "fused": bool(fused.value),
                "tuned": bool(tuned.value), "ms": ms.value}

    def close(self):
        if self.handle:
            self.lib.grhd_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def new_state(box: BoxSpec, ncomp: int = 11, device="cuda"):
    """Zeroed FArrayBox tensor for `box`."""
    return torch.zeros((ncomp,) + box.storage_shape, dtype=torch.float64,
                       device=device)


def new_rhs(box: BoxSpec, ncomp: int = 11, device="cuda"):
    return torch.zeros((ncomp,) + box.interior_shape, dtype=torch.float64,
                       device=device)


def _check_fab(t, shape, name):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("%s must be a contiguous float64 HIP tensor" % name)
    if tuple(t.shape[1:]) != tuple(shape):
        raise ValueError("%s has shape %s, box needs (ncomp,)+%s"
                         % (name, tuple(t.shape), tuple(shape)))


class FixedBGLevelDataFiniteVolumeOperator:
    """One-box mirror of FixedBGLevelDataFiniteVolumeOperator
    (FixedBGFArrayBoxFiniteVolume.hpp:446-614), plus advance_rk4."""

    time_dependent = False

    def __init__(self, background, box: BoxSpec,
                 options: OperatorOptions = None, comp_base: int = 0,
                 context: Context = None):
        if not isinstance(background, (KerrSchildParams,
                                       CircularBinaryPNParams)):
            raise TypeError("background must be KerrSchildParams or "
                            "CircularBinaryPNParams")
        self.background = background
        self.box = box
        self.options = options or OperatorOptions()
        self.comp_base = comp_base
        self.ctx = context or Context()
        self.lib = self.ctx.lib
        self._cbox = box.to_c()
        self._cbg = background.to_c()
        self._set_params()

    def _set_params(self):
        opts = self.options
        if not self.time_dependent:
            opts = OperatorOptions(**{**opts.__dict__,
                                      "use_metric_volume_time_sources": False})
        self._cparams = opts.to_c()

    def _args(self, state, time):
        _check_fab(state, self.box.storage_shape, "state")
        ncomp = state.shape[0]
        return (self.ctx.handle, ctypes.byref(self._cbox),
                ctypes.byref(self._cparams), ctypes.byref(self._cbg),
                ctypes.c_double(time)), ncomp

    def compute_rhs(self, rhs, state, time=0.0, source_mask=None,
                    stream=None):
        head, ncomp = self._args(state, time)
        _check_fab(rhs, self.box.interior_shape, "rhs")
        if rhs.shape[0] != ncomp:
            raise ValueError("rhs and state must have the same components")
        err = abi.GrhdError()
        rc = self.lib.grhd_compute_rhs_box(
            *head, _ptr(state), ncomp, self.comp_base, _ptr(rhs),
            _ptr(source_mask), ctypes.byref(err), _stream(stream))
        _raise(rc, err, "compute_rhs")

    def recover_primitives(self, state, time=0.0, mask=None, stream=None):
        head, ncomp = self._args(state, time)
        diag = abi.GrhdDiagnostics()
        rc = self.lib.grhd_recover_primitives_box(
            *head, _ptr(state), ncomp, self.comp_base, _ptr(mask),
            ctypes.byref(diag), _stream(stream))
        _raise(rc, None, "recover_primitives")
        return diag.as_dict()

    def update_conserved(self, state, rhs, dt, recover_primitives_after_update
                         =True, recovery_time=0.0, mask=None, stream=None):
        _check_fab(state, self.box.storage_shape, "state")
        _check_fab(rhs, self.box.interior_shape, "rhs")
        diag = abi.GrhdDiagnostics()
        rc = self.lib.grhd_update_conserved_box(
            self.ctx.handle, ctypes.byref(self._cbox),
            ctypes.byref(self._cparams), ctypes.byref(self._cbg),
            _ptr(state), state.shape[0], self.comp_base, _ptr(rhs),
            ctypes.c_double(dt), int(bool(recover_primitives_after_update)),
            ctypes.c_double(recovery_time), _ptr(mask), ctypes.byref(diag),
            _stream(stream))
        _raise(rc, None, "update_conserved")
        return diag.as_dict()

    def compute_max_inverse_dt(self, state, time=0.0, stream=None):
        head, ncomp = self._args(state, time)
        out = ctypes.c_double()
        err = abi.GrhdError()
        rc = self.lib.grhd_compute_max_inverse_dt_box(
            *head, _ptr(state), ncomp, self.comp_base, ctypes.byref(out),
            ctypes.byref(err), _stream(stream))
        _raise(rc, err, "compute_max_inverse_dt")
        return out.value

    def compute_stable_dt(self, state, cfl, time=0.0, stream=None):
        if cfl <= 0.0:
            raise ValueError("GRHD CFL number must be positive")
        head, ncomp = self._args(state, time)
        out = ctypes.c_double()
        err = abi.GrhdError()
        rc = self.lib.grhd_compute_stable_dt_box(
            *head, _ptr(state), ncomp, self.comp_base, ctypes.c_double(cfl),
            ctypes.byref(out), ctypes.byref(err), _stream(stream))
        _raise(rc, err, "compute_stable_dt")
        return out.value

    def advance_ssprk2(self, state, dt, time=0.0, mask=None, stream=None):
        head, ncomp = self._args(state, time)
        diag = abi.GrhdDiagnostics()
        err = abi.GrhdError()
        rc = self.lib.grhd_advance_ssprk2_box(
            *head, ctypes.c_double(dt), _ptr(state), ncomp, self.comp_base,
            _ptr(mask), ctypes.byref(diag), ctypes.byref(err),
            _stream(stream))
        _raise(rc, err, "advance_ssprk2")
        return diag.as_dict()

    def advance_rk4(self, state, dt, nsteps=1, time=0.0, mask=None,
                    stream=None):
        head, ncomp = self._args(state, time)
        diag = abi.GrhdDiagnostics()
        err = abi.GrhdError()
        rc = self.lib.grhd_advance_rk4_box(
            *head, ctypes.c_double(dt), int(nsteps), _ptr(state), ncomp,
            self.comp_base, _ptr(mask), ctypes.byref(diag),
            ctypes.byref(err), _stream(stream))
        _raise(rc, err, "advance_rk4")
        return diag.as_dict()

    def rk4_adjoint(self, state, dt, nsteps, shell_radius, first_step,
                    checkpoint_every=20, stream=None):
        """Discrete adjoint of `nsteps` RK4 steps (grhd_rk4_adjoint_box):
        J = sum over steps n >= first_step of dt * mdot(U_n), the rest-mass
        flux through |x| = shell_radius.  Returns (J, dJ/dU0 as a (5,) +
        interior-shape tensor, dJ/dGamma, dJ/da) at fixed U0."""
        _check_fab(state, self.box.storage_shape, "state")
        ncomp = state.shape[0]
        grad = torch.zeros((5,) + self.box.interior_shape,
                           dtype=torch.float64, device=state.device)
        J = ctypes.c_double()
        par = (ctypes.c_double * 2)()
        err = abi.GrhdError()
        rc = self.lib.grhd_rk4_adjoint_box(
            self.ctx.handle, ctypes.byref(self._cbox),
            ctypes.byref(self._cparams), ctypes.byref(self._cbg),
            ctypes.c_double(dt), int(nsteps), _ptr(state), ncomp,
            self.comp_base, ctypes.c_double(shell_radius), int(first_step),
            int(checkpoint_every), ctypes.byref(J), _ptr(grad), par,
            ctypes.byref(err), _stream(stream))
        _raise(rc, err, "rk4_adjoint")
        return J.value, grad, par[0], par[1]

    def rk4_objective(self, state, dt, nsteps, shell_radius, first_step,
                      stream=None):
        """The objective J of rk4_adjoint alone (forward sweep only)."""
        _check_fab(state, self.box.storage_shape, "state")
        J = ctypes.c_double()
        err = abi.GrhdError()
        rc = self.lib.grhd_rk4_adjoint_box(
            self.ctx.handle, ctypes.byref(self._cbox),
            ctypes.byref(self._cparams), ctypes.byref(self._cbg),
            ctypes.c_double(dt), int(nsteps), _ptr(state), state.shape[0],
            self.comp_base, ctypes.c_double(shell_radius), int(first_step),
            int(nsteps), ctypes.byref(J), None, None, ctypes.byref(err),
            _stream(stream))
        _raise(rc, err, "rk4_objective")
        return J.value

    def cell_geometry(self, time=0.0, stream=None):
        out = torch.empty((16,) + self.box.storage_shape, dtype=torch.float64,
                          device="cuda")
        rc = self.lib.grhd_cell_geometry_box(
            self.ctx.handle, ctypes.byref(self._cbox), ctypes.byref(self._cbg),
            ctypes.c_double(time), _ptr(out), _stream(stream))
        _raise(rc, None, "cell_geometry")
        return out


class TimeDependentFixedBGLevelDataFiniteVolumeOperator(
        FixedBGLevelDataFiniteVolumeOperator):
    """One-box mirror of TimeDependentFixedBGLevelDataFiniteVolumeOperator
    (FixedBGFArrayBoxFiniteVolume.hpp:677-830): the background is evaluated
    at the call's time and the metric volume-time sources are added when
    enabled in the options."""

    time_dependent = True
