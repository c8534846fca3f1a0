"""ctypes view of the C-ABI declared in include/grhd_abi.h.

The structs mirror the header field for field.  `load_library()` loads the
in-tree HIP build (grdzhadzha_amd/libgrhd_hip.so) and raises if it is
missing: there is no CPU fallback for the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrhd_hip.so")

GRHD_NUM_COMPS = 11
(GRHD_D, GRHD_S1, GRHD_S2, GRHD_S3, GRHD_TAU, GRHD_RHO, GRHD_EPS,
 GRHD_PRESSURE, GRHD_V1, GRHD_V2, GRHD_V3) = range(11)

GRHD_OK = 0
GRHD_ERR_RECOVERY_GHOST_FILL = 1
GRHD_ERR_RECOVERY_RHS = 2
GRHD_ERR_RECOVERY_SOURCES = 3
GRHD_ERR_RECOVERY_CFL = 4
GRHD_ERR_INVALID_ARGUMENT = 5
GRHD_ERR_DEVICE = 6

GRHD_BG_KERR_SCHILD = 0
GRHD_BG_CIRCULAR_BINARY_PN = 1

GRHD_MASK_FAILED = 1
GRHD_MASK_FLOORED = 2
GRHD_MASK_RESET = 4
GRHD_MASK_SOURCE_ATMOSPHERE = 8

GRHD_NOALLOC_SCRATCH_BYTES = 128 << 20

_I3 = ctypes.c_int * 3
_D3 = ctypes.c_double * 3


class GrhdBox(ctypes.Structure):
    _fields_ = [("lo", _I3), ("hi", _I3), ("glo", _I3), ("ghi", _I3),
                ("dlo", _I3), ("dhi", _I3), ("periodic", _I3),
                ("dx", ctypes.c_double), ("center", _D3),
                ("inverse_dx", _D3)]


class GrhdParams(ctypes.Structure):
    _fields_ = [("adiabatic_index", ctypes.c_double),
                ("rho_floor", ctypes.c_double),
                ("pressure_floor", ctypes.c_double),
                ("max_velocity_squared", ctypes.c_double),
                ("metric_source_atmosphere_factor", ctypes.c_double),
                ("max_iterations", ctypes.c_int),
                ("tolerance", ctypes.c_double),
                ("recovery_pressure_floor", ctypes.c_double),
                ("recovery_max_velocity_squared", ctypes.c_double),
                ("limiter_theta", ctypes.c_double),
                ("use_reconstruction", ctypes.c_int),
                ("use_static_metric_sources", ctypes.c_int),
                ("use_metric_volume_time_sources", ctypes.c_int),
                ("metric_time_derivative_step", ctypes.c_double),
                ("mixed_precision", ctypes.c_int),
                ("excise_horizon", ctypes.c_int),
                ("excision_inner_buffer", ctypes.c_double),
                ("excision_outer_buffer", ctypes.c_double),
                ("mixed_fp64_margin", ctypes.c_double),
                ("mixed_fp64_shift", ctypes.c_double)]


class GrhdBackground(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("mass", ctypes.c_double),
                ("spin", ctypes.c_double), ("mass_1", ctypes.c_double),
                ("mass_2", ctypes.c_double), ("separation", ctypes.c_double),
                ("angular_velocity", ctypes.c_double),
                ("phase", ctypes.c_double), ("time", ctypes.c_double),
                ("softening_radius", ctypes.c_double)]


class GrhdDiagnostics(ctypes.Structure):
    _fields_ = [("num_cells", ctypes.c_longlong),
                ("num_failed_recoveries", ctypes.c_longlong),
                ("num_floored_primitives", ctypes.c_longlong),
                ("num_conserved_resets", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GrhdLevelBox(ctypes.Structure):
    _fields_ = [("box", GrhdBox), ("state_offset", ctypes.c_longlong)]


class GrhdError(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("iv", _I3),
                ("residual", ctypes.c_double)]


# Every entry point of include/grhd_abi.h, with its ctypes signature.
_P = ctypes.POINTER
_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int
SIGNATURES = {
    "grhd_ctx_create": (_i, [_P(_vp)]),
    "grhd_ctx_destroy": (None, [_vp]),
    "grhd_compute_rhs_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                  _P(GrhdBackground), _d, _vp, _i, _i, _vp,
                                  _vp, _P(GrhdError), _vp]),
    "grhd_recover_primitives_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                         _P(GrhdBackground), _d, _vp, _i, _i,
                                         _vp, _P(GrhdDiagnostics), _vp]),
    "grhd_update_conserved_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                       _P(GrhdBackground), _vp, _i, _i, _vp,
                                       _d, _i, _d, _vp, _P(GrhdDiagnostics),
                                       _vp]),
    "grhd_compute_max_inverse_dt_box": (_i, [_vp, _P(GrhdBox),
                                             _P(GrhdParams),
                                             _P(GrhdBackground), _d, _vp, _i,
                                             _i, _P(_d), _P(GrhdError), _vp]),
    "grhd_compute_stable_dt_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                        _P(GrhdBackground), _d, _vp, _i, _i,
                                        _d, _P(_d), _P(GrhdError), _vp]),
    "grhd_advance_ssprk2_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                     _P(GrhdBackground), _d, _d, _vp, _i, _i,
                                     _vp, _P(GrhdDiagnostics), _P(GrhdError),
                                     _vp]),
    "grhd_advance_rk4_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                  _P(GrhdBackground), _d, _d, _i, _vp, _i, _i,
                                  _vp, _P(GrhdDiagnostics), _P(GrhdError),
                                  _vp]),
    "grhd_advance_rk4_box_noalloc": (_i, [_P(GrhdBox), _P(GrhdParams),
                                          _P(GrhdBackground), _d, _d, _i,
                                          _vp, _i, _i, _vp,
                                          _P(GrhdDiagnostics), _P(GrhdError),
                                          _vp]),
    "grhd_cell_geometry_box": (_i, [_vp, _P(GrhdBox), _P(GrhdBackground), _d,
                                    _vp, _vp]),
    "grhd_fp64_div_sqrt": (_i, [_vp, _vp, _vp, _vp, ctypes.c_longlong, _vp]),
    "grhd_level_compute_rhs": (_i, [_vp, _P(GrhdLevelBox), _i,
                                    _P(GrhdParams), _P(GrhdBackground), _d,
                                    _vp, _i, _i, _vp, _vp, _P(GrhdError),
                                    _vp]),
    "grhd_level_add_conserved_rhs": (_i, [_vp, _P(GrhdLevelBox), _i,
                                          _P(GrhdParams), _P(GrhdBackground),
                                          _d, _vp, _i, _i, _vp,
                                          _P(GrhdError), _vp]),
    "grhd_level_recover_primitives": (_i, [_vp, _P(GrhdLevelBox), _i,
                                           _P(GrhdParams), _P(GrhdBackground),
                                           _d, _vp, _i, _i, _vp,
                                           _P(GrhdDiagnostics), _vp]),
    "grhd_level_compute_max_inverse_dt": (_i, [_vp, _P(GrhdLevelBox), _i,
                                               _P(GrhdParams),
                                               _P(GrhdBackground), _d, _vp,
                                               _i, _i, _P(_d), _P(GrhdError),
                                               _vp]),
    "grhd_level_update_conserved": (_i, [_vp, _P(GrhdLevelBox), _i,
                                         _P(GrhdParams), _P(GrhdBackground),
                                         _vp, _i, _i, _vp, _d, _i, _d, _vp,
                                         _P(GrhdDiagnostics), _vp]),
    "grhd_level_conserved_sum": (_i, [_vp, _P(GrhdLevelBox), _i, _vp, _i,
                                      _i, _P(_d), _vp]),
    "grhd_level_conserved_state_is_finite": (_i, [_vp, _P(GrhdLevelBox), _i,
                                                  _vp, _i, _i, _P(_i),
                                                  _vp]),
    "grhd_rk4_adjoint_box": (_i, [_vp, _P(GrhdBox), _P(GrhdParams),
                                  _P(GrhdBackground), _d, _i, _vp, _i, _i,
                                  _d, _i, _i, _P(_d), _vp, _P(_d),
                                  _P(GrhdError), _vp]),
    "grhd_level_checkpoint_pack": (_i, [_vp, _P(GrhdLevelBox), _i, _vp, _i,
                                        _P(_i), _P(_i), _i, _vp, _vp]),
    "grhd_level_checkpoint_unpack": (_i, [_vp, _P(GrhdLevelBox), _i, _vp,
                                          _i, _P(_i), _P(_i), _i, _vp, _vp]),
    "grhd_level_advance_ssprk2": (_i, [_vp, _P(GrhdLevelBox), _i,
                                       _P(GrhdParams), _P(GrhdBackground), _d,
                                       _d, _vp, _i, _i, _vp,
                                       _P(GrhdDiagnostics), _P(GrhdError),
                                       _vp]),
    "grhd_level_advance_rk4": (_i, [_vp, _P(GrhdLevelBox), _i,
                                    _P(GrhdParams), _P(GrhdBackground), _d,
                                    _d, _i, _vp, _i, _i, _vp,
                                    _P(GrhdDiagnostics), _P(GrhdError), _vp]),
    "grhd_ctx_set_timing": (_i, [_vp, _i]),
    "grhd_ctx_get_timing": (_i, [_vp, _P(_d), _P(ctypes.c_longlong)]),
    "grhd_ctx_set_autotune": (_i, [_vp, _i]),
    "grhd_ctx_set_face_launch": (_i, [_vp, _i, _i, _i]),
    "grhd_ctx_get_face_launch": (_i, [_vp, _i, _P(_i), _P(_i), _P(_i),
                                      _P(_d)]),
}

KERNEL_CLASSES = ("prims", "faces", "cells", "other")

_lib = None


def load_library(path=None):
    """Load the HIP build of the operator.  Raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "grdzhadzha_amd: HIP library %s is not built; run "
            "__graft_entry__.build() (hipcc --offload-arch=gfx950)" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
