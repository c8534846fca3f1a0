"""Option and geometry records mirroring the reference's C++ structs.

AtmosphereOptions   Source/Matter/GRHD/Atmosphere.hpp:16-22
RecoveryOptions     Source/Matter/GRHD/Valencia.hpp:17-23
KerrSchildParams    Source/Background/KerrSchild.hpp:26-31
CircularBinaryPNParams  Source/Background/CircularBinaryPN.hpp:32-42
BoxSpec             one Chombo box of a level: interior Box, FAB storage box
                    (interior grown by `ghosts`), ProblemDomain, dx and the
                    Coordinates center.
Defaults are the reference's defaults.
"""
from dataclasses import dataclass, field
from typing import Sequence, Tuple

from . import abi


@dataclass
class AtmosphereOptions:
    rho_floor: float = 1.0e-12
    pressure_floor: float = 1.0e-12
    max_velocity_squared: float = 1.0 - 1.0e-12
    metric_source_atmosphere_factor: float = 1.0 + 1.0e-12


@dataclass
class RecoveryOptions:
    max_iterations: int = 50
    tolerance: float = 1.0e-12
    pressure_floor: float = 1.0e-12
    max_velocity_squared: float = 1.0 - 1.0e-12


@dataclass
class KerrSchildParams:
    mass: float = 1.0
    spin: float = 0.0

    def to_c(self):
        b = abi.GrhdBackground()
        b.kind = abi.GRHD_BG_KERR_SCHILD
        b.mass = self.mass
        b.spin = self.spin
        return b


@dataclass
class CircularBinaryPNParams:
    mass_1: float = 0.5
    mass_2: float = 0.5
    separation: float = 10.0
    angular_velocity: float = 0.0
    phase: float = 0.0
    time: float = 0.0
    softening_radius: float = 0.5

    def to_c(self):
        b = abi.GrhdBackground()
        b.kind = abi.GRHD_BG_CIRCULAR_BINARY_PN
        b.mass_1 = self.mass_1
        b.mass_2 = self.mass_2
        b.separation = self.separation
        b.angular_velocity = self.angular_velocity
        b.phase = self.phase
        b.time = self.time
        b.softening_radius = self.softening_radius
        return b


@dataclass
class OperatorOptions:
    """Everything the level operator's constructor takes besides the
    background (FixedBGFArrayBoxFiniteVolume.hpp:450-464, :681-699)."""
    adiabatic_index: float = 5.0 / 3.0
    atmosphere: AtmosphereOptions = field(default_factory=AtmosphereOptions)
    recovery: RecoveryOptions = field(default_factory=RecoveryOptions)
    limiter_theta: float = 1.5
    use_reconstruction: bool = True
    use_static_metric_sources: bool = True
    use_metric_volume_time_sources: bool = False
    metric_time_derivative_step: float = 1.0e-3
    # no counterpart in the reference: FP32 reconstruction, face metric and
    # HLLE, and static-metric sources; recovery, the divergence sum and the
    # update FP64 (False: the bit-exact FP64 path)
    mixed_precision: bool = False
    # no counterpart in the reference: with mixed_precision, a margin m >= 1
    # computes in FP64 the faces beside cells with rho <= m rho_floor,
    # P <= m pressure_floor, 1 - v^2 <= m (1 - max_v2) or a failed recovery
    # (0: every face FP32)
    mixed_fp64_margin: float = 0.0
    # no counterpart in the reference: with mixed_precision, s > 0 also
    # computes in FP64 the faces beside cells whose shift has
    # gamma_ij beta^i beta^j >= s^2 (0: no shift test)
    mixed_fp64_shift: float = 0.0
    # KerrSchild::check_if_excised (KerrSchild.hpp:247-280) with
    # ExcisionTest's zero right-hand side (ExcisionTest.hpp:40-63) inside
    # the buffered horizons; Kerr-Schild backgrounds only
    excise_horizon: bool = False
    excision_inner_buffer: float = 1.05
    excision_outer_buffer: float = 0.9

    def to_c(self):
        p = abi.GrhdParams()
        p.adiabatic_index = self.adiabatic_index
        p.rho_floor = self.atmosphere.rho_floor
        p.pressure_floor = self.atmosphere.pressure_floor
        p.max_velocity_squared = self.atmosphere.max_velocity_squared
        p.metric_source_atmosphere_factor = \
            self.atmosphere.metric_source_atmosphere_factor
        p.max_iterations = self.recovery.max_iterations
        p.tolerance = self.recovery.tolerance
        p.recovery_pressure_floor = self.recovery.pressure_floor
        p.recovery_max_velocity_squared = self.recovery.max_velocity_squared
        p.limiter_theta = self.limiter_theta
        p.use_reconstruction = int(self.use_reconstruction)
        p.use_static_metric_sources = int(self.use_static_metric_sources)
        p.use_metric_volume_time_sources = int(
            self.use_metric_volume_time_sources)
        p.metric_time_derivative_step = self.metric_time_derivative_step
        p.mixed_precision = int(self.mixed_precision)
        p.excise_horizon = int(self.excise_horizon)
        p.excision_inner_buffer = self.excision_inner_buffer
        p.excision_outer_buffer = self.excision_outer_buffer
        p.mixed_fp64_margin = self.mixed_fp64_margin
        p.mixed_fp64_shift = self.mixed_fp64_shift
        return p


@dataclass
class BoxSpec:
    lo: Tuple[int, int, int]
    hi: Tuple[int, int, int]
    ghosts: int
    dx: float
    center: Tuple[float, float, float]
    domain_lo: Tuple[int, int, int] = None
    domain_hi: Tuple[int, int, int] = None
    periodic: Tuple[int, int, int] = (0, 0, 0)
    # inverse_dx_U of the reference's operator overloads: the divergence
    # and CFL scale of each direction (None: 1/dx in every direction)
    inverse_dx: Tuple[float, float, float] = None

    def __post_init__(self):
        self.lo = tuple(int(v) for v in self.lo)
        self.hi = tuple(int(v) for v in self.hi)
        if self.domain_lo is None:
            self.domain_lo = self.lo
        if self.domain_hi is None:
            self.domain_hi = self.hi

    @property
    def glo(self):
        return tuple(v - self.ghosts for v in self.lo)

    @property
    def ghi(self):
        return tuple(v + self.ghosts for v in self.hi)

    @property
    def interior_shape(self):
        """numpy/torch shape (nz, ny, nx) of the interior box"""
        return tuple(self.hi[d] - self.lo[d] + 1 for d in (2, 1, 0))

    @property
    def storage_shape(self):
        """numpy/torch shape (nz, ny, nx) of the FAB storage box"""
        return tuple(self.hi[d] - self.lo[d] + 1 + 2 * self.ghosts
                     for d in (2, 1, 0))

    @property
    def num_interior(self):
        nz, ny, nx = self.interior_shape
        return nz * ny * nx

    def interior_slices(self):
        g = self.ghosts
        return tuple(slice(g, g + n) for n in self.interior_shape)

    def to_c(self):
        b = abi.GrhdBox()
        for d in range(3):
            b.lo[d] = self.lo[d]
            b.hi[d] = self.hi[d]
            b.glo[d] = self.lo[d] - self.ghosts
            b.ghi[d] = self.hi[d] + self.ghosts
            b.dlo[d] = self.domain_lo[d]
            b.dhi[d] = self.domain_hi[d]
            b.periodic[d] = int(self.periodic[d])
            b.center[d] = self.center[d]
            b.inverse_dx[d] = (0.0 if self.inverse_dx is None
                               else float(self.inverse_dx[d]))
        b.dx = self.dx
        return b


def cube_box(n: int, ghosts: int, half_width: float) -> BoxSpec:
    """Box [0, n-1]^3 covering [-half_width, half_width]^3 (domain = box):
    dx = 2 half_width / n and center = (half_width,)*3, so cell centres are
    x = (i + 1/2) dx - half_width."""
    dx = 2.0 * half_width / n
    return BoxSpec(lo=(0, 0, 0), hi=(n - 1,) * 3, ghosts=ghosts, dx=dx,
                   center=(half_width,) * 3)
