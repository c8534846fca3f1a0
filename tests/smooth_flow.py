"""A smooth, supersonic, floor-free flow for temporal-order checks.

Every cell (ghosts included) holds the primitives
    rho = 1 + a sin(2 pi x / L) cos(2 pi y / L) + a sin(2 pi z / L),
    P   = 0.01 rho^(4/3),  eps = P / ((Gamma - 1) rho),
    v   = (0.3 + 0.02 cos(2 pi y / L), 0.3 + 0.02 sin(2 pi z / L), 0.3)
and the conserved variables prim2con gives on the cell's metric.  The
velocity is supersonic along every axis (each characteristic speed is
positive on every face), so the HLLE flux is always the upwind left flux;
without reconstruction the right-hand side is then a smooth function of
the state.  Densities sit far above the floors and the velocity far below
the v^2 cap, so no floor, reset or atmosphere branch fires: the time
integrator's truncation error is the only error left.
"""
import numpy as np

import oracle
from grdzhadzha_amd.options import (AtmosphereOptions, BoxSpec,
                                    CircularBinaryPNParams, KerrSchildParams,
                                    OperatorOptions, RecoveryOptions)

WAVELENGTH = 4.0
AMPLITUDE = 0.05


def options(time_dependent=False):
    """No reconstruction (no limiter branches), Gamma = 4/3, a recovery
    tolerance of 1e-14 so that con2prim's stopping rule adds no noise
    above the temporal errors being measured."""
    return OperatorOptions(
        adiabatic_index=4.0 / 3.0, use_reconstruction=False,
        use_metric_volume_time_sources=time_dependent,
        recovery=RecoveryOptions(tolerance=1.0e-14),
        atmosphere=AtmosphereOptions(rho_floor=1.0e-10,
                                     pressure_floor=1.0e-12))


def cases():
    """(name, background, box, time_dependent): a static Kerr-Schild
    a = 0.5 box about (-24, -24, -24) M, and a box near a circular PN binary
    (separation 6 M, orbital period ~92 M) whose metric changes in time, so
    stage times enter the result."""
    return [
        ("kerr_schild", KerrSchildParams(1.0, 0.5),
         BoxSpec((0, 0, 0), (7, 7, 7), 3, 0.5, (24.0, 24.0, 24.0)), False),
        ("binary_pn", CircularBinaryPNParams(0.5, 0.5, 6.0,
                                             softening_radius=0.5),
         BoxSpec((0, 0, 0), (7, 7, 7), 3, 0.5, (-8.0, 2.0, 2.0)), True),
    ]


def state(box, background, gamma=4.0 / 3.0, ncomp=11):
    nz, ny, nx = box.storage_shape
    g = box.ghosts
    x = (np.arange(nx) + box.lo[0] - g + 0.5) * box.dx - box.center[0]
    y = (np.arange(ny) + box.lo[1] - g + 0.5) * box.dx - box.center[1]
    z = (np.arange(nz) + box.lo[2] - g + 0.5) * box.dx - box.center[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    k = 2.0 * np.pi / WAVELENGTH
    rho = (1.0 + AMPLITUDE * np.sin(k * X) * np.cos(k * Y)
           + AMPLITUDE * np.sin(k * Z))
    P = 0.01 * rho ** (4.0 / 3.0)
    eps = P / ((gamma - 1.0) * rho)
    v1 = 0.3 + 0.02 * np.cos(k * Y)
    v2 = 0.3 + 0.02 * np.sin(k * Z)
    geo = oracle.Oracle(background, box).cell_geometry()
    st = np.zeros((ncomp,) + box.storage_shape)
    for idx in np.ndindex(*box.storage_shape):
        g6 = geo[(slice(0, 6),) + idx]
        gm = np.array([[g6[0], g6[1], g6[2]], [g6[1], g6[3], g6[4]],
                       [g6[2], g6[4], g6[5]]])
        prim = [rho[idx], eps[idx], P[idx], v1[idx], v2[idx], 0.3]
        st[(slice(0, 5),) + idx] = oracle.prim2con(prim, gamma, gm)
        st[(slice(5, 11),) + idx] = prim
    return st


def observed_orders(results, ref):
    """log2 of successive error ratios (max |U - U_ref| over the interior
    conserved components), for results at dt, dt/2, dt/4, ..."""
    errs = [float(np.max(np.abs(r - ref))) for r in results]
    return errs, [float(np.log2(errs[i] / errs[i + 1]))
                  for i in range(len(errs) - 1)]
