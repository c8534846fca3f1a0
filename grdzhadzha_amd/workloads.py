"""Synthetic workloads: seeded initial states for the benchmark configurations.

Per-cell noise xi = U(-1, 1) comes from SplitMix64(seed XOR linear index of
the cell in the FAB storage box), so the same state is produced on any
machine.  Cell centres follow the reference's Coordinates convention
x = (i + 1/2) dx - center.  Primitives pass through the atmosphere floors
(Source/Matter/GRHD/Atmosphere.hpp:24-46) on the cell's metric and are
converted to conserved variables (Valencia.hpp:55-73); both primitive and
conserved components are stored, as the reference's drivers do
(Tests/GRHDLevelDataMichelAccretion/GRHDLevelDataMichelAccretion.cpp:403-436).

The workloads are host-side numpy (inputs are built once, outside any timed
region) and use a vectorised Kerr-Schild metric.

  michel_box_32   32^3 + 3 ghosts on [-8, 8]^3, Schwarzschild (KS, M = 1),
                  Michel-like inflow, Gamma = 4/3, MC limiter (theta = 2).
  torus_box(n)    n^3 + 3 ghosts on [-n/4, n/4]^3 (dx = 0.5), Kerr a = 0.9,
                  Gaussian torus at R = 12, Keplerian Omega = 1/(R^1.5 + a),
                  Gamma = 4/3, minmod limiter; n = 128 is the 128^3 case.
  binary_pn_levels  3 refinement levels x 180 boxes of seeded uneven sizes
                  on the circular binary PN background, circumbinary ring.
"""
import numpy as np

from .options import (AtmosphereOptions, BoxSpec, KerrSchildParams,
                      OperatorOptions, RecoveryOptions, cube_box)

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser of uint64 array x (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed, n):
    """xi_i = U(-1, 1) for linear indices i = 0..n-1."""
    idx = np.arange(n, dtype=np.uint64) ^ np.uint64(seed)
    z = splitmix64(idx)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def cell_centres(box: BoxSpec):
    """x, y, z arrays of shape storage_shape (nz, ny, nx)."""
    nz, ny, nx = box.storage_shape
    glo = box.glo
    i = np.arange(nx) + glo[0]
    j = np.arange(ny) + glo[1]
    k = np.arange(nz) + glo[2]
    x = (i + 0.5) * box.dx - box.center[0]
    y = (j + 0.5) * box.dx - box.center[1]
    z = (k + 0.5) * box.dx - box.center[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def kerr_schild_metric(M, a, X, Y, Z):
    """gamma_ij (3, 3, ...), lapse, shift^i (3, ...) of Kerr-Schild
    (Source/Background/KerrSchild.hpp:65-116)."""
    a2 = a * a
    rho = np.maximum(np.sqrt(X * X + Y * Y + Z * Z), 1e-6)
    rho2 = rho * rho
    r2 = 0.5 * (rho2 - a2) + np.sqrt(0.25 * (rho2 - a2) ** 2 + a2 * Z * Z)
    r = np.sqrt(r2)
    cos_t = Z / r
    H = M * r / (r2 + a2 * cos_t * cos_t)
    el = np.stack([(r * X + a * Y) / (r2 + a2), (r * Y - a * X) / (r2 + a2),
                   Z / r])
    g = np.empty((3, 3) + X.shape)
    for i in range(3):
        for j in range(3):
            g[i, j] = (1.0 if i == j else 0.0) + 2.0 * H * el[i] * el[j]
    lapse = 1.0 / np.sqrt(1.0 + 2.0 * H)
    gu = np.linalg.inv(np.moveaxis(g, (0, 1), (-2, -1)))
    gu = np.moveaxis(gu, (-2, -1), (0, 1))
    shift = np.einsum("ij...,j...->i...", gu, 2.0 * H * el)
    return g, lapse, shift


def apply_floors(rho, eps, p, v, g, gamma, atm: AtmosphereOptions):
    rho = np.maximum(rho, atm.rho_floor)
    p = np.maximum(p, atm.pressure_floor)
    v2 = np.einsum("ij...,i...,j...->...", g, v, v)
    scale = np.where(v2 >= atm.max_velocity_squared,
                     np.sqrt(atm.max_velocity_squared / np.maximum(v2, 1e-300)),
                     1.0)
    v = v * scale
    eps = p / ((gamma - 1.0) * rho)
    return rho, eps, p, v


def prim_to_cons(rho, eps, p, v, g, gamma):
    vL = np.einsum("ij...,j...->i...", g, v)
    v2 = np.einsum("i...,i...->...", vL, v)
    W = 1.0 / np.sqrt(1.0 - v2)
    h = 1.0 + eps + p / rho
    rhW2 = rho * h * W * W
    D = rho * W
    tau = rhW2 - p - D
    S = rhW2 * vL
    return D, S, tau


def pack_state(box, rho, eps, p, v, D, S, tau, ncomp=11, comp_base=0):
    st = np.zeros((ncomp,) + box.storage_shape)
    cb = comp_base
    st[cb + 0] = D
    st[cb + 1:cb + 4] = S
    st[cb + 4] = tau
    st[cb + 5] = rho
    st[cb + 6] = eps
    st[cb + 7] = p
    st[cb + 8:cb + 11] = v
    return st


class Workload:
    """A box, its background, operator options, an initial state, and the
    step size / count the configuration prescribes."""

    def __init__(self, name, box, background, options, state, dt, nsteps):
        self.name = name
        self.box = box
        self.background = background
        self.options = options
        self.state = state
        self.dt = dt
        self.nsteps = nsteps


def michel_box_32(seed=1, n=32, ghosts=3, half_width=8.0):
    """Plain FP64 one-box Michel-like case: 32^3 on [-8,8]^3 (dx = 0.5 M),
    Schwarzschild M = 1 in Kerr-Schild coordinates, Gamma = 4/3,
    rho0 = 1e-3 (r/2M)^-1.5 (1 + 0.05 xi), P = 0.1 rho0^(4/3),
    v^r = -0.9 sqrt(2M/r), v^i = v^r x^i / r, atmosphere (rho 1e-10,
    P 1e-12) for r < 1.5 M, MC limiter, one RK4 step with dt = 0.25 dx.
    The radial seed is superluminal for 1.5 M < r < ~2.8 M in these
    coordinates, so it passes through the floors with max v^2 = 0.8 (the
    value the reference's Michel test uses)."""
    box = cube_box(n, ghosts, half_width)
    M = 1.0
    gamma = 4.0 / 3.0
    atm = AtmosphereOptions(rho_floor=1e-10, pressure_floor=1e-12,
                            max_velocity_squared=0.8)
    rec = RecoveryOptions(max_iterations=100, tolerance=1e-12,
                          pressure_floor=1e-12,
                          max_velocity_squared=1.0 - 1.0e-12)
    opts = OperatorOptions(adiabatic_index=gamma, atmosphere=atm,
                           recovery=rec, limiter_theta=2.0)
    X, Y, Z = cell_centres(box)
    r = np.sqrt(X * X + Y * Y + Z * Z)
    xi = uniform_pm1(seed, X.size).reshape(X.shape)
    rho = 1e-3 * (r / (2 * M)) ** -1.5 * (1.0 + 0.05 * xi)
    p = 0.1 * rho ** (4.0 / 3.0)
    vr = -0.9 * np.sqrt(2 * M / r)
    v = np.stack([vr * X / r, vr * Y / r, vr * Z / r])
    atmos = r < 1.5 * M
    rho = np.where(atmos, atm.rho_floor, rho)
    p = np.where(atmos, atm.pressure_floor, p)
    v = np.where(atmos[None], 0.0, v)
    eps = p / ((gamma - 1.0) * rho)
    g, _, _ = kerr_schild_metric(M, 0.0, X, Y, Z)
    rho, eps, p, v = apply_floors(rho, eps, p, v, g, gamma, atm)
    D, S, tau = prim_to_cons(rho, eps, p, v, g, gamma)
    state = pack_state(box, rho, eps, p, v, D, S, tau)
    return Workload("michel_32^3_ks_a0", box, KerrSchildParams(M, 0.0), opts,
                    state, 0.25 * box.dx, 1)


def torus_box(n=128, seed=2, ghosts=3, spin=0.9, R0=12.0, sigma=3.0,
              nsteps=100):
    """Plain FP64 at scale: n^3 box (+3 ghosts) with dx = 0.5 M on
    [-n/4, n/4]^3 (n = 128: [-32, 32]^3), Kerr a = 0.9 M in Kerr-Schild
    coordinates, rho0 = 1e-2 exp(-((R-12)^2 + z^2) / (2 3^2)) (1 + 0.01 xi)
    floored at 1e-10, P = 0.01 rho0^(4/3), Gamma = 4/3, Keplerian
    Omega = 1/(R^1.5 + a) with Eulerian v^i = (Omega (-y, x, 0) + beta^i) /
    alpha, minmod limiter, RK4 with dt = 0.25 dx.  Velocities are capped at
    v^2 = 0.95 by the floors (the fluid inside the horizon cannot be
    rotating at the seed's coordinate velocity)."""
    half = 0.25 * n
    box = cube_box(n, ghosts, half)
    M = 1.0
    opts, state = torus_state(box, M, spin, R0, sigma, seed)
    return Workload("torus_%d^3_kerr_a%.1f" % (n, spin), box,
                    KerrSchildParams(M, spin), opts, state, 0.25 * box.dx,
                    nsteps)


def binary_pn_levels(seed=3, nlevels=3, nbox=180, base_cells=64,
                     base_dx=2.0, sides=(8, 16, 24, 32, 48, 64), ghosts=3,
                     ring_R=40.0, ring_sigma=8.0, nsteps=50):
    """Batched box set: circular binary PN background (m1 = m2 = 0.5,
    separation 20 M, circular PN orbit), `nlevels` refinement levels
    (ratio 2) of `nbox` boxes each.  Level l has dx = base_dx / 2^l and an
    index region of base_cells^3 cells centred on the origin; each box's
    side lengths are drawn independently per axis from `sides` and its
    corner uniformly in that region (SplitMix64 stream `seed`).  Every box
    is its own ProblemDomain (outflow ghosts).  Circumbinary ring
    rho0 = 1e-2 exp(-((R-40)^2 + z^2) / (2 8^2)) floored at 1e-10,
    P = 0.01 rho0^(4/3), Keplerian Omega = (M/R^3)^(1/2) about the total
    mass, Eulerian v^i = (Omega (-y, x, 0) + beta^i) / alpha, Gamma = 4/3,
    MC limiter (theta = 2), volume-time sources on.  Returns
    (background, options, levels) with levels = [(boxes, fabs, dt)],
    dt = 0.25 dx of the level."""
    from .options import CircularBinaryPNParams
    bg = CircularBinaryPNParams(mass_1=0.5, mass_2=0.5, separation=20.0)
    gamma = 4.0 / 3.0
    atm = AtmosphereOptions(rho_floor=1e-10, pressure_floor=1e-12,
                            max_velocity_squared=0.95)
    rec = RecoveryOptions(max_iterations=100, tolerance=1e-12,
                          pressure_floor=1e-12,
                          max_velocity_squared=1.0 - 1.0e-12)
    opts = OperatorOptions(adiabatic_index=gamma, atmosphere=atm,
                           recovery=rec, limiter_theta=2.0,
                           use_metric_volume_time_sources=True,
                           metric_time_derivative_step=1.0e-3)
    draws = splitmix64(np.arange(nlevels * nbox * 6, dtype=np.uint64) ^
                       np.uint64(seed))
    draws = draws.reshape(nlevels, nbox, 6)
    levels = []
    for lev in range(nlevels):
        dx = base_dx / (2 ** lev)
        half = 0.5 * base_cells * dx
        boxes, fabs = [], []
        for b in range(nbox):
            side = [sides[int(draws[lev, b, d] % np.uint64(len(sides)))]
                    for d in range(3)]
            lo = [int(draws[lev, b, 3 + d] %
                      np.uint64(base_cells - side[d] + 1)) for d in range(3)]
            hi = [lo[d] + side[d] - 1 for d in range(3)]
            box = BoxSpec(tuple(lo), tuple(hi), ghosts, dx, (half,) * 3)
            X, Y, Z = cell_centres(box)
            fabs.append(_pn_ring_state(box, bg, X, Y, Z, gamma, atm,
                                       ring_R, ring_sigma))
            boxes.append(box)
        levels.append((boxes, fabs, 0.25 * dx))
    return bg, opts, levels, nsteps


def binary_pn_metric(bg, X, Y, Z, time=0.0):
    """gamma_ij, lapse, shift^i of the softened circular binary
    (Source/Background/CircularBinaryPN.hpp:70-130, 209-297), vectorised."""
    M = bg.mass_1 + bg.mass_2
    om = bg.angular_velocity if bg.angular_velocity != 0.0 else \
        np.sqrt(M / bg.separation ** 3)
    ph = bg.phase + om * (bg.time + time)
    r1 = bg.mass_2 / M * bg.separation
    r2 = bg.mass_1 / M * bg.separation
    srcs = [(bg.mass_1, (r1 * np.cos(ph), r1 * np.sin(ph)),
             (-om * r1 * np.sin(ph), om * r1 * np.cos(ph))),
            (bg.mass_2, (-r2 * np.cos(ph), -r2 * np.sin(ph)),
             (om * r2 * np.sin(ph), -om * r2 * np.cos(ph)))]
    U = np.zeros_like(X)
    V = np.zeros((3,) + X.shape)
    for m, (px, py), (vx, vy) in srcs:
        R = np.sqrt(bg.softening_radius ** 2 + (X - px) ** 2 + (Y - py) ** 2 +
                    Z ** 2)
        U += m / R
        V[0] += m * vx / R
        V[1] += m * vy / R
    cf = 1.0 + 2.0 * U
    g = np.zeros((3, 3) + X.shape)
    for i in range(3):
        g[i, i] = cf
    lapse = np.sqrt(1.0 - 2.0 * U + 2.0 * U * U +
                    16.0 * np.einsum("i...,i...->...", V, V) / cf)
    shift = -4.0 * V / cf
    return g, lapse, shift


def _pn_ring_state(box, bg, X, Y, Z, gamma, atm, ring_R, ring_sigma):
    R = np.sqrt(X * X + Y * Y)
    rho = 1e-2 * np.exp(-((R - ring_R) ** 2 + Z * Z) /
                        (2.0 * ring_sigma ** 2))
    rho = np.maximum(rho, 1e-10)
    p = 0.01 * rho ** (4.0 / 3.0)
    M = bg.mass_1 + bg.mass_2
    omega = np.sqrt(M / np.maximum(R, 1e-6) ** 3)
    g, lapse, shift = binary_pn_metric(bg, X, Y, Z)
    V = np.stack([-omega * Y, omega * X, np.zeros_like(X)])
    v = (V + shift) / lapse
    eps = p / ((gamma - 1.0) * rho)
    rho, eps, p, v = apply_floors(rho, eps, p, v, g, gamma, atm)
    D, S, tau = prim_to_cons(rho, eps, p, v, g, gamma)
    return pack_state(box, rho, eps, p, v, D, S, tau)


def torus_state(box, M=1.0, spin=0.9, R0=12.0, sigma=3.0, seed=2):
    """Options and state FAB of torus_box's torus on any box."""
    gamma = 4.0 / 3.0
    atm = AtmosphereOptions(rho_floor=1e-10, pressure_floor=1e-12,
                            max_velocity_squared=0.95)
    rec = RecoveryOptions(max_iterations=100, tolerance=1e-12,
                          pressure_floor=1e-12,
                          max_velocity_squared=1.0 - 1.0e-12)
    opts = OperatorOptions(adiabatic_index=gamma, atmosphere=atm,
                           recovery=rec, limiter_theta=1.0)
    X, Y, Z = cell_centres(box)
    R = np.sqrt(X * X + Y * Y)
    xi = uniform_pm1(seed, X.size).reshape(X.shape)
    rho = 1e-2 * np.exp(-((R - R0) ** 2 + Z * Z) / (2 * sigma ** 2)) * \
        (1.0 + 0.01 * xi)
    rho = np.maximum(rho, 1e-10)
    p = 0.01 * rho ** (4.0 / 3.0)
    omega = 1.0 / (R ** 1.5 + spin)
    g, lapse, shift = kerr_schild_metric(M, spin, X, Y, Z)
    V = np.stack([-omega * Y, omega * X, np.zeros_like(X)])
    v = (V + shift) / lapse
    eps = p / ((gamma - 1.0) * rho)
    rho, eps, p, v = apply_floors(rho, eps, p, v, g, gamma, atm)
    D, S, tau = prim_to_cons(rho, eps, p, v, g, gamma)
    state = pack_state(box, rho, eps, p, v, D, S, tau)
    return opts, state


def mass_accretion_rate(box, state, M, spin, radius):
    """Rest-mass accretion rate through the coordinate sphere |x| = radius:
    minus the outward D flux sqrt(gamma) D (alpha v^i - beta^i) x_i/|x|
    (the Valencia flux of D, Valencia.hpp:230-249) summed over the interior
    cells with ||x| - radius| < dx/2, times dx^3, divided by the shell width
    dx."""
    X, Y, Z = cell_centres(box)
    sl = box.interior_slices()
    X, Y, Z = X[sl], Y[sl], Z[sl]
    rr = np.sqrt(X * X + Y * Y + Z * Z)
    sel = np.abs(rr - radius) < 0.5 * box.dx
    x = np.stack([X[sel], Y[sel], Z[sel]])
    g, lapse, shift = kerr_schild_metric(M, spin, x[0], x[1], x[2])
    det = (g[0, 0] * (g[1, 1] * g[2, 2] - g[1, 2] * g[2, 1]) -
           g[0, 1] * (g[1, 0] * g[2, 2] - g[1, 2] * g[2, 0]) +
           g[0, 2] * (g[1, 0] * g[2, 1] - g[1, 1] * g[2, 0]))
    D = state[0][sl][sel]
    v = np.stack([state[8 + i][sl][sel] for i in range(3)])
    flux = D * (lapse * v - shift)
    radial = np.einsum("i...,i...->...", flux, x) / rr[sel]
    return -float(np.sum(np.sqrt(det) * radial)) * box.dx ** 2
