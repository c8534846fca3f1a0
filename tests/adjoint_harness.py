"""Test harness of the discrete adjoint (grhd_rk4_adjoint_box).

`RK4Objective` is a torch.autograd.Function around the C-ABI adjoint:
forward runs the RK4 evolution and returns J; backward hands back
dJ/dU0, dJ/dGamma and dJ/da.  `michel_inflow_U0` builds the initial
conserved state from the initial rest-mass density field in torch, so that
autograd carries the gradient from U0 back to rho0 and adds the initial
data's own dependence on Gamma and a:

  rho0 = rho0 field (interior and ghosts), P = 0.1 rho0^(4/3),
  v^i = v^r x^i / r with v^r = -0.9 sqrt(2M/r), atmosphere (rho 1e-10,
  P 1e-12, v = 0) for r < 1.5 M, floors (rho, P, v^2 <= max_v2) on the
  Kerr-Schild metric of spin a, eps = P / ((Gamma - 1) rho), then
  D = rho W, S_i = rho h W^2 v_i, tau = rho h W^2 - P - D.
"""
import math

import numpy as np
import torch

from grdzhadzha_amd import workloads
from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator
from grdzhadzha_amd.options import (AtmosphereOptions, BoxSpec,
                                    KerrSchildParams, OperatorOptions,
                                    RecoveryOptions)

M = 1.0


def options(gamma, max_v2=0.8, theta=2.0, tolerance=1e-18, excise=False):
    """Michel-like inflow options.  The recovery tolerance (absolute, on the
    pressure residual) defaults to 1e-18: pressures here are ~1e-6, so the
    reference's 1e-12 leaves ~1e-7 relative iteration noise in the
    recovered pressure -- as large as what an h = 1e-6 perturbation of tau
    changes -- and central differences would measure that noise, not the
    derivative of the root that the adjoint (implicit-function theorem)
    returns."""
    atm = AtmosphereOptions(rho_floor=1e-10, pressure_floor=1e-12,
                            max_velocity_squared=max_v2)
    rec = RecoveryOptions(max_iterations=100, tolerance=tolerance,
                          pressure_floor=1e-12,
                          max_velocity_squared=1.0 - 1.0e-12)
    return OperatorOptions(adiabatic_index=float(gamma), atmosphere=atm,
                           recovery=rec, limiter_theta=theta,
                           excise_horizon=excise)


def ks_metric_torch(a, X, Y, Z):
    """gamma_ij (3, 3, ...) of Kerr-Schild with spin tensor a (autograd)."""
    a2 = a * a
    rho = torch.clamp(torch.sqrt(X * X + Y * Y + Z * Z), min=1e-6)
    rho2 = rho * rho
    r2 = 0.5 * (rho2 - a2) + torch.sqrt(0.25 * (rho2 - a2) ** 2 + a2 * Z * Z)
    r = torch.sqrt(r2)
    cos_t = Z / r
    H = M * r / (r2 + a2 * cos_t * cos_t)
    el = [(r * X + a * Y) / (r2 + a2), (r * Y - a * X) / (r2 + a2), Z / r]
    return [[(1.0 if i == j else 0.0) + 2.0 * H * el[i] * el[j]
             for j in range(3)] for i in range(3)]


def michel_inflow_U0(box, rho0, gamma, spin, opts, init_max_v2=None):
    """Conserved initial state (5, storage shape) as a differentiable
    function of rho0 (storage-shape tensor), gamma and spin (0-d).  The
    velocity clamp of the initial data is init_max_v2 (default: the
    atmosphere's max_velocity_squared)."""
    X, Y, Z = [torch.from_numpy(c).to(rho0) for c in
               workloads.cell_centres(box)]
    r = torch.sqrt(X * X + Y * Y + Z * Z)
    atm = opts.atmosphere
    atmos = r < 1.5 * M
    rho = torch.where(atmos, torch.full_like(rho0, atm.rho_floor), rho0)
    p = torch.where(atmos, torch.full_like(rho0, atm.pressure_floor),
                    0.1 * rho0 ** (4.0 / 3.0))
    vr = -0.9 * torch.sqrt(2 * M / r)
    v = [torch.where(atmos, torch.zeros_like(X), vr * c / r)
         for c in (X, Y, Z)]
    g = ks_metric_torch(spin, X, Y, Z)
    rho = torch.clamp(rho, min=atm.rho_floor)
    p = torch.clamp(p, min=atm.pressure_floor)
    v2 = sum(g[i][j] * v[i] * v[j] for i in range(3) for j in range(3))
    cap = atm.max_velocity_squared if init_max_v2 is None else init_max_v2
    over = v2 >= cap
    scale = torch.where(over, torch.sqrt(cap / torch.clamp(v2, min=1e-300)),
                        torch.ones_like(v2))
    v = [c * scale for c in v]
    eps = p / ((gamma - 1.0) * rho)
    vL = [sum(g[i][j] * v[j] for j in range(3)) for i in range(3)]
    v2 = sum(vL[i] * v[i] for i in range(3))
    W = 1.0 / torch.sqrt(1.0 - v2)
    h = 1.0 + eps + p / rho
    rhW2 = rho * h * W * W
    D = rho * W
    S = [rhW2 * c for c in vL]
    tau = rhW2 - p - D
    return torch.stack([D, S[0], S[1], S[2], tau])


def seed_rho0(box, seed):
    """rho0 = 1e-3 (r / 2M)^-1.5 (1 + 0.05 xi), xi ~ U(-1, 1) (SplitMix64)."""
    X, Y, Z = workloads.cell_centres(box)
    r = np.sqrt(X * X + Y * Y + Z * Z)
    xi = workloads.uniform_pm1(seed, X.size).reshape(X.shape)
    return 1e-3 * (r / (2 * M)) ** -1.5 * (1.0 + 0.05 * xi)


def full_state(box, U0):
    """11-component FAB with the conserved part U0 (primitives unused)."""
    st = torch.zeros((11,) + box.storage_shape, dtype=torch.float64,
                     device=U0.device)
    st[0:5] = U0
    return st.contiguous()


class RK4Objective(torch.autograd.Function):
    """J(U0, Gamma, a) of grhd_rk4_adjoint_box with its adjoint."""

    @staticmethod
    def forward(ctx, U0, gamma, spin, box, opts_fn, dt, nsteps, r0,
                first_step, checkpoint_every, context):
        op = FixedBGLevelDataFiniteVolumeOperator(
            KerrSchildParams(M, float(spin.detach())), box,
            opts_fn(float(gamma.detach())), context=context)
        J, g, dg, da = op.rk4_adjoint(full_state(box, U0.detach()), dt,
                                      nsteps, r0, first_step,
                                      checkpoint_every)
        grad = torch.zeros_like(U0)
        sl = (slice(None),) + box.interior_slices()
        grad[sl] = g
        ctx.save_for_backward(grad)
        ctx.dg, ctx.da = dg, da
        return U0.new_tensor(J)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return (gJ * grad, gJ * ctx.dg, gJ * ctx.da, None, None, None, None,
                None, None, None, None)


def objective(box, rho0, gamma, spin, dt, nsteps, r0, first_step,
              checkpoint_every, context, max_v2=0.8, **opt_kw):
    """J through the initial-data map; opt_kw go to `options`."""
    opts_fn = lambda gm: options(gm, max_v2, **opt_kw)  # noqa: E731
    U0 = michel_inflow_U0(box, rho0, gamma, spin,
                          opts_fn(float(gamma.detach())))
    return RK4Objective.apply(U0, gamma, spin, box, opts_fn, dt, nsteps, r0,
                              first_step, checkpoint_every, context)


def offset_box(n, dx, lo_x, ghosts=3):
    """n^3 cells with x from lo_x, y and z centred on 0."""
    w = n * dx
    return BoxSpec((0, 0, 0), (n - 1,) * 3, ghosts, dx,
                   (-lo_x, 0.5 * w, 0.5 * w))


def smooth_field(box, rng):
    """Seeded smooth field over the storage box: sum of 6 products of sin
    modes 1..2 along each axis (random phases) with U(-1, 1) weights."""
    X, Y, Z = workloads.cell_centres(box)
    L = box.dx * np.array(box.storage_shape[::-1])
    f = np.zeros(X.shape)
    for _ in range(6):
        w = rng.uniform(-1, 1)
        m = rng.integers(1, 3, size=3)
        ph = rng.uniform(0, 2 * math.pi, size=3)
        f += w * np.sin(math.pi * m[0] * X / L[0] + ph[0]) * \
            np.sin(math.pi * m[1] * Y / L[1] + ph[1]) * \
            np.sin(math.pi * m[2] * Z / L[2] + ph[2])
    return torch.from_numpy(f).cuda()


def objective_only(box, rho0, gamma, spin, dt, nsteps, r0, first_step,
                   context, max_v2=0.8, **opt_kw):
    """J alone (forward sweep only), for finite differences."""
    opts = options(gamma, max_v2, **opt_kw)
    with torch.no_grad():
        U0 = michel_inflow_U0(box, rho0, torch.tensor(gamma,
                                                      dtype=torch.float64),
                              torch.tensor(spin, dtype=torch.float64), opts)
    op = FixedBGLevelDataFiniteVolumeOperator(KerrSchildParams(M, spin), box,
                                              opts, context=context)
    return op.rk4_objective(full_state(box, U0), dt, nsteps, r0, first_step)
