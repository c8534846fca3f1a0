"""The device accretion rate (inside grhd_rk4_adjoint_box) against the
exact Michel flux and against workloads.mass_accretion_rate.

J of one RK4 step with first_step = nsteps = 1 is dt * mdot(U_1); at
dt = 1e-9 the step moves the state by ~1e-9 of its truncation error, so
J / dt is mdot of the initial state to ~1e-12.
- On the exact Michel state (Schwarzschild, tests/michel.py) every shell
  term equals the Michel constant over r^2, so J / dt must equal
  -F dx^2 sum_shell 1/r^2 (F = r^2 rho u^r) to rounding and the con2prim
  round trip.
- On a Kerr a = 0.5 inflow the device value equals the numpy diagnostic
  of the oracle's state after the same step (same con2prim, independent
  metric code).
"""
import numpy as np
import pytest
import torch

import adjoint_harness as ah
import michel
import oracle
from grdzhadzha_amd import workloads
from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator
from grdzhadzha_amd.options import KerrSchildParams

pytestmark = pytest.mark.gpu

DT = 1e-9


@pytest.mark.parametrize("r0", [13.6, 14.3])
def test_device_mdot_equals_michel_flux(hip_context, r0):
    cfg = michel.MichelConfig(sonic_density=1.0)
    box, st = michel.accretion_box_state(cfg, 64)
    opts = michel.convergence_options(cfg)
    op = FixedBGLevelDataFiniteVolumeOperator(
        KerrSchildParams(cfg.mass, 0.0), box, opts, context=hip_context)
    J = op.rk4_objective(torch.from_numpy(st).cuda(), DT, 1, r0, 1)
    F = michel.make_constants(cfg)["mass_flux"]
    r = michel.shell_cells(box, r0)
    want = -F * box.dx ** 2 * float(np.sum(1.0 / r ** 2))
    got = J / DT
    print("r0", r0, "device mdot", got, "closed form", want, flush=True)
    assert abs(got - want) <= 1e-9 * abs(want)


def test_device_mdot_equals_numpy_diagnostic_kerr(hip_context):
    gamma, spin = 4.0 / 3.0, 0.5
    box = ah.offset_box(16, 0.5, 4.0)
    rho0 = torch.from_numpy(ah.seed_rho0(box, 5)).cuda()
    opts = ah.options(gamma, tolerance=1e-12)
    U0 = ah.michel_inflow_U0(box, rho0, gamma, spin, opts)
    st = ah.full_state(box, U0.contiguous())
    bg = KerrSchildParams(1.0, spin)
    op = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                              context=hip_context)
    r0 = 7.0
    J = op.rk4_objective(st, DT, 1, r0, 1)
    ref = st.cpu().numpy().copy()
    oracle.Oracle(bg, box, opts).advance_rk4(ref, DT, 1)
    want = workloads.mass_accretion_rate(box, ref, 1.0, spin, r0)
    print("device", J / DT, "numpy", want, flush=True)
    assert want > 0.0
    assert abs(J / DT - want) <= 1e-11 * abs(want)
