"""Michel-accretion convergence study (the reference has none; built here).

The stationary Michel flow (Schwarzschild, Kerr-Schild coordinates) is an
exact solution of the continuum equations, so the discrete right-hand side
evaluated on it is the scheme's truncation error.  Every storage cell holds
the exact state and the ProblemDomain encloses the ghosts, so there is no
boundary error.  The test fluid is scaled to sonic density 1 so that the
con2prim tolerance (1e-12, absolute on the pressure residual) sits far
below the truncation error; at the reference's 1e-6 the recovered pressure
carries ~1e-5 relative noise, which grows as 1/dx in the rhs and stalls the
momentum and tau orders by n = 128.  On x in [12, 20], y, z in [-4, 4]
(outside the sonic radius 8) the volume-weighted L1 norm of the rhs falls
at second order with MUSCL + MC (oracle, n = 8 -> 64: D 1.75, 1.83, 1.90;
S_1 1.39, 1.62, 1.80; tau 1.38, 1.63, 1.80) and at first order with
reconstruction off.  The GPU rhs equals the oracle's bit for bit, so the
CPU and GPU orders are the same numbers."""
import dataclasses
import math

import numpy as np
import pytest

import michel
import oracle
from grdzhadzha_amd.options import KerrSchildParams

CFG = michel.MichelConfig(sonic_density=1.0)
BG = KerrSchildParams(CFG.mass, 0.0)


def errors(ns, opts, rhs_fn):
    out = []
    for n in ns:
        box = michel.convergence_box(n)
        st = michel.michel_box_state(CFG, box, opts)
        out.append(michel.rhs_l1(box, rhs_fn(box, st)))
    return out


def orders(errs):
    return [[math.log2(a / b) for a, b in zip(p, q)]
            for p, q in zip(errs, errs[1:])]


def oracle_rhs(opts):
    return lambda box, st: oracle.Oracle(BG, box, opts).compute_rhs(st.copy())


def test_exact_state_is_stationary_to_truncation_error():
    # the state solves the Bernoulli relation the reference's profile uses
    c = michel.make_constants(CFG)
    r = np.array([12.5, 15.0, 19.7])
    rho = michel.solve_density_many(c, CFG, r)
    for ri, rh in zip(r, rho):
        assert abs(rh - michel.solve_density(c, CFG, ri)) <= 1e-12 * rh


def test_muscl_second_order_oracle():
    opts = michel.convergence_options(CFG, theta=2.0)
    o = orders(errors((8, 16, 32), opts, oracle_rhs(opts)))
    assert 1.75 < o[1][0] < 2.2, o         # D
    assert all(p > 1.55 for p in o[1]), o  # every component
    assert all(b > a for a, b in zip(o[0], o[1])), o  # rising towards 2


def test_first_order_without_reconstruction_oracle():
    opts = dataclasses.replace(michel.convergence_options(CFG),
                               use_reconstruction=False)
    o = orders(errors((8, 16, 32), opts, oracle_rhs(opts)))
    assert all(0.85 < p < 1.1 for p in o[1]), o


@pytest.mark.gpu
def test_gpu_convergence_matches_oracle_and_order(hip_context):
    import torch

    from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator
    opts = michel.convergence_options(CFG, theta=2.0)

    def gpu_rhs(box, st):
        op = FixedBGLevelDataFiniteVolumeOperator(BG, box, opts,
                                                  context=hip_context)
        s = torch.from_numpy(st.copy()).cuda()
        r = torch.zeros((11,) + box.interior_shape, dtype=torch.float64,
                        device="cuda")
        op.compute_rhs(r, s)
        return r.cpu().numpy()

    for n in (16, 32):
        box = michel.convergence_box(n)
        st = michel.michel_box_state(CFG, box, opts)
        assert np.array_equal(gpu_rhs(box, st)[:5],
                              oracle_rhs(opts)(box, st)[:5])
    o = orders(errors((32, 64, 128), opts, gpu_rhs))
    assert all(1.75 < p < 2.2 for p in o[1]), o
