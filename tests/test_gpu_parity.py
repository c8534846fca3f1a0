"""HIP path vs the CPU oracle, called through the C-ABI (ctypes).

Bar: bit for bit.  The rhs, the state (interior after a step; every storage
cell after an rhs, whose ghost fill mutates it), recovery masks, reset
counts and dt equal the oracle's exactly: both evaluate the reference's
formulas in the same order without contraction.  (The FP64 path's bar is
1e-12 relative per conserved variable; bit-equality is stronger.)  On a
mismatch the message reports the ulp distance and the max-normalised
error.
"""
import dataclasses

import numpy as np
import pytest
import torch

import oracle
from grdzhadzha_amd import workloads
from grdzhadzha_amd.operator import (
    FixedBGLevelDataFiniteVolumeOperator,
    TimeDependentFixedBGLevelDataFiniteVolumeOperator)
from grdzhadzha_amd.options import (BoxSpec, CircularBinaryPNParams,
                                    KerrSchildParams, OperatorOptions,
                                    RecoveryOptions)
from helpers import assert_bitwise, conserved_errors

pytestmark = pytest.mark.gpu



def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def small_torus():
    return workloads.torus_box(n=32, R0=5.0, sigma=1.5, nsteps=2)


def reference_test_state(box, background, gamma, ncomp=11):
    """State of Tests/GRHDFixedBGTest/GRHDFixedBGTest.cpp:84-158: linear
    primitive field, conserved from prim2con on the cell metric."""
    st = np.zeros((ncomp,) + box.storage_shape)
    geo = oracle.Oracle(background, box).cell_geometry()
    nz, ny, nx = box.storage_shape
    glo = box.glo
    for kk in range(nz):
        for jj in range(ny):
            for ii in range(nx):
                i, j, k = glo[0] + ii, glo[1] + jj, glo[2] + kk
                rho = 0.7 + 1.0e-3 * (i + 2 * j + 3 * k)
                eps = 0.12 + 1.0e-4 * (2 * i + j + k)
                p = (gamma - 1.0) * rho * eps
                v = (1.0e-3 * (1.0 + i), -7.5e-4 * (1.0 + j),
                     5.0e-4 * (1.0 + k))
                g6 = geo[0:6, kk, jj, ii]
                g = np.array([[g6[0], g6[1], g6[2]], [g6[1], g6[3], g6[4]],
                              [g6[2], g6[4], g6[5]]])
                cons = oracle.prim2con([rho, eps, p, *v], gamma, g)
                st[0:5, kk, jj, ii] = cons
                st[5:11, kk, jj, ii] = [rho, eps, p, *v]
    return st


# ---------------------------------------------------------------- geometry


@pytest.mark.parametrize("bg", [KerrSchildParams(1.0, 0.0),
                                KerrSchildParams(1.0, 0.9),
                                CircularBinaryPNParams(0.55, 0.45, 8.0, 0.0,
                                                       0.25, 0.4, 0.8)])
def test_cell_geometry_bitwise(hip_context, bg):
    box = BoxSpec((0, 0, 0), (15, 13, 11), 3, 0.37, (3.1, 2.7, 1.9))
    gpu = FixedBGLevelDataFiniteVolumeOperator(bg, box,
                                               context=hip_context)
    g = gpu.cell_geometry(time=0.3).cpu().numpy()
    ref = oracle.Oracle(bg, box).cell_geometry(time=0.3)
    assert np.array_equal(g, ref), np.max(np.abs(g - ref))


# ---------------------------------------------------------------- RHS


def _rhs_case(w, ctx, time=0.0, time_dependent=False):
    cls = (TimeDependentFixedBGLevelDataFiniteVolumeOperator
           if time_dependent else FixedBGLevelDataFiniteVolumeOperator)
    op = cls(w.background, w.box, w.options, context=ctx)
    O = oracle.Oracle(w.background, w.box, w.options,
                      time_dependent=time_dependent)
    st_ref = np.ascontiguousarray(w.state).copy()
    rhs_ref, mask_ref = O.compute_rhs(st_ref, time=time, with_mask=True)
    st = to_dev(w.state)
    rhs = torch.zeros((st.shape[0],) + w.box.interior_shape,
                      dtype=torch.float64, device="cuda")
    mask = torch.zeros(w.box.interior_shape, dtype=torch.uint8, device="cuda")
    op.compute_rhs(rhs, st, time=time, source_mask=mask)
    torch.cuda.synchronize()
    got = rhs.cpu().numpy()
    assert_bitwise(got, rhs_ref, "rhs")
    assert np.array_equal(mask.cpu().numpy(), mask_ref)
    # the ghost fill mutates the state exactly like the reference
    assert_bitwise(st.cpu().numpy(), st_ref, "state after the ghost fill")
    assert np.all(got[5:] == 0.0)
    return conserved_errors(got, rhs_ref)


def test_rhs_michel_32(hip_context):
    _rhs_case(workloads.michel_box_32(), hip_context)


def test_rhs_torus_kerr(hip_context):
    _rhs_case(small_torus(), hip_context)


def test_rhs_torus_96(hip_context):
    _rhs_case(workloads.torus_box(n=96, nsteps=1), hip_context)


FACE_LAUNCHES = [(walk, fused) for walk in (0, 1, 2, 4, 8)
                 for fused in (True, False)]


def test_rhs_torus_every_face_launch(hip_context):
    """The ten face launches pinned in turn on a 64^3 torus box: the y/z
    faces formed in 8 x 8 tiles (walk 0) or by walkers of 1, 2, 4 or 8
    faces, each carrying a cell's metric and slopes to its next face; the
    three directions in one fused launch or three on streams.  Every one
    gives the oracle's rhs, mask and ghost-filled state bit for bit."""
    w = workloads.torus_box(n=64, nsteps=1)
    O = oracle.Oracle(w.background, w.box, w.options)
    st_ref = np.ascontiguousarray(w.state).copy()
    rhs_ref, mask_ref = O.compute_rhs(st_ref, time=0.0, with_mask=True)
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=hip_context)
    try:
        for walk, fused in FACE_LAUNCHES:
            hip_context.set_face_launch(walk, fused)
            st = to_dev(w.state)
            rhs = torch.zeros((st.shape[0],) + w.box.interior_shape,
                              dtype=torch.float64, device="cuda")
            mask = torch.zeros(w.box.interior_shape, dtype=torch.uint8,
                               device="cuda")
            op.compute_rhs(rhs, st, source_mask=mask)
            torch.cuda.synchronize()
            assert hip_context.face_launch() == {
                "walk": walk, "fused": fused, "tuned": False, "ms": 0.0}
            what = "walk %d, fused %s" % (walk, fused)
            assert_bitwise(rhs.cpu().numpy(), rhs_ref, "rhs, " + what)
            assert np.array_equal(mask.cpu().numpy(), mask_ref), what
            assert_bitwise(st.cpu().numpy(), st_ref,
                           "state after the ghost fill, " + what)
    finally:
        hip_context.set_face_launch(None)


def random_box_case(seed, binary=False):
    """A seeded random box away from the hole: sides 2..19 cells, 3
    ghosts, dx 0.2..0.5, a Kerr-Schild spin in [0, 0.95) (or, with
    `binary`, a circular binary of random masses, separation and phase
    at a random time, with the volume-time sources on), a random
    adiabatic index and limiter, and a perturbed linear primitive field
    with 5% of its cells at the density and pressure floors.  Returns
    (box, background, options, state, time)."""
    rng = np.random.default_rng(1000 + seed)
    n = rng.integers(2, 20, 3)
    lo = rng.integers(25, 40, 3)
    box = BoxSpec(tuple(lo), tuple(lo + n - 1), 3,
                  float(rng.uniform(0.2, 0.5)),
                  tuple(float(v) for v in rng.uniform(-2.0, 2.0, 3)))
    spin = float(rng.uniform(0.0, 0.95))
    gamma = float(rng.choice([4.0 / 3.0, 1.4, 5.0 / 3.0]))
    theta = float(rng.choice([1.0, 1.5, 2.0]))
    time = 0.0
    if binary:
        m1 = float(rng.uniform(0.3, 0.7))
        bg = CircularBinaryPNParams(
            mass_1=m1, mass_2=1.0 - m1,
            separation=float(rng.uniform(4.0, 10.0)),
            phase=float(rng.uniform(0.0, 6.0)),
            softening_radius=float(rng.uniform(0.3, 1.0)))
        time = float(rng.uniform(0.0, 5.0))
        opts = OperatorOptions(adiabatic_index=gamma, limiter_theta=theta,
                               use_metric_volume_time_sources=True,
                               metric_time_derivative_step=1.0e-3)
    else:
        bg = KerrSchildParams(1.0, spin)
        opts = OperatorOptions(adiabatic_index=gamma, limiter_theta=theta)
    st = reference_test_state(box, bg, gamma)
    geo = oracle.Oracle(bg, box).cell_geometry(time=time)
    shape = box.storage_shape
    xi = rng.uniform(-1.0, 1.0, (5,) + shape)
    prim = st[5:11].copy()
    prim[0] *= 1.0 + 0.3 * xi[0]
    prim[1] *= 1.0 + 0.3 * xi[1]
    for m in range(3):
        prim[3 + m] += 0.05 * xi[2 + m]
    prim[2] = (gamma - 1.0) * prim[0] * prim[1]
    # a few cells at the floors: density around rho_floor (some below it),
    # pressure between the floor and twice it
    low = rng.uniform(0.0, 1.0, shape) < 0.05
    nlow = int(low.sum())
    prim[0][low] = 1.0e-12 * rng.uniform(0.5, 2.0, nlow)
    prim[2][low] = 1.0e-12 * rng.uniform(1.0, 2.0, nlow)
    prim[1][low] = prim[2][low] / ((gamma - 1.0) * prim[0][low])
    for idx in np.ndindex(shape):
        g6 = geo[0:6][(slice(None),) + idx]
        g = np.array([[g6[0], g6[1], g6[2]], [g6[1], g6[3], g6[4]],
                      [g6[2], g6[4], g6[5]]])
        pv = prim[(slice(None),) + idx]
        st[(slice(0, 5),) + idx] = oracle.prim2con(pv, gamma, g)
        st[(slice(5, 11),) + idx] = pv
    return box, bg, opts, st, time


@pytest.mark.parametrize("seed", range(8))
def test_rhs_random_boxes(hip_context, seed):
    """Seeded random boxes (sizes, positions, dx, Gamma, limiter,
    perturbed states with cells at and below the floors), on Kerr-Schild
    (seeds 0-3) or on a random circular binary at a random time with the
    volume-time sources (seeds 4-7, the time-dependent operator), each
    with one of the ten face launches pinned: the rhs, mask and
    ghost-filled state equal the oracle's bit for bit."""
    box, bg, opts, st0, t = random_box_case(seed, binary=seed >= 4)
    walk, fused = FACE_LAUNCHES[seed % len(FACE_LAUNCHES)]
    td = seed >= 4
    O = oracle.Oracle(bg, box, opts, time_dependent=td)
    st_ref = st0.copy()
    rhs_ref, mask_ref = O.compute_rhs(st_ref, time=t, with_mask=True)
    cls = (TimeDependentFixedBGLevelDataFiniteVolumeOperator if td
           else FixedBGLevelDataFiniteVolumeOperator)
    op = cls(bg, box, opts, context=hip_context)
    try:
        hip_context.set_face_launch(walk, fused)
        st = to_dev(st0)
        rhs = torch.zeros((st.shape[0],) + box.interior_shape,
                          dtype=torch.float64, device="cuda")
        mask = torch.zeros(box.interior_shape, dtype=torch.uint8,
                           device="cuda")
        op.compute_rhs(rhs, st, time=t, source_mask=mask)
        torch.cuda.synchronize()
    finally:
        hip_context.set_face_launch(None)
    what = "seed %d, box %s..%s, walk %d fused %s" % (seed, box.lo, box.hi,
                                                     walk, fused)
    assert_bitwise(rhs.cpu().numpy(), rhs_ref, "rhs, " + what)
    assert np.array_equal(mask.cpu().numpy(), mask_ref), what
    assert_bitwise(st.cpu().numpy(), st_ref, "ghost-filled state, " + what)


def test_rhs_no_reconstruction_no_sources(hip_context):
    w = small_torus()
    w.options.use_reconstruction = False
    w.options.use_static_metric_sources = False
    _rhs_case(w, hip_context)


def test_rhs_reference_helper_box(hip_context):
    """GRHDFixedBGTest.cpp:190-237: interior (20..22, 15..17, 10..12), 2
    ghosts, ProblemDomain = ghosted box (no physical ghosts), KS a = 0.3,
    dx = 0.25, theta = 1.5, Gamma = 5/3, max_iterations = 80."""
    box = BoxSpec((20, 15, 10), (22, 17, 12), 2, 0.25, (0.0, 0.0, 0.0),
                  domain_lo=(18, 13, 8), domain_hi=(24, 19, 14))
    bg = KerrSchildParams(1.0, 0.3)
    opts = OperatorOptions(adiabatic_index=5.0 / 3.0,
                           recovery=RecoveryOptions(max_iterations=80))
    st = reference_test_state(box, bg, 5.0 / 3.0)
    w = workloads.Workload("ref", box, bg, opts, st, 0.0, 0)
    _rhs_case(w, hip_context)


def test_rhs_reference_leveldata_box_binary(hip_context):
    """GRHDFixedBGTest.cpp:463-551: one box = domain (2 ghosts), binary PN
    time factory (m1 0.55, m2 0.45, sep 8, phase 0.25, time 0.4, soft 0.8),
    volume-time sources with step 1e-4, evaluated at t = 0.35."""
    box = BoxSpec((20, 15, 10), (22, 17, 12), 2, 0.25, (0.0, 0.0, 0.0))
    bg = CircularBinaryPNParams(0.55, 0.45, 8.0, 0.0, 0.25, 0.4, 0.8)
    opts = OperatorOptions(adiabatic_index=5.0 / 3.0,
                           recovery=RecoveryOptions(max_iterations=80),
                           use_metric_volume_time_sources=True,
                           metric_time_derivative_step=1.0e-4)
    st = reference_test_state(box, bg, 5.0 / 3.0)
    w = workloads.Workload("ref-pn", box, bg, opts, st, 0.0, 0)
    _rhs_case(w, hip_context, time=0.35, time_dependent=True)


def test_anisotropic_inverse_dx_rhs_cfl_and_rk4(hip_context):
    """inverse_dx_U with a different spacing per direction (the reference's
    overloads, FixedBGFArrayBoxFiniteVolume.hpp:468-484, :535-589): the
    rhs, the CFL rate and an RK4 step equal the oracle's bit for bit."""
    w = small_torus()
    inv = 1.0 / w.box.dx
    w.box = dataclasses.replace(w.box,
                                inverse_dx=(inv, 1.25 * inv, 0.8 * inv))
    _rhs_case(w, hip_context)
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=hip_context)
    O = oracle.Oracle(w.background, w.box, w.options)
    st = np.ascontiguousarray(w.state)
    assert op.compute_max_inverse_dt(to_dev(st)) == \
        O.compute_max_inverse_dt(st.copy())
    _step_case(w, hip_context, "rk4", 1)


# ---------------------------------------------------------------- recovery


def test_recover_masks_bitexact(hip_context):
    w = workloads.michel_box_32()
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=hip_context)
    O = oracle.Oracle(w.background, w.box, w.options)
    st_ref = np.ascontiguousarray(w.state).copy()
    # perturb so that every branch fires: failures (D <= 0), floors, resets
    sl = w.box.interior_slices()
    st_ref[0][sl][3, 4, 5] = -1.0
    st_ref[0][sl][10, 10, 10] = 1e-14
    st_ref[4][sl][7, 8, 9] = -1.0
    diag_ref, mask_ref = O.recover_primitives(st_ref2 := st_ref.copy())
    st = to_dev(st_ref)
    mask = torch.zeros(w.box.interior_shape, dtype=torch.uint8, device="cuda")
    diag = op.recover_primitives(st, mask=mask)
    assert diag == diag_ref
    assert diag["num_failed_recoveries"] >= 1
    assert np.array_equal(mask.cpu().numpy(), mask_ref)
    assert np.array_equal(st.cpu().numpy(), st_ref2)


def test_max_inverse_dt(hip_context):
    for w in (workloads.michel_box_32(), small_torus()):
        op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box,
                                                  w.options,
                                                  context=hip_context)
        O = oracle.Oracle(w.background, w.box, w.options)
        st = np.ascontiguousarray(w.state)
        ref = O.compute_max_inverse_dt(st.copy())
        got = op.compute_max_inverse_dt(to_dev(st))
        assert got == ref
        assert op.compute_stable_dt(to_dev(st), 0.2) == \
            O.compute_stable_dt(st.copy(), 0.2)
    with pytest.raises(ValueError):
        op.compute_stable_dt(to_dev(st), 0.0)


@pytest.mark.parametrize("seed", range(8))
def test_max_inverse_dt_and_ssprk2_random_boxes(hip_context, seed):
    """The random boxes of test_rhs_random_boxes: max 1/dt equals the
    oracle's exactly, and one SSPRK2 step at the CFL dt (0.3) gives the
    oracle's state, counts and masks bit for bit."""
    import types
    box, bg, opts, st0, t = random_box_case(seed, binary=seed >= 4)
    td = seed >= 4
    cls = (TimeDependentFixedBGLevelDataFiniteVolumeOperator if td
           else FixedBGLevelDataFiniteVolumeOperator)
    op = cls(bg, box, opts, context=hip_context)
    O = oracle.Oracle(bg, box, opts, time_dependent=td)
    ref = O.compute_max_inverse_dt(st0.copy(), time=t)
    assert op.compute_max_inverse_dt(to_dev(st0), time=t) == ref
    dt = O.compute_stable_dt(st0.copy(), 0.3, time=t)
    w = types.SimpleNamespace(background=bg, box=box, options=opts,
                              state=st0, dt=dt)
    if td:
        st_ref = st0.copy()
        st = to_dev(st_ref)
        mask = torch.zeros(box.interior_shape, dtype=torch.uint8,
                           device="cuda")
        diag_ref, mask_ref = O.advance_ssprk2(st_ref, dt, time=t)
        diag = op.advance_ssprk2(st, dt, time=t, mask=mask)
        sl = (slice(None),) + box.interior_slices()
        assert_bitwise(st.cpu().numpy()[sl], st_ref[sl], "ssprk2 state")
        assert diag == diag_ref
        assert np.array_equal(mask.cpu().numpy(), mask_ref)
    else:
        _step_case(w, hip_context, "ssprk2")


# ---------------------------------------------------------------- steps


def _step_case(w, ctx, method, nsteps=1, time=0.0):
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=ctx)
    O = oracle.Oracle(w.background, w.box, w.options)
    st_ref = np.ascontiguousarray(w.state).copy()
    st = to_dev(st_ref)
    mask = torch.zeros(w.box.interior_shape, dtype=torch.uint8, device="cuda")
    if method == "rk4":
        diag_ref, mask_ref = O.advance_rk4(st_ref, w.dt, nsteps, time=time)
        diag = op.advance_rk4(st, w.dt, nsteps, time=time, mask=mask)
    else:
        diag_ref, mask_ref = O.advance_ssprk2(st_ref, w.dt, time=time)
        diag = op.advance_ssprk2(st, w.dt, time=time, mask=mask)
    got = st.cpu().numpy()
    sl = (slice(None),) + w.box.interior_slices()
    print(method, "max-normalised error", conserved_errors(got, st_ref))
    assert_bitwise(got[sl], st_ref[sl], method + " interior state")
    assert diag == diag_ref
    assert np.array_equal(mask.cpu().numpy(), mask_ref)


@pytest.mark.parametrize("inside", [1, 2])
def test_rhs_box_short_of_its_domain_face(hip_context, inside):
    """A box whose low x face lies 1 or 2 cells inside its ProblemDomain:
    its ghost layers inside the domain are internal ghosts (the caller's
    values), and the ones past the domain face are outflow ghosts filled
    from the box's own edge cells, as the reference does
    (FixedBGFArrayBoxFiniteVolume.hpp:130-146), not from the domain's
    edge.  The rhs, mask and ghost-filled state equal the oracle's bit for
    bit."""
    import types
    box0, bg, opts, st0, _ = random_box_case(3)
    dlo = (box0.lo[0] - inside, box0.lo[1], box0.lo[2])
    box = dataclasses.replace(box0, domain_lo=dlo)
    w = types.SimpleNamespace(background=bg, box=box, options=opts,
                              state=st0)
    _rhs_case(w, hip_context)


@pytest.mark.parametrize("seed", range(4))
def test_rk4_random_boxes(hip_context, seed):
    """Two RK4 steps (dt = 0.2 dx) of the seeded random boxes of
    test_rhs_random_boxes under the size rule's face launch: state,
    recovery counts and masks equal the oracle's bit for bit."""
    import types
    box, bg, opts, st0, _ = random_box_case(seed)
    w = types.SimpleNamespace(background=bg, box=box, options=opts,
                              state=st0, dt=0.2 * box.dx)
    _step_case(w, hip_context, "rk4", 2)


def test_rk4_michel_one_step(hip_context):
    _step_case(workloads.michel_box_32(), hip_context, "rk4", 1)


def test_rk4_torus_two_steps(hip_context):
    _step_case(small_torus(), hip_context, "rk4", 2)


def test_rk4_torus_128_full_size_one_step(hip_context):
    """The benchmark workload itself at full size: one RK4 step of the
    128^3 Kerr a = 0.9 torus (2.1M interior cells, minmod, dt = 0.25 dx),
    the size at which the face kernels form the y/z faces in tiles and run
    fused.  State, masks and recovery counts equal the oracle's bit for
    bit.  The oracle needs a few seconds on
    the host's cores."""
    _step_case(workloads.torus_box(n=128, nsteps=1), hip_context, "rk4", 1)


def test_ssprk2_michel(hip_context):
    _step_case(workloads.michel_box_32(), hip_context, "ssprk2")


# ------------------------------------------------------- horizon excision


def excised(w):
    return workloads.Workload(
        w.name + "_excised", w.box, w.background,
        dataclasses.replace(w.options, excise_horizon=True), w.state, w.dt,
        w.nsteps)


def test_rhs_michel_32_excised(hip_context):
    _rhs_case(excised(workloads.michel_box_32()), hip_context)


def test_rhs_torus_kerr_excised(hip_context):
    _rhs_case(excised(small_torus()), hip_context)


def test_rk4_michel_excised_three_steps(hip_context):
    _step_case(excised(workloads.michel_box_32()), hip_context, "rk4", 3)


def test_ssprk2_torus_excised(hip_context):
    _step_case(excised(small_torus()), hip_context, "ssprk2")


def test_excision_needs_a_kerr_schild_background(hip_context):
    w = workloads.michel_box_32(n=16, half_width=4.0)
    op = FixedBGLevelDataFiniteVolumeOperator(
        CircularBinaryPNParams(0.5, 0.5, 10.0, 0.0, 0.0, 0.1, 0.4), w.box,
        dataclasses.replace(w.options, excise_horizon=True),
        context=hip_context)
    st = to_dev(w.state)
    rhs = torch.zeros((st.shape[0],) + w.box.interior_shape,
                      dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        op.compute_rhs(rhs, st)


# ---------------------------------------------------------------- errors


def test_rhs_recovery_error_reports_cell(hip_context):
    w = workloads.michel_box_32()
    st = np.ascontiguousarray(w.state).copy()
    g = w.box.ghosts
    st[0, g + 5, g + 6, g + 7] = -1.0  # D <= 0 at interior (7, 6, 5)
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=hip_context)
    rhs = torch.zeros((11,) + w.box.interior_shape, dtype=torch.float64,
                      device="cuda")
    with pytest.raises(RuntimeError, match=r"failed at \(7,6,5\)"):
        op.compute_rhs(rhs, to_dev(st))
    with pytest.raises(oracle.OracleError) as e:
        oracle.Oracle(w.background, w.box, w.options).compute_rhs(st.copy())
    assert e.value.iv == (7, 6, 5)


def test_ghost_fill_error_names_source(hip_context):
    w = workloads.michel_box_32()
    st = np.ascontiguousarray(w.state).copy()
    g = w.box.ghosts
    st[0, g + 0, g + 4, g + 9] = -1.0  # boundary cell (9, 4, 0) is a source
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, w.options,
                                              context=hip_context)
    rhs = torch.zeros((11,) + w.box.interior_shape, dtype=torch.float64,
                      device="cuda")
    with pytest.raises(RuntimeError, match=r"outflow source \(9,4,0\)"):
        op.compute_rhs(rhs, to_dev(st))
    with pytest.raises(oracle.OracleError) as e:
        oracle.Oracle(w.background, w.box, w.options).compute_rhs(st.copy())
    assert e.value.code == 1 and e.value.iv == (9, 4, 0)


def _horizon_box_with_strict_floor():
    """8^3 torus-state box whose x-low face touches the Kerr horizon; the
    recovery's pressure floor (2e-12) sits above the atmosphere pressure
    (1e-12).  The first RHS recovers everywhere; cells reset to atmosphere
    by the first update cannot be recovered again, so the next evaluation's
    ghost fill fails at an outflow source before its interior recovery."""
    import dataclasses
    box = BoxSpec((0, 0, 0), (7, 7, 7), 3, 0.5, (0.0, 2.0, 2.0))
    opts, st = workloads.torus_state(box)
    rec = dataclasses.replace(opts.recovery, pressure_floor=2.0e-12)
    return box, dataclasses.replace(opts, recovery=rec), st


@pytest.mark.parametrize("scheme", ["rk4", "ssprk2"])
def test_later_evaluation_ghost_fill_error(hip_context, scheme):
    box, opts, st = _horizon_box_with_strict_floor()
    bg = KerrSchildParams(1.0, 0.9)
    O = oracle.Oracle(bg, box, opts)
    O.compute_rhs(st.copy())  # evaluation 0 recovers everywhere
    with pytest.raises(oracle.OracleError) as e:
        if scheme == "rk4":
            O.advance_rk4(st.copy(), 0.125, 1)
        else:
            O.advance_ssprk2(st.copy(), 0.125)
    assert e.value.code == 1
    i, j, k = e.value.iv
    op = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                              context=hip_context)
    with pytest.raises(RuntimeError,
                       match=r"outflow source \(%d,%d,%d\)" % (i, j, k)):
        if scheme == "rk4":
            op.advance_rk4(to_dev(st.copy()), 0.125, 1)
        else:
            op.advance_ssprk2(to_dev(st.copy()), 0.125)


# ------------------------------------------------------------ tiny boxes


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 2, 3), (3, 1, 2),
                                   (2, 5, 1), (9, 1, 1), (13, 11, 9),
                                   (8, 15, 22)])
def test_tiny_and_ragged_boxes_rhs_and_rk4(hip_context, shape):
    """Interiors thinner than the ghost width along some axis, or covered
    by partial y/z face tiles (8 cells along i by 8 along y or z, 7 faces
    per column): the x-face wave packing, the tiles and the exterior-cell
    enumeration against the oracle (rhs and one RK4 step)."""
    bg = KerrSchildParams(1.0, 0.5)
    box = BoxSpec((2, 3, 4), tuple(l + n - 1 for l, n in
                                   zip((2, 3, 4), shape)), 3, 0.37,
                  (-4.1, 1.3, -0.7))
    opts = OperatorOptions(adiabatic_index=4.0 / 3.0)
    st = reference_test_state(box, bg, 4.0 / 3.0)
    op = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                              context=hip_context)
    O = oracle.Oracle(bg, box, opts)
    ref = st.copy()
    rhs_ref = O.compute_rhs(ref)
    d = to_dev(st)
    rhs = torch.zeros((11,) + box.interior_shape, dtype=torch.float64,
                      device="cuda")
    op.compute_rhs(rhs, d)
    assert_bitwise(rhs.cpu().numpy(), rhs_ref, "rhs")
    ref = st.copy()
    diag_ref, mask_ref = O.advance_rk4(ref, 0.05, 1)
    d = to_dev(st)
    mask = torch.zeros(box.interior_shape, dtype=torch.uint8, device="cuda")
    diag = op.advance_rk4(d, 0.05, 1, mask=mask)
    sl = (slice(None),) + box.interior_slices()
    assert_bitwise(d.cpu().numpy()[sl], ref[sl], "rk4 interior state")
    assert diag == diag_ref
    assert np.array_equal(mask.cpu().numpy(), mask_ref)
