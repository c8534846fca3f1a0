"""Level (batched) path on the GPU, through the level C-ABI.

- A level of uneven boxes advanced in one launch per kernel equals each box
  advanced alone, bit for bit, and matches the oracle per box.
- A box split in two that share a ProblemDomain, with internal ghosts
  exchanged on the device, gives the single-box RHS and steps bit for bit;
  poisoned internal ghosts are repaired by the exchange
  (GRHDLevelDataBinaryPNSmoke.cpp:204-233, logged error 0).
- The reference's own 2-box LevelData Michel run (GRHD_PROGRESS.md:423-428)
  and its 2-box moving-binary ring run (GRHD_PROGRESS.md:308-311, :339-343)
  reproduce their printed numbers on the GPU.
"""
import dataclasses

import numpy as np
import pytest
import torch

import michel
import oracle
from grdzhadzha_amd import workloads
from grdzhadzha_amd.level import LevelOperator, split_box, sub_fab
from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def small_pn_level(nbox=10):
    bg, opts, levels, _ = workloads.binary_pn_levels(
        nlevels=2, nbox=nbox, base_cells=32, sides=(8, 16), base_dx=2.0)
    boxes, fabs, dt = levels[1]
    return bg, opts, boxes, fabs, dt


def test_level_rk4_equals_each_box_alone(hip_context):
    bg, opts, boxes, fabs, dt = small_pn_level()
    lev = LevelOperator(bg, boxes, opts, time_dependent=True,
                        context=hip_context)
    pool = lev.pack(fabs)
    mask = torch.zeros(lev.num_interior, dtype=torch.uint8, device="cuda")
    diag = lev.advance_rk4(pool, dt, nsteps=2, time=0.0, mask=mask)
    got = lev.unpack(pool)
    masks = lev.split_interior(mask)
    total = {k: 0 for k in diag}
    for n, (box, fab) in enumerate(zip(boxes, fabs)):
        from grdzhadzha_amd.operator import \
            TimeDependentFixedBGLevelDataFiniteVolumeOperator as TD
        op = TD(bg, box, opts, context=hip_context)
        st = torch.from_numpy(np.ascontiguousarray(fab)).cuda()
        m1 = torch.zeros(box.interior_shape, dtype=torch.uint8, device="cuda")
        d1 = op.advance_rk4(st, dt, 2, time=0.0, mask=m1)
        for k in total:
            total[k] += d1[k]
        assert np.array_equal(st.cpu().numpy(), got[n]), n
        assert np.array_equal(m1.cpu().numpy(), masks[n]), n
    assert total == diag
    # and the oracle agrees on two of the boxes
    for n in (0, len(boxes) - 1):
        ref = np.ascontiguousarray(fabs[n]).copy()
        dref, mref = oracle.Oracle(bg, boxes[n], opts,
                                   time_dependent=True).advance_rk4(
            ref, dt, 2, time=0.0)
        sl = (slice(None),) + boxes[n].interior_slices()
        assert_bitwise(got[n][sl], ref[sl], "box %d interior" % n)
        assert np.array_equal(masks[n], mref)


@pytest.fixture(scope="module")
def michel_split():
    cfg, box, bg, opts, state = michel.leveldata_setup()
    O = oracle.Oracle(bg, box, opts)
    O.recover_primitives(state)
    halves = split_box(box, 0, 9)
    return box, bg, opts, state, halves


def test_level_rhs_split_equals_single_box(hip_context, michel_split):
    box, bg, opts, state, halves = michel_split
    single = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                                  context=hip_context)
    st = torch.from_numpy(state.copy()).cuda()
    rhs = torch.zeros((11,) + box.interior_shape, dtype=torch.float64,
                      device="cuda")
    single.compute_rhs(rhs, st)
    rhs_full = rhs.cpu().numpy()
    lev = LevelOperator(bg, halves, opts, context=hip_context)
    fabs = [sub_fab(state, box, h) for h in halves]
    pool = lev.pack(fabs)
    r = torch.zeros(11 * lev.num_interior, dtype=torch.float64, device="cuda")
    lev.compute_rhs(r, pool)
    parts = lev.split_interior(r, 11)
    assert np.array_equal(parts[0][:5], rhs_full[:5, :, :, :9])
    assert np.array_equal(parts[1][:5], rhs_full[:5, :, :, 9:])

    # poisoned internal ghosts: the exchange repairs them
    poisoned = []
    for h, f in zip(halves, fabs):
        f = f.copy()
        x0 = h.glo[0]
        for x in range(h.glo[0], h.ghi[0] + 1):
            if (x < h.lo[0] or x > h.hi[0]) and 0 <= x <= 17:
                f[:, :, :, x - x0] = 1e30 + np.arange(11)[:, None, None]
        poisoned.append(f)
    pool = lev.pack(poisoned)
    r2 = torch.zeros_like(r)
    lev.compute_rhs(r2, pool)
    assert torch.equal(r2, r)


def test_level_michel_anchor_on_gpu(hip_context, michel_split):
    """The reference's 2-box LevelData Michel run: 18^3, dx 2, theta 1.2,
    SSPRK2 with dt = min(CFL 0.2 stable dt, 0.02) to t = 0.1."""
    box, bg, opts, state, halves = michel_split
    lev = LevelOperator(bg, halves, opts, context=hip_context)
    pool = lev.pack([sub_fab(state, box, h) for h in halves])

    def mass(p):
        return sum(f[0][h.interior_slices()].sum()
                   for f, h in zip(lev.unpack(p), halves))

    m0 = mass(pool)
    t, steps = 0.0, 0
    while t < 0.1 - 1e-14:
        dt = min(min(lev.compute_stable_dt(pool, 0.2), 0.02), 0.1 - t)
        lev.advance_ssprk2(pool, dt)
        t += dt
        steps += 1
    m1 = mass(pool)
    fabs = lev.unpack(pool)
    max_rho = max(f[5][h.interior_slices()].max()
                  for f, h in zip(fabs, halves))
    max_p = max(f[7][h.interior_slices()].max()
                for f, h in zip(fabs, halves))
    assert steps == 5
    assert float("%.6g" % m0) == 0.00172117
    assert float("%.6g" % m1) == 0.00172335
    assert float("%.6g" % ((m1 - m0) / m0)) == 0.00126778
    assert float("%.6g" % max_rho) == 2.06653e-06
    assert float("%.6g" % max_p) == 1.75256e-07
    # and the split level equals the single box stepped the same way
    single = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                                  context=hip_context)
    st = torch.from_numpy(state.copy()).cuda()
    t = 0.0
    while t < 0.1 - 1e-14:
        dt = min(min(single.compute_stable_dt(st, 0.2), 0.02), 0.1 - t)
        single.advance_ssprk2(st, dt)
        t += dt
    full = st.cpu().numpy()
    for f, h in zip(fabs, halves):
        sl = (slice(None),) + h.interior_slices()
        ref = sub_fab(full, box, h)[sl]
        assert np.array_equal(f[sl], ref)


def test_level_binary_pn_ring_anchor_on_gpu(hip_context):
    """The reference's binary-PN ring smoke run (tests/binary_ring.py) as
    the 2-box level it is: time-dependent operator, internal ghosts
    exchanged on the device, CFL from the device max 1/dt, 4 SSPRK2 steps
    to t = 0.2.  The printed numbers come out, and the state equals the
    oracle's single-box run bit for bit."""
    import binary_ring
    box, halves, bg, opts, state = binary_ring.setup()
    lev = LevelOperator(bg, halves, opts, time_dependent=True,
                        context=hip_context)
    pool = lev.pack([binary_ring.half_fab(state, box, h) for h in halves])
    d0 = lev.recover_primitives(pool, time=0.0)
    assert d0["num_failed_recoveries"] == 0
    assert d0["num_conserved_resets"] == 0
    geo0 = [oracle.Oracle(bg, h, opts).cell_geometry(time=0.0)
            for h in halves]
    di = binary_ring.diagnostics(halves, lev.unpack(pool), geo0, opts)
    O = oracle.Oracle(bg, box, opts, time_dependent=True)
    ref = state.copy()
    O.recover_primitives(ref, time=0.0)
    t = 0.0
    for _ in range(4):
        dt = min(0.05, 0.25 * lev.compute_stable_dt(pool, 0.2, time=t))
        dt_ref = min(0.05, 0.25 * O.compute_stable_dt(ref, 0.2, time=t))
        assert dt == dt_ref
        lev.advance_ssprk2(pool, dt, time=t)
        O.advance_ssprk2(ref, dt, time=t)
        t += dt
    fabs = lev.unpack(pool)
    geo1 = [oracle.Oracle(bg, h, opts).cell_geometry(time=t)
            for h in halves]
    df = binary_ring.diagnostics(halves, fabs, geo1, opts)
    a = binary_ring.ANCHOR
    assert float("%.6g" % t) == a["final_time"]
    assert float("%.6g" % di["mass"]) == a["initial_mass"]
    assert float("%.6g" % df["mass"]) == a["final_mass"]
    assert float("%.6g" % df["max_rho"]) == a["max_rho"]
    assert float("%.6g" % df["max_v2"]) == a["max_v2"]
    for f, h in zip(fabs, halves):
        sl = (slice(None),) + h.interior_slices()
        assert np.array_equal(f[sl], binary_ring.half_fab(ref, box, h)[sl])


def test_level_rk4_split_equals_single_box(hip_context, michel_split):
    box, bg, opts, state, halves = michel_split
    single = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                                  context=hip_context)
    st = torch.from_numpy(state.copy()).cuda()
    d_single = single.advance_rk4(st, 0.01, 3)
    lev = LevelOperator(bg, halves, opts, context=hip_context)
    pool = lev.pack([sub_fab(state, box, h) for h in halves])
    d_level = lev.advance_rk4(pool, 0.01, 3)
    assert d_single == d_level
    full = st.cpu().numpy()
    for f, h in zip(lev.unpack(pool), halves):
        sl = (slice(None),) + h.interior_slices()
        assert np.array_equal(f[sl], sub_fab(full, box, h)[sl])


def test_level_rejects_unowned_internal_ghosts(hip_context, michel_split):
    box, bg, opts, state, halves = michel_split
    lev = LevelOperator(bg, halves[:1], opts, context=hip_context)
    pool = lev.pack([sub_fab(state, box, halves[0])])
    with pytest.raises(ValueError):
        lev.advance_rk4(pool, 0.01, 1)


# ------------------------------------------------------------ diagnostics
def _interior(boxes, fabs, m):
    return np.concatenate([f[m][b.interior_slices()].ravel()
                           for b, f in zip(boxes, fabs)])


def test_level_conserved_sum_and_finiteness(hip_context):
    import math
    from grdzhadzha_amd.level import (compute_leveldata_conserved_sum,
                                      leveldata_conserved_state_is_finite)
    bg, opts, boxes, fabs, dt = small_pn_level(nbox=24)
    lev = LevelOperator(bg, boxes, opts, context=hip_context)
    pool = lev.pack(fabs)
    got = compute_leveldata_conserved_sum(lev, pool)
    assert got == lev.conserved_sum(pool)  # deterministic
    seq = oracle.leveldata_conserved_sum(boxes, fabs)
    eps = np.finfo(float).eps
    for m in range(5):
        v = _interior(boxes, fabs, m)
        exact = math.fsum(v)
        # double-double accumulation rounds the exact sum once
        assert abs(got[m] - exact) <= 0.5 * np.spacing(abs(exact)) + 1e-300
        # and differs from the reference's sequential sum by no more than
        # that sum's own rounding bound
        bound = (len(v) - 1) * eps / 2 * np.abs(v).sum()
        assert abs(got[m] - seq[m]) <= bound
    assert leveldata_conserved_state_is_finite(lev, pool)

    # a NaN in an interior cell of one box; a ghost-only inf does not count
    bad = [f.copy() for f in fabs]
    b = boxes[5]
    bad[5][2, b.lo[2] - b.glo[2], b.hi[1] - b.glo[1], b.lo[0] - b.glo[0]] = \
        np.nan
    assert not lev.conserved_state_is_finite(lev.pack(bad))
    assert not oracle.leveldata_conserved_state_is_finite(boxes, bad)
    s = lev.conserved_sum(lev.pack(bad))
    assert np.isnan(s[2]) and np.isnan(oracle.leveldata_conserved_sum(
        boxes, bad)[2])
    ghost = [f.copy() for f in fabs]
    ghost[7][0, 0, 0, 0] = np.inf
    p = lev.pack(ghost)
    assert lev.conserved_state_is_finite(p)
    assert lev.conserved_sum(p) == got


# ---------------------------------------------------------- periodic domains


def periodic_box(periodic):
    """16^3 box (dx 0.5) that is its own domain, on a flat background
    (Kerr-Schild with M = 0: the only metric that wraps consistently), with
    a smooth state of period 16 cells in every direction, so a periodic
    wrap is continuous.  Returns (background, options, box, state)."""
    from grdzhadzha_amd.options import BoxSpec, KerrSchildParams
    w = workloads.michel_box_32(n=16, half_width=4.0)
    b = w.box
    box = BoxSpec(b.lo, b.hi, b.ghosts, b.dx, b.center, periodic=periodic)
    gamma = w.options.adiabatic_index
    k, j, i = np.meshgrid(*[np.arange(box.glo[d], box.ghi[d] + 1)
                            for d in (2, 1, 0)], indexing="ij")
    ph = 2 * np.pi / 16
    rho = 1.0 + 0.2 * np.sin(ph * i) * np.cos(ph * j) + 0.1 * np.sin(ph * k)
    eps = 0.1 + 0.02 * np.cos(ph * (i + k))
    p = (gamma - 1.0) * rho * eps
    v = [0.1 * np.sin(ph * j), -0.05 * np.cos(ph * k), 0.03 * np.sin(ph * i)]
    v2 = sum(c * c for c in v)
    W = 1.0 / np.sqrt(1.0 - v2)
    rhW2 = rho * (1.0 + eps + p / rho) * W * W
    st = np.zeros((11,) + box.storage_shape)
    st[0] = rho * W
    for m in range(3):
        st[1 + m] = rhW2 * v[m]
    st[4] = rhW2 - p - st[0]
    st[5], st[6], st[7] = rho, eps, p
    for m in range(3):
        st[8 + m] = v[m]
    return KerrSchildParams(0.0, 0.0), w.options, box, st


def wrap_exchange(state, box):
    """Chombo's exchange on one box that is its own periodic domain: a ghost
    cell is filled when every coordinate in a non-periodic direction lies in
    the box, from the interior cell one period away; every other ghost is
    the outflow fill's."""
    st = state.copy()
    maps, inside = [], []
    for d in range(3):
        g = np.arange(box.glo[d], box.ghi[d] + 1)
        n = box.hi[d] - box.lo[d] + 1
        if box.periodic[d]:
            maps.append((g - box.lo[d]) % n + box.lo[d] - box.glo[d])
            inside.append(np.ones(g.size, bool))
        else:
            maps.append(g - box.glo[d])
            inside.append((g >= box.lo[d]) & (g <= box.hi[d]))
    out = [(np.arange(box.glo[d], box.ghi[d] + 1) < box.lo[d]) |
           (np.arange(box.glo[d], box.ghi[d] + 1) > box.hi[d])
           for d in range(3)]
    ok = (inside[2][:, None, None] & inside[1][None, :, None] &
          inside[0][None, None, :])
    periodic_out = np.zeros(ok.shape, bool)
    for d, sh in ((0, (1, 1, -1)), (1, (1, -1, 1)), (2, (-1, 1, 1))):
        if box.periodic[d]:
            periodic_out |= out[d].reshape(sh)
    kk, jj, ii = np.nonzero(ok & periodic_out)
    st[:, kk, jj, ii] = state[:, maps[2][kk], maps[1][jj], maps[0][ii]]
    return st


@pytest.mark.parametrize("periodic", [(1, 0, 0), (1, 0, 1), (1, 1, 1)])
def test_level_rhs_periodic_box_matches_oracle(hip_context, periodic):
    """A single box that is its own domain, periodic in some directions:
    the library wraps its own ghosts (device exchange) and fills the rest
    by outflow -- including the corner ghosts whose outflow source is a
    periodic ghost -- exactly as the oracle does after Chombo's exchange."""
    bg, opts, box, state = periodic_box(periodic)
    lev = LevelOperator(bg, [box], opts, context=hip_context)
    pool = lev.pack([state])
    r = torch.zeros(11 * lev.num_interior, dtype=torch.float64,
                    device="cuda")
    smask = torch.zeros(lev.num_interior, dtype=torch.uint8, device="cuda")
    lev.compute_rhs(r, pool, source_mask=smask)
    got = lev.split_interior(r, 11)[0]
    ref_state = wrap_exchange(state, box)
    rhs_ref, mask_ref = oracle.Oracle(bg, box, opts).compute_rhs(
        ref_state, with_mask=True)
    assert np.abs(rhs_ref[0:5]).max() > 0.0
    assert_bitwise(got, rhs_ref, "periodic rhs")
    assert np.array_equal(lev.split_interior(smask)[0], mask_ref)
    # the state's ghosts end as the oracle's: wrapped, then outflow-filled
    assert_bitwise(lev.unpack(pool)[0], ref_state, "periodic ghosts")


def test_level_rk4_periodic_split_equals_single_box(hip_context):
    """Periodic in x: two boxes splitting the domain along x exchange across
    the wrap every stage; the result equals the single periodic box bit for
    bit."""
    bg, opts, box, state = periodic_box((1, 0, 0))
    dt = 0.25 * box.dx
    single = LevelOperator(bg, [box], opts, context=hip_context)
    pool = single.pack([state])
    single.advance_rk4(pool, dt, nsteps=2)
    whole = single.unpack(pool)[0]
    halves = split_box(box, 0, 5)
    lev = LevelOperator(bg, halves, opts, context=hip_context)
    p2 = lev.pack([sub_fab(state, box, h) for h in halves])
    lev.advance_rk4(p2, dt, nsteps=2)
    for h, f in zip(halves, lev.unpack(p2)):
        sl = (slice(None),) + tuple(
            slice(h.lo[d] - box.glo[d], h.hi[d] - box.glo[d] + 1)
            for d in (2, 1, 0))
        hs = (slice(None),) + h.interior_slices()
        assert np.array_equal(f[hs], whole[sl])


def periodic_flat_box(n):
    """n^3 box (dx 0.5) that is its own fully periodic domain on a flat
    background, with a smooth state of period n cells: every source term
    vanishes and no floor or reset acts."""
    from grdzhadzha_amd.options import BoxSpec, KerrSchildParams
    w = workloads.michel_box_32(n=16, half_width=4.0)
    box = BoxSpec((0, 0, 0), (n - 1,) * 3, 3, 0.5, (0.25 * n,) * 3,
                  periodic=(1, 1, 1))
    gamma = w.options.adiabatic_index
    k, j, i = np.meshgrid(*[np.arange(box.glo[d], box.ghi[d] + 1)
                            for d in (2, 1, 0)], indexing="ij")
    ph = 2 * np.pi / n
    rho = 1.0 + 0.2 * np.sin(ph * i) * np.cos(ph * j) + 0.1 * np.sin(ph * k)
    eps = 0.1 + 0.02 * np.cos(ph * (i + k))
    p = (gamma - 1.0) * rho * eps
    v = [0.1 * np.sin(ph * j), -0.05 * np.cos(ph * k), 0.03 * np.sin(ph * i)]
    W = 1.0 / np.sqrt(1.0 - sum(c * c for c in v))
    rhW2 = rho * (1.0 + eps + p / rho) * W * W
    st = np.zeros((11,) + box.storage_shape)
    st[0] = rho * W
    for m in range(3):
        st[1 + m] = rhW2 * v[m]
    st[4] = rhW2 - p - st[0]
    st[5], st[6], st[7] = rho, eps, p
    for m in range(3):
        st[8 + m] = v[m]
    return KerrSchildParams(0.0, 0.0), w.options, box, st


@pytest.mark.parametrize("n", [128, 256])
def test_periodic_flat_box_conserves_all_five_totals_at_full_size(
        hip_context, n):
    """At the 128^3 and 256^3 sizes of the throughput configurations, a
    size-independent property: on a flat, fully periodic box the flux
    divergence telescopes and the sources vanish, so the level sums of D,
    S_i and tau stay constant through RK4 steps to rounding (relative to
    sum |U_c|), with no floors, resets or failures."""
    bg, opts, box, state = periodic_flat_box(n)
    lev = LevelOperator(bg, [box], opts, context=hip_context)
    pool = lev.pack([state])
    del state
    before = lev.conserved_sum(pool)
    G = int(np.prod(box.storage_shape))
    fab = pool[:11 * G].view((11,) + box.storage_shape)
    sl = (slice(0, 5),) + box.interior_slices()
    scale = fab[sl].abs().sum(dim=(1, 2, 3)).cpu().numpy()
    d_before = fab[0].clone()
    diag = lev.advance_rk4(pool, 0.25 * box.dx, nsteps=2)
    after = lev.conserved_sum(pool)
    # the flow moved: D changed by far more than rounding
    assert float((fab[0] - d_before).abs().max()) > 1e-5
    assert diag["num_failed_recoveries"] == 0
    assert diag["num_conserved_resets"] == 0
    assert lev.conserved_state_is_finite(pool)
    for c in range(5):
        assert abs(after[c] - before[c]) <= 1e-13 * scale[c], \
            (c, before[c], after[c], scale[c])


def tiles(box, edge):
    """The boxes of edge^3 cells tiling `box`, sharing its domain."""
    from grdzhadzha_amd.options import BoxSpec
    out = []
    n = [(box.hi[d] - box.lo[d] + 1) // edge for d in range(3)]
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                lo = (box.lo[0] + i * edge, box.lo[1] + j * edge,
                      box.lo[2] + k * edge)
                out.append(BoxSpec(lo, tuple(v + edge - 1 for v in lo),
                                   box.ghosts, box.dx, box.center,
                                   box.domain_lo, box.domain_hi,
                                   box.periodic))
    return out


def test_level_64_box_tiling_equals_single_box(hip_context):
    """64 boxes of 8^3 tiling the 32^3 Michel domain: every internal ghost
    comes from a neighbour through the exchange list (owner lookup by the
    bucket index), and two RK4 steps equal the single box bit for bit."""
    w = workloads.michel_box_32()
    single = LevelOperator(w.background, [w.box], w.options,
                           context=hip_context)
    pool = single.pack([w.state])
    single.advance_rk4(pool, w.dt, nsteps=2)
    whole = single.unpack(pool)[0]
    parts = tiles(w.box, 8)
    lev = LevelOperator(w.background, parts, w.options, context=hip_context)
    p2 = lev.pack([sub_fab(w.state, w.box, t) for t in parts])
    lev.advance_rk4(p2, w.dt, nsteps=2)
    for t, f in zip(parts, lev.unpack(p2)):
        sl = (slice(None),) + tuple(
            slice(t.lo[d] - w.box.glo[d], t.hi[d] - w.box.glo[d] + 1)
            for d in (2, 1, 0))
        assert np.array_equal(f[(slice(None),) + t.interior_slices()],
                              whole[sl])


def test_face_launch_forms_agree(hip_context):
    """Calls below 3 * 2^20 interior cells launch the three face directions
    as one fused kernel with the y/z faces in tiles (on a 32^3 box); larger
    calls launch them as three kernels on concurrent streams with walks of
    8.
    The 32^3 Michel box advanced alone equals the same box advanced in a
    level whose 160^3 second box (its own domain, at shifted indices) puts
    the call above the threshold, bit for bit."""
    from grdzhadzha_amd.options import BoxSpec
    w = workloads.michel_box_32()
    alone = LevelOperator(w.background, [w.box], w.options,
                          context=hip_context)
    pool = alone.pack([w.state])
    alone.advance_rk4(pool, w.dt, nsteps=2)
    ref = alone.unpack(pool)[0]
    big = workloads.michel_box_32(n=160, half_width=40.0)
    off = 1000
    b = big.box
    shifted = BoxSpec((b.lo[0] + off, b.lo[1], b.lo[2]),
                      (b.hi[0] + off, b.hi[1], b.hi[2]), b.ghosts, b.dx,
                      (b.center[0] + off * b.dx, b.center[1], b.center[2]))
    assert big.options == w.options and big.background == w.background
    lev = LevelOperator(w.background, [w.box, shifted], w.options,
                        context=hip_context)
    assert lev.num_interior >= 3 << 20
    p2 = lev.pack([w.state, big.state])
    del big
    lev.advance_rk4(p2, w.dt, nsteps=2)
    got = lev.unpack(p2)[0]
    assert np.array_equal(got, ref)


def test_level_plan_for_4096_boxes_builds_quickly(hip_context):
    """A 4096-box level (8^3 boxes tiling 128^3, shared domain): the launch
    plan, with its exchange list of 9.1M ghost cells, builds in seconds;
    the RHS is finite."""
    import time
    from grdzhadzha_amd.options import BoxSpec
    big = BoxSpec((0, 0, 0), (127, 127, 127), 3, 0.5, (32.0, 32.0, 32.0))
    parts = tiles(big, 8)
    assert len(parts) == 4096
    # an 8-periodic smooth state on a flat background tiles consistently
    bg, opts, _, fab = periodic_flat_box(8)
    lev = LevelOperator(bg, parts, opts, context=hip_context)
    pool = lev.pack([fab] * len(parts))
    r = torch.zeros(11 * lev.num_interior, dtype=torch.float64,
                    device="cuda")
    t0 = time.perf_counter()
    lev.compute_rhs(r, pool)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print("4096-box level: first compute_rhs (plan build included) %.2f s"
          % elapsed)
    assert elapsed < 20.0
    assert bool(torch.isfinite(r).all())


@pytest.mark.parametrize("mixed", [False, True])
def test_tuned_face_launch_equals_rule_launch(hip_context, mixed):
    """A tuning context times the ten face launches (walk length 1, 2, 4
    or 8, or tiles; fused or on streams) on its first call over a 10-box binary-PN
    level and keeps the fastest.  Its results equal those of a context on
    the size rules bit for bit (state, mask, counts), on the tuning call and
    on the next call, which reuses the choice without tuning again."""
    from grdzhadzha_amd.operator import Context
    bg, opts, levels, _ = workloads.binary_pn_levels(nbox=10, nlevels=1)
    boxes, fabs, dt = levels[0]
    opts = dataclasses.replace(opts, mixed_precision=mixed)
    tuned_ctx = Context()
    try:
        tuned_ctx.set_autotune(True)
        out = []
        for ctx in (hip_context, tuned_ctx):
            lev = LevelOperator(bg, boxes, opts, time_dependent=True,
                                context=ctx)
            pool = lev.pack(fabs)
            mask = torch.zeros(lev.num_interior, dtype=torch.uint8,
                               device="cuda")
            d1 = lev.advance_rk4(pool, dt, nsteps=1, time=0.0)
            d2 = lev.advance_rk4(pool, dt, nsteps=1, time=dt, mask=mask)
            out.append((pool.cpu().numpy(), mask.cpu().numpy(), d1, d2))
        choice = tuned_ctx.face_launch(mixed=mixed)
        rule = hip_context.face_launch(mixed=mixed)
    finally:
        tuned_ctx.close()
    print("face launch: rule", rule, "tuned", choice)
    assert choice["tuned"] and not rule["tuned"]
    assert choice["walk"] in (0, 1, 2, 4, 8) and choice["ms"] > 0.0
    (p0, m0, a0, b0), (p1, m1, a1, b1) = out
    assert np.array_equal(p0, p1)
    assert np.array_equal(m0, m1)
    assert (a0, b0) == (a1, b1)


@pytest.mark.parametrize("mixed", [False, True])
def test_every_pinned_face_launch_equals_rule_launch(hip_context, mixed):
    """Each of the ten face launches pinned on a context (tiles or walks of
    1, 2, 4 or 8 faces; fused or on streams) advances a 10-box binary-PN
    level exactly like the size rules do (state, mask, counts), in FP64
    and with FP32 faces; unpinning returns to the rule."""
    from grdzhadzha_amd.operator import Context
    bg, opts, levels, _ = workloads.binary_pn_levels(nbox=10, nlevels=1)
    boxes, fabs, dt = levels[0]
    opts = dataclasses.replace(opts, mixed_precision=mixed)

    def run(ctx):
        lev = LevelOperator(bg, boxes, opts, time_dependent=True,
                            context=ctx)
        pool = lev.pack(fabs)
        mask = torch.zeros(lev.num_interior, dtype=torch.uint8,
                           device="cuda")
        d = lev.advance_rk4(pool, dt, nsteps=1, time=0.0, mask=mask)
        return pool.cpu().numpy(), mask.cpu().numpy(), d

    p0, m0, d0 = run(hip_context)
    rule = hip_context.face_launch(mixed=mixed)
    ctx = Context()
    try:
        for walk in (0, 1, 2, 4, 8):
            for fused in (True, False):
                ctx.set_face_launch(walk, fused, mixed=mixed)
                p1, m1, d1 = run(ctx)
                assert ctx.face_launch(mixed=mixed) == {
                    "walk": walk, "fused": fused, "tuned": False, "ms": 0.0}
                assert np.array_equal(p1, p0), (walk, fused)
                assert np.array_equal(m1, m0), (walk, fused)
                assert d1 == d0
        ctx.set_face_launch(None, mixed=mixed)
        run(ctx)
        assert ctx.face_launch(mixed=mixed) == rule
    finally:
        ctx.close()


def test_tuning_is_remembered_per_level(hip_context):
    """A tuning context shared by two levels tunes each once: after a call
    on another level, returning to the first restores the choice measured
    on it (the same milliseconds, not a new measurement)."""
    from grdzhadzha_amd.operator import Context
    bg, opts, levels, _ = workloads.binary_pn_levels(nbox=6, nlevels=2)
    ctx = Context()
    try:
        ctx.set_autotune(True)
        levs = [LevelOperator(bg, boxes, opts, time_dependent=True,
                              context=ctx) for boxes, _, _ in levels]
        pools = [lev.pack(fabs) for lev, (_, fabs, _) in zip(levs, levels)]
        first = []
        for lev, pool, (_, _, dt) in zip(levs, pools, levels):
            lev.advance_rk4(pool, dt, nsteps=1, time=0.0)
            first.append(ctx.face_launch())
        for lev, pool, (_, _, dt), want in zip(levs, pools, levels, first):
            lev.advance_rk4(pool, dt, nsteps=1, time=dt)
            assert ctx.face_launch() == want
        assert all(f["tuned"] and f["ms"] > 0.0 for f in first)
    finally:
        ctx.close()


def test_failed_plan_is_not_reused(hip_context):
    """A level whose plan cannot be built (here: more storage cells than
    the failure key can number, refused before any allocation or launch)
    fails the same way on the next call with the same boxes; a half-built
    plan is never taken for a good one."""
    import ctypes
    from grdzhadzha_amd import abi
    from grdzhadzha_amd.options import BoxSpec, KerrSchildParams, \
        OperatorOptions
    big = BoxSpec((0, 0, 0), (16383, 16383, 16383), 3, 0.5, (0.0,) * 3)
    lb = (abi.GrhdLevelBox * 1)()
    lb[0].box = big.to_c()
    lb[0].state_offset = 0
    params = OperatorOptions().to_c()
    bg = KerrSchildParams(1.0, 0.0).to_c()
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    diag = abi.GrhdDiagnostics()
    for _ in range(2):
        rc = hip_context.lib.grhd_level_recover_primitives(
            hip_context.handle, lb, 1, ctypes.byref(params),
            ctypes.byref(bg), ctypes.c_double(0.0),
            ctypes.c_void_p(dummy.data_ptr()), 11, 0, None,
            ctypes.byref(diag), None)
        assert rc == abi.GRHD_ERR_INVALID_ARGUMENT


def test_level_calls_with_no_boxes(hip_context):
    """A process of a distributed level may own none of its boxes (the
    reference's loops over the DataIterator then do nothing).  Every level
    entry point with nbox = 0 and NULL boxes and pools succeeds with an
    empty level's outputs (zero counts, max 1/dt 0, zero sums, a finite
    state, a zero checkpoint grid, an untouched one on unpack), and still
    rejects bad parameters."""
    import ctypes
    from grdzhadzha_amd import abi
    from grdzhadzha_amd.options import KerrSchildParams, OperatorOptions
    lib, h = hip_context.lib, hip_context.handle
    params = OperatorOptions().to_c()
    bg = KerrSchildParams(1.0, 0.5).to_c()
    P = ctypes.byref
    OK = abi.GRHD_OK

    def fresh_diag():
        d = abi.GrhdDiagnostics()
        d.num_cells = d.num_failed_recoveries = 7
        return d

    err = abi.GrhdError()
    assert lib.grhd_level_compute_rhs(h, None, 0, P(params), P(bg), 0.0,
                                      None, 11, 0, None, None, P(err),
                                      None) == OK
    assert lib.grhd_level_add_conserved_rhs(h, None, 0, P(params), P(bg),
                                            0.0, None, 11, 0, None, P(err),
                                            None) == OK
    for call in ("recover", "update", "ssprk2", "rk4"):
        d = fresh_diag()
        if call == "recover":
            rc = lib.grhd_level_recover_primitives(
                h, None, 0, P(params), P(bg), 0.0, None, 11, 0, None, P(d),
                None)
        elif call == "update":
            rc = lib.grhd_level_update_conserved(
                h, None, 0, P(params), P(bg), None, 11, 0, None, 0.1, 1, 0.1,
                None, P(d), None)
        elif call == "ssprk2":
            rc = lib.grhd_level_advance_ssprk2(
                h, None, 0, P(params), P(bg), 0.0, 0.1, None, 11, 0, None,
                P(d), P(err), None)
        else:
            rc = lib.grhd_level_advance_rk4(
                h, None, 0, P(params), P(bg), 0.0, 0.1, 3, None, 11, 0, None,
                P(d), P(err), None)
        assert rc == OK, call
        assert (d.num_cells, d.num_failed_recoveries) == (0, 0), call
    m = ctypes.c_double(5.0)
    assert lib.grhd_level_compute_max_inverse_dt(
        h, None, 0, P(params), P(bg), 0.0, None, 11, 0, P(m), P(err),
        None) == OK
    assert m.value == 0.0
    sums = (ctypes.c_double * 5)(*([1.0] * 5))
    assert lib.grhd_level_conserved_sum(h, None, 0, None, 11, 0, sums,
                                        None) == OK
    assert list(sums) == [0.0] * 5
    fin = ctypes.c_int(0)
    assert lib.grhd_level_conserved_state_is_finite(h, None, 0, None, 11, 0,
                                                    P(fin), None) == OK
    assert fin.value == 1
    lo, hi = (ctypes.c_int * 3)(0, 0, 0), (ctypes.c_int * 3)(3, 2, 1)
    ck = torch.full((2 * 10 * 9 * 8,), 3.0, dtype=torch.float64,
                    device="cuda")
    assert lib.grhd_level_checkpoint_unpack(h, None, 0, ck.data_ptr(), 2, lo,
                                            hi, 3, None, None) == OK
    assert bool((ck == 3.0).all())
    assert lib.grhd_level_checkpoint_pack(h, None, 0, None, 2, lo, hi, 3,
                                          ck.data_ptr(), None) == OK
    assert bool((ck == 0.0).all())
    # parameters are still checked
    assert lib.grhd_level_compute_rhs(h, None, 0, P(params), P(bg), 0.0,
                                      None, 10, 0, None, None, P(err),
                                      None) == abi.GRHD_ERR_INVALID_ARGUMENT
    assert lib.grhd_level_advance_rk4(
        h, None, 0, P(params), P(bg), 0.0, 0.1, -1, None, 11, 0, None,
        None, P(err), None) == abi.GRHD_ERR_INVALID_ARGUMENT
    assert lib.grhd_level_advance_rk4(
        h, None, -1, P(params), P(bg), 0.0, 0.1, 1, None, 11, 0, None,
        None, P(err), None) == abi.GRHD_ERR_INVALID_ARGUMENT


def test_python_level_with_no_boxes(hip_context):
    """The Python level mirror over an empty box list (a rank with no
    boxes): every operation runs through the C-ABI's nbox = 0 path."""
    from grdzhadzha_amd.options import KerrSchildParams, OperatorOptions
    lev = LevelOperator(KerrSchildParams(1.0, 0.5), [], OperatorOptions(),
                        context=hip_context)
    pool = lev.new_pool()
    assert pool.numel() == 0 and lev.num_interior == 0
    rhs = torch.zeros(0, dtype=torch.float64, device="cuda")
    lev.compute_rhs(rhs, pool)
    lev.recover_primitives(pool)
    assert lev.compute_max_inverse_dt(pool) == 0.0
    assert lev.compute_stable_dt(pool, 0.5) == 0.5 / 1e-14
    d = lev.advance_rk4(pool, 0.1, nsteps=2)
    d2 = lev.advance_ssprk2(pool, 0.1)
    assert d == d2
    assert list(lev.conserved_sum(pool)) == [0.0] * 5
    assert lev.conserved_state_is_finite(pool)


@pytest.mark.parametrize("seed", range(4))
def test_random_split_level_equals_single_box(hip_context, seed):
    """A seeded random box (random_box_case of the parity tests) made its
    own ProblemDomain with random periodic directions, split at random
    planes into 2 to 4 boxes: two RK4 steps of the level -- internal
    ghosts exchanged on the device, wrapped where periodic -- equal the
    single box's bit for bit.

    Along a non-periodic direction every plane lies at least `ghosts`
    cells from the domain faces.  The reference fills a box's physical
    ghosts from the box's own edge cells
    (FixedBGFArrayBoxFiniteVolume.hpp:130-146), so a box that ends 1 or 2
    cells short of a domain face fills the ghosts past that face from
    itself, not from the cells the single box uses, and the two differ
    there (as they do in the reference)."""
    import dataclasses as dc
    from test_gpu_parity import random_box_case
    rng = np.random.default_rng(2000 + seed)
    box0, bg, opts, state, _ = random_box_case(seed)
    per = tuple(int(v) for v in rng.integers(0, 2, 3))
    box = dc.replace(box0, periodic=per)
    parts = [box]
    g = box.ghosts

    def planes(b, d):
        """Where box b may be cut along d: cells < at and >= at."""
        lo, hi = b.lo[d] + 1, b.hi[d]
        if not per[d]:
            lo, hi = max(lo, box.lo[d] + g), min(hi, box.hi[d] + 1 - g)
        return list(range(lo, hi + 1))

    for _ in range(int(rng.integers(1, 4))):
        cand = [(n, d) for n, b in enumerate(parts) for d in range(3)
                if planes(b, d)]
        if not cand:
            break
        n, d = cand[int(rng.integers(0, len(cand)))]
        b = parts.pop(n)
        opts_at = planes(b, d)
        at = opts_at[int(rng.integers(0, len(opts_at)))]
        parts.extend(split_box(b, d, at))
    dt = 0.2 * box.dx
    single = LevelOperator(bg, [box], opts, context=hip_context)
    pool = single.pack([state])
    single.advance_rk4(pool, dt, nsteps=2)
    whole = single.unpack(pool)[0]
    lev = LevelOperator(bg, parts, opts, context=hip_context)
    p2 = lev.pack([sub_fab(state, box, h) for h in parts])
    lev.advance_rk4(p2, dt, nsteps=2)
    assert len(parts) >= 2
    for h, f in zip(parts, lev.unpack(p2)):
        sl = (slice(None),) + tuple(
            slice(h.lo[d] - box.glo[d], h.hi[d] - box.glo[d] + 1)
            for d in (2, 1, 0))
        hs = (slice(None),) + h.interior_slices()
        assert np.array_equal(f[hs], whole[sl]), (seed, per, h.lo, h.hi)
