"""Guarded mixed precision (grhd_params.mixed_fp64_margin /
mixed_fp64_shift): FP32 faces except beside cells near a floor or the
velocity cap, or where the shift is large, whose faces are computed in FP64
from a device-built list.

- Guarding every cell gives the FP64 path bit for bit (single box and a
  level of binary-PN boxes): every face then goes through the FP64 list
  kernel, whose flux must equal the FP64 face kernels'.
- With the shift test alone, each cell's right-hand side equals the FP64
  path's bit for bit when all six of its faces are FP64 (the cell itself
  is guarded), and the plain mixed path's bit for bit when none is (the
  cell and its six neighbours unguarded).  The guard is classified on the
  host from the oracle's geometry; cells within 1e-9 of the threshold are
  left out.
- Results do not depend on the order in which the list fills.
- Out-of-range parameters are rejected.
"""
import dataclasses

import numpy as np
import pytest
import torch

import oracle
from grdzhadzha_amd import workloads
from grdzhadzha_amd.level import LevelOperator
from grdzhadzha_amd.operator import (
    FixedBGLevelDataFiniteVolumeOperator,
    TimeDependentFixedBGLevelDataFiniteVolumeOperator, new_rhs)

pytestmark = pytest.mark.gpu


def step(w, opts, steps, ctx):
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, opts,
                                              context=ctx)
    st = torch.from_numpy(np.ascontiguousarray(w.state).copy()).cuda()
    m = torch.zeros(w.box.interior_shape, dtype=torch.uint8, device="cuda")
    d = op.advance_rk4(st, w.dt, steps, mask=m)
    return st.cpu().numpy(), m.cpu().numpy(), d


def guarded(opts, **kw):
    return dataclasses.replace(opts, mixed_precision=True, **kw)


def test_guarding_every_cell_is_the_fp64_path(hip_context):
    w = workloads.torus_box(n=32, R0=8.0, sigma=2.0)
    ref = step(w, w.options, 4, hip_context)
    assert ref[2]["num_conserved_resets"] > 0
    for kw in (dict(mixed_fp64_margin=1e300), dict(mixed_fp64_shift=1e-300)):
        got = step(w, guarded(w.options, **kw), 4, hip_context)
        print("guard everything:", kw, flush=True)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
        assert got[2] == ref[2]


def test_guarding_every_cell_of_a_level_is_the_fp64_path(hip_context):
    bg, opts, levels, _ = workloads.binary_pn_levels(nbox=10, nlevels=1)
    boxes, fabs, dt = levels[0]
    out = []
    for o in (opts, guarded(opts, mixed_fp64_margin=1e300)):
        lev = LevelOperator(bg, boxes, o, time_dependent=True,
                            context=hip_context)
        pool = lev.pack(fabs)
        mask = torch.zeros(lev.num_interior, dtype=torch.uint8,
                           device="cuda")
        d = lev.advance_rk4(pool, dt, nsteps=2, time=0.0, mask=mask)
        out.append((lev.unpack(pool), mask.cpu().numpy(), d))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def shift_guard(w, s):
    """Storage-cell guard of the shift test (gamma_ij beta^i beta^j >= s^2)
    from the oracle's geometry, and the cells too close to call."""
    geo = oracle.Oracle(w.background, w.box, w.options).cell_geometry()
    g = np.empty((3, 3) + geo.shape[1:])
    for n, (i, j) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2),
                                (2, 2))):
        g[i, j] = g[j, i] = geo[n]
    beta = geo[13:16]
    b2 = np.einsum("ij...,i...,j...->...", g, beta, beta)
    close = np.abs(b2 / (s * s) - 1.0) < 1e-9
    return b2 >= s * s, close


def rhs(w, opts, ctx):
    op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box, opts,
                                              context=ctx)
    st = torch.from_numpy(np.ascontiguousarray(w.state).copy()).cuda()
    r = new_rhs(w.box, ncomp=st.shape[0])
    op.compute_rhs(r, st)
    return r.cpu().numpy()[:5]


def test_shift_guard_picks_fp64_or_fp32_per_face(hip_context):
    w = workloads.torus_box(n=32, R0=5.0, sigma=1.5)
    s = 0.3
    r64 = rhs(w, w.options, hip_context)
    r32 = rhs(w, dataclasses.replace(w.options, mixed_precision=True),
              hip_context)
    rg = rhs(w, guarded(w.options, mixed_fp64_shift=s), hip_context)
    G, close = shift_guard(w, s)
    lo, hi = w.box.lo, w.box.hi
    g0 = [l - gl for l, gl in zip(lo, w.box.glo)]

    def view(a, dk, dj, di):
        k0, j0, i0 = g0[2] + dk, g0[1] + dj, g0[0] + di
        nz, ny, nx = w.box.interior_shape
        return a[k0:k0 + nz, j0:j0 + ny, i0:i0 + nx]

    self_g = view(G, 0, 0, 0)
    nbrs = [view(G, *o) for o in ((0, 0, -1), (0, 0, 1), (0, -1, 0),
                                  (0, 1, 0), (-1, 0, 0), (1, 0, 0))]
    amb = view(close, 0, 0, 0).copy()
    for o in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0),
              (1, 0, 0)):
        amb |= view(close, *o)
    all64 = self_g & ~amb
    all32 = ~self_g & ~np.logical_or.reduce(nbrs) & ~amb
    print("cells with six FP64 faces %d, six FP32 faces %d, of %d" %
          (all64.sum(), all32.sum(), all64.size), flush=True)
    assert all64.sum() > 1000 and all32.sum() > 1000
    for c in range(5):
        assert np.array_equal(rg[c][all64], r64[c][all64]), c
        assert np.array_equal(rg[c][all32], r32[c][all32]), c
    # and the FP32 faces really differ from FP64 there
    assert not np.array_equal(r32[0][all32], r64[0][all32])


def test_guarded_steps_are_deterministic(hip_context):
    w = workloads.torus_box(n=48, R0=12.0, sigma=3.0)
    o = guarded(w.options, mixed_fp64_shift=0.1, mixed_fp64_margin=2.0)
    a = step(w, o, 3, hip_context)
    b = step(w, o, 3, hip_context)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2] == b[2]
    ref = step(w, w.options, 3, hip_context)
    sl = w.box.interior_slices()
    errs = [float(np.abs(a[0][c][sl] - ref[0][c][sl]).sum() /
                  np.abs(ref[0][c][sl]).sum()) for c in range(5)]
    print("guarded 48^3 three steps, L1 rel", errs, flush=True)
    assert 0.0 < max(errs) < 1e-7


@pytest.mark.parametrize("kw", [dict(mixed_fp64_margin=0.5),
                                dict(mixed_fp64_margin=float("nan")),
                                dict(mixed_fp64_margin=float("inf")),
                                dict(mixed_fp64_shift=-1.0),
                                dict(mixed_fp64_shift=float("nan"))])
def test_out_of_range_guard_parameters_are_rejected(hip_context, kw):
    w = workloads.torus_box(n=8)
    with pytest.raises(ValueError):
        step(w, guarded(w.options, **kw), 1, hip_context)


def test_time_dependent_guarded_binary_pn(hip_context):
    """The binary-PN background through the time-dependent operator with
    both guards on: finite and within 1e-6 of FP64 after two steps."""
    bg, opts, levels, _ = workloads.binary_pn_levels(
        nlevels=1, nbox=2, base_cells=32, sides=(16,), base_dx=2.0)
    boxes, fabs, dt = levels[0]
    o = guarded(opts, mixed_fp64_margin=10.0, mixed_fp64_shift=0.05)
    for box, fab in zip(boxes, fabs):
        out = []
        for oo in (opts, o):
            op = TimeDependentFixedBGLevelDataFiniteVolumeOperator(
                bg, box, oo, context=hip_context)
            st = torch.from_numpy(np.ascontiguousarray(fab).copy()).cuda()
            op.advance_rk4(st, dt, 2)
            out.append(st.cpu().numpy())
        sl = box.interior_slices()
        assert np.all(np.isfinite(out[1]))
        errs = [float(np.abs(out[1][c][sl] - out[0][c][sl]).sum() /
                      np.abs(out[0][c][sl]).sum()) for c in range(5)]
        assert max(errs) < 1e-6, errs
