"""Mixed precision (FP32 reconstruction, face metric and HLLE; FP64
recovery, divergence sum, sources and update) against the bit-exact FP64
path on the GPU.  Tolerances: after a few RK4 steps the relative L1 error
of every conserved variable stays below 1e-6 -- FP32 rounding (6e-8 per
operation) enters each face flux at a few ulp and the state through
dt * (F_hi - F_lo) / dx, i.e. at about CFL x few ulp per step."""
import dataclasses

import numpy as np
import pytest
import torch

from grdzhadzha_amd import workloads
from grdzhadzha_amd.operator import (
    FixedBGLevelDataFiniteVolumeOperator,
    TimeDependentFixedBGLevelDataFiniteVolumeOperator)

pytestmark = pytest.mark.gpu


def rel_l1(a, b, box):
    sl = box.interior_slices()
    out = []
    for c in range(5):
        ref = b[c][sl]
        out.append(float(np.abs(a[c][sl] - ref).sum() / np.abs(ref).sum()))
    return out


def run(w, opts, steps, ctx, td=False):
    cls = TimeDependentFixedBGLevelDataFiniteVolumeOperator if td else \
        FixedBGLevelDataFiniteVolumeOperator
    op = cls(w.background, w.box, opts, context=ctx)
    st = torch.from_numpy(np.ascontiguousarray(w.state).copy()).cuda()
    d = op.advance_rk4(st, w.dt, steps)
    return st.cpu().numpy(), d


def test_mixed_torus_close_to_fp64(hip_context):
    w = workloads.torus_box(n=32, R0=5.0, sigma=1.5)
    m = dataclasses.replace(w.options, mixed_precision=True)
    s64, d64 = run(w, w.options, 4, hip_context)
    s32, d32 = run(w, m, 4, hip_context)
    errs = rel_l1(s32, s64, w.box)
    assert max(errs) < 1e-6, errs
    assert min(errs) > 0.0  # the FP32 faces really ran
    assert d32["num_cells"] == d64["num_cells"]
    # deterministic
    s32b, _ = run(w, m, 4, hip_context)
    assert np.array_equal(s32, s32b)


def test_mixed_binary_pn_volume_sources(hip_context):
    bg, opts, levels, _ = workloads.binary_pn_levels(
        nlevels=1, nbox=3, base_cells=32, sides=(16,), base_dx=2.0)
    boxes, fabs, dt = levels[0]
    m = dataclasses.replace(opts, mixed_precision=True)
    for box, fab in zip(boxes, fabs):
        w = workloads.Workload("pn", box, bg, opts, fab, dt, 2)
        s64, _ = run(w, opts, 2, hip_context, td=True)
        s32, _ = run(w, m, 2, hip_context, td=True)
        assert np.all(np.isfinite(s32))
        assert max(rel_l1(s32, s64, box)) < 1e-6


def test_fp64_path_ignores_nothing_when_flag_off(hip_context):
    """The default options run the bit-exact FP64 faces: two operators, one
    with an explicit mixed_precision=False, agree bit for bit."""
    w = workloads.torus_box(n=16)
    a, _ = run(w, w.options, 1, hip_context)
    b, _ = run(w, dataclasses.replace(w.options, mixed_precision=False), 1,
               hip_context)
    assert np.array_equal(a, b)


def test_batched_mixed_level_equals_each_box_alone(hip_context):
    """The box-set configuration in mixed precision: one level call over a
    seeded binary-PN box table gives, box for box, exactly what the per-box
    mixed path gives (same kernels, same operation order), masks and counts
    included; and it stays within 1e-6 of FP64 after two steps."""
    from grdzhadzha_amd.level import LevelOperator
    from grdzhadzha_amd.operator import \
        TimeDependentFixedBGLevelDataFiniteVolumeOperator as TD
    bg, opts, levels, _ = workloads.binary_pn_levels(nbox=10, nlevels=1)
    boxes, fabs, dt = levels[0]
    m = dataclasses.replace(opts, mixed_precision=True)
    lev = LevelOperator(bg, boxes, m, time_dependent=True,
                        context=hip_context)
    pool = lev.pack(fabs)
    mask = torch.zeros(lev.num_interior, dtype=torch.uint8, device="cuda")
    diag = lev.advance_rk4(pool, dt, nsteps=2, time=0.0, mask=mask)
    got = lev.unpack(pool)
    masks = lev.split_interior(mask)
    total = {k: 0 for k in diag}
    num = den = 0.0
    for n, (box, fab) in enumerate(zip(boxes, fabs)):
        st = torch.from_numpy(np.ascontiguousarray(fab)).cuda()
        m1 = torch.zeros(box.interior_shape, dtype=torch.uint8, device="cuda")
        d1 = TD(bg, box, m, context=hip_context).advance_rk4(
            st, dt, 2, time=0.0, mask=m1)
        for k in total:
            total[k] += d1[k]
        assert np.array_equal(st.cpu().numpy(), got[n]), n
        assert np.array_equal(m1.cpu().numpy(), masks[n]), n
        ref = torch.from_numpy(np.ascontiguousarray(fab)).cuda()
        TD(bg, box, opts, context=hip_context).advance_rk4(ref, dt, 2,
                                                           time=0.0)
        sl = (slice(0, 5),) + box.interior_slices()
        r = ref.cpu().numpy()[sl]
        num += float(np.abs(got[n][sl] - r).sum())
        den += float(np.abs(r).sum())
    assert total == diag
    assert num / den < 1e-6, num / den
