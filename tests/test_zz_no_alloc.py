"""grhd_advance_rk4_box_noalloc (no context, scratch in a static device
array) against grhd_advance_rk4_box, through the C-ABI.

Two calls in a row on one GPU, the second on a smaller box (the array
keeps the first call's contents, which the second must not see): state,
mask and recovery counts equal the context path's bit for bit.  A box whose
scratch exceeds GRHD_NOALLOC_SCRATCH_BYTES, or that is not its own
non-periodic domain, is rejected with GRHD_ERR_INVALID_ARGUMENT.
"""
import ctypes
import dataclasses

import numpy as np
import pytest
import torch

from grdzhadzha_amd import abi, workloads
from grdzhadzha_amd.operator import FixedBGLevelDataFiniteVolumeOperator
from grdzhadzha_amd.options import BoxSpec

pytestmark = pytest.mark.gpu


def noalloc_rk4(w, state, nsteps, mask):
    lib = abi.load_library()
    diag = abi.GrhdDiagnostics()
    err = abi.GrhdError()
    box = w.box.to_c()
    params = w.options.to_c()
    bg = w.background.to_c()
    rc = lib.grhd_advance_rk4_box_noalloc(
        ctypes.byref(box), ctypes.byref(params), ctypes.byref(bg),
        ctypes.c_double(0.0), ctypes.c_double(w.dt), int(nsteps),
        ctypes.c_void_p(state.data_ptr()), state.shape[0], 0,
        ctypes.c_void_p(mask.data_ptr()) if mask is not None else None,
        ctypes.byref(diag), ctypes.byref(err), None)
    return rc, diag.as_dict(), err


@pytest.mark.parametrize("first,second,guarded", [(48, 16, False),
                                                  (32, 8, False),
                                                  (48, 16, True)])
def test_noalloc_rk4_equals_context_rk4(hip_context, first, second,
                                        guarded):
    """guarded: mixed precision with FP64 faces near floors and where the
    shift is large, whose face list also lives in the static array."""
    for n, nsteps in ((first, 2), (second, 3)):
        w = workloads.torus_box(n=n, R0=0.3 * n, sigma=0.08 * n, nsteps=1)
        if guarded:
            w.options = dataclasses.replace(w.options, mixed_precision=True,
                                            mixed_fp64_margin=2.0,
                                            mixed_fp64_shift=0.2)
        print("noalloc case n", n, "steps", nsteps, "guarded", guarded,
              flush=True)
        op = FixedBGLevelDataFiniteVolumeOperator(w.background, w.box,
                                                  w.options,
                                                  context=hip_context)
        a = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
        ma = torch.zeros(w.box.interior_shape, dtype=torch.uint8,
                         device="cuda")
        da = op.advance_rk4(a, w.dt, nsteps, mask=ma)
        b = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
        mb = torch.zeros(w.box.interior_shape, dtype=torch.uint8,
                         device="cuda")
        rc, db, _ = noalloc_rk4(w, b, nsteps, mb)
        assert rc == abi.GRHD_OK
        assert db == da
        if n >= 32:
            assert da["num_conserved_resets"] > 0  # the reset branch ran
        assert torch.equal(a, b)
        assert torch.equal(ma, mb)


@pytest.mark.parametrize("seed", range(4))
def test_noalloc_random_boxes(hip_context, seed):
    """The seeded random boxes of the parity tests (2..19 cells a side,
    random spin, Gamma, limiter, near-floor cells), one after another on
    the same static array: two RK4 steps through the allocation-free entry
    equal the context path's state, mask and counts bit for bit."""
    import types
    from test_gpu_parity import random_box_case
    box, bg, opts, st0, _ = random_box_case(seed)
    w = types.SimpleNamespace(box=box, background=bg, options=opts,
                              dt=0.2 * box.dx)
    print("noalloc random box", seed, box.lo, box.hi, flush=True)
    op = FixedBGLevelDataFiniteVolumeOperator(bg, box, opts,
                                              context=hip_context)
    a = torch.from_numpy(st0.copy()).cuda()
    ma = torch.zeros(box.interior_shape, dtype=torch.uint8, device="cuda")
    da = op.advance_rk4(a, w.dt, 2, mask=ma)
    b = torch.from_numpy(st0.copy()).cuda()
    mb = torch.zeros(box.interior_shape, dtype=torch.uint8, device="cuda")
    rc, db, _ = noalloc_rk4(w, b, 2, mb)
    assert rc == abi.GRHD_OK
    assert db == da
    assert torch.equal(a, b)
    assert torch.equal(ma, mb)


def test_noalloc_rejects_boxes_beyond_its_scratch(hip_context):
    w = workloads.torus_box(n=16, nsteps=1)
    big = BoxSpec((0, 0, 0), (95, 95, 95), 3, 0.5, (24.0, 24.0, 24.0))
    w.box = big
    state = torch.zeros((11,) + big.storage_shape, dtype=torch.float64,
                        device="cuda")
    rc, _, _ = noalloc_rk4(w, state, 1, None)
    assert rc == abi.GRHD_ERR_INVALID_ARGUMENT
    # and a box that is not its own domain (the context path refuses it too)
    w = workloads.torus_box(n=16, nsteps=1)
    w.box = BoxSpec(w.box.lo, w.box.hi, 3, w.box.dx, w.box.center,
                    periodic=(1, 0, 0))
    state = torch.from_numpy(np.ascontiguousarray(w.state)).cuda()
    rc, _, _ = noalloc_rk4(w, state, 1, None)
    assert rc == abi.GRHD_ERR_INVALID_ARGUMENT


def test_noalloc_reports_recovery_errors_like_the_context_path(
        hip_context):
    w = workloads.michel_box_32(n=16, half_width=4.0)
    st = np.ascontiguousarray(w.state).copy()
    g = w.box.ghosts
    st[0, g + 5, g + 6, g + 7] = -1.0  # D <= 0 at interior (7, 6, 5)
    b = torch.from_numpy(st).cuda()
    rc, _, err = noalloc_rk4(w, b, 1, None)
    assert rc == abi.GRHD_ERR_RECOVERY_RHS
    assert tuple(err.iv) == (7, 6, 5)
