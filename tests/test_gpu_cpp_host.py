"""The C++ host mirror (include/grhd_operator.hpp) on the GPU: a C++
program drives FixedBGLevelDataFiniteVolumeOperator through the C-ABI over
the reference's 2-box LevelData Michel run (GRHD_PROGRESS.md:423-428) and
must reproduce the printed anchors and the Python path's final state bit
for bit."""
import os
import subprocess

import numpy as np
import pytest

import michel
import oracle
from grdzhadzha_amd.level import LevelOperator, split_box, sub_fab

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "level_driver")


def test_cpp_operator_michel_leveldata(hip_context, tmp_path):
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    cfg, box, bg, opts, state = michel.leveldata_setup()
    oracle.Oracle(bg, box, opts).recover_primitives(state)
    halves = split_box(box, 0, 9)
    lev = LevelOperator(bg, halves, opts, context=hip_context)
    fabs = [sub_fab(state, box, h) for h in halves]
    host = lev.pack(fabs).cpu().numpy()

    d = str(tmp_path)
    b0 = halves[0]
    with open(os.path.join(d, "level.txt"), "w") as f:
        f.write("%d %d %d %r %r %r %r\n" % ((b0.ghosts, 11, len(halves),
                                            b0.dx) + tuple(b0.center)))
        f.write("%d %d %d %d %d %d\n" % (tuple(b0.domain_lo) +
                                         tuple(b0.domain_hi)))
        for h in halves:
            f.write("%d %d %d %d %d %d\n" % (tuple(h.lo) + tuple(h.hi)))
    open(os.path.join(d, "params.bin"), "wb").write(bytes(opts.to_c()))
    open(os.path.join(d, "background.bin"), "wb").write(bytes(bg.to_c()))
    host.tofile(os.path.join(d, "state.bin"))
    res = subprocess.run([DRIVER, d, "0.2", "0.02", "0.1"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    steps, t, m0, m1, finite, dt_none, m_none = open(
        os.path.join(d, "result.txt")).read().split()
    assert int(steps) == 5 and int(finite) == 1
    # the same operator on an empty layout (a rank with no boxes)
    assert float(dt_none) == 0.2 / 1.0e-14 and float(m_none) == 0.0
    assert float("%.6g" % float(m0)) == 0.00172117
    assert float("%.6g" % float(m1)) == 0.00172335
    got = np.fromfile(os.path.join(d, "state_out.bin"))

    # the same loop through the Python mirror
    pool = lev.pack(fabs)
    tt = 0.0
    while tt < 0.1 - 1e-14:
        dt = min(min(lev.compute_stable_dt(pool, 0.2), 0.02), 0.1 - tt)
        lev.advance_ssprk2(pool, dt)
        tt += dt
    assert np.array_equal(got, pool.cpu().numpy())
    assert float(t) == tt

    # the same run with the face launch pinned through the C++ context
    # (walks of 2 faces, three launches on streams): the same state
    res = subprocess.run([DRIVER, d, "0.2", "0.02", "0.1", "2", "0"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert np.array_equal(np.fromfile(os.path.join(d, "state_out.bin")), got)


HOOKS = os.path.join(ROOT, "tests", "cpp", "_build", "hooks_driver")


def _hook_level(time_dependent, time):
    """Tests/GRHDFixedBGAMRHookTest/GRHDFixedBGAMRHookTest.cpp:592-616: a
    12 x 8 x 6 domain split in x at 5 | 6, 2 ghosts, dx 1, center
    (6, 4, 3), Gamma 5/3; make_primitive (:78-89) on every storage cell,
    floored with the default atmosphere (fill_state :91-119), theta 1.4,
    80 recovery iterations; Schwarzschild (:213-219) or the binary at
    t = 0.3 (m 0.5 + 0.5, separation 5, phase 0.2, softening 1;
    :283-296)."""
    from grdzhadzha_amd.options import (BoxSpec, CircularBinaryPNParams,
                                        KerrSchildParams, OperatorOptions,
                                        RecoveryOptions)
    gamma = 5.0 / 3.0
    opts = OperatorOptions(adiabatic_index=gamma, limiter_theta=1.4,
                           recovery=RecoveryOptions(max_iterations=80),
                           use_metric_volume_time_sources=time_dependent,
                           metric_time_derivative_step=1e-3)
    bg = (CircularBinaryPNParams(0.5, 0.5, 5.0, 0.0, 0.2, 0.0, 1.0)
          if time_dependent else KerrSchildParams(1.0, 0.0))
    center = (6.0, 4.0, 3.0)
    box = BoxSpec((0, 0, 0), (11, 7, 5), 2, 1.0, center)
    halves = [BoxSpec((0, 0, 0), (5, 7, 5), 2, 1.0, center,
                      domain_lo=(0, 0, 0), domain_hi=(11, 7, 5)),
              BoxSpec((6, 0, 0), (11, 7, 5), 2, 1.0, center,
                      domain_lo=(0, 0, 0), domain_hi=(11, 7, 5))]
    geo = oracle.Oracle(bg, box, opts).cell_geometry(time=time)
    st = np.zeros((11,) + box.storage_shape)
    nz, ny, nx = box.storage_shape
    for kk in range(nz):
        for jj in range(ny):
            for ii in range(nx):
                i, j, k = box.glo[0] + ii, box.glo[1] + jj, box.glo[2] + kk
                rho = 0.2 + 5.0e-4 * (i + 2 * j + 3 * k)
                eps = 0.08 + 1.0e-4 * (i + j + k)
                p = (gamma - 1.0) * rho * eps
                prim = [rho, eps, p, 8.0e-4 * (1.0 + i), -6.0e-4 * (1.0 + j),
                        4.0e-4 * (1.0 + k)]
                g6 = geo[0:6, kk, jj, ii]
                g = np.array([[g6[0], g6[1], g6[2]], [g6[1], g6[3], g6[4]],
                              [g6[2], g6[4], g6[5]]])
                prim = oracle.floors(prim, OperatorOptions(), g)
                st[0:5, kk, jj, ii] = oracle.prim2con(prim, gamma, g)
                st[5:11, kk, jj, ii] = prim
    return box, halves, bg, opts, st


def _cut(full, box, half, ghosts=True):
    g = half.ghosts if ghosts else 0
    x0 = half.lo[0] - box.lo[0] + (box.ghosts - g if ghosts else 0)
    w = half.hi[0] - half.lo[0] + 1 + 2 * g
    return np.ascontiguousarray(full[..., x0:x0 + w])


@pytest.mark.parametrize("time_dependent", [False, True],
                         ids=["kerr_schild", "binary_pn"])
@pytest.mark.parametrize("inverse_dx", [(1.0, 1.0, 1.0), (1.0, 1.25, 0.8)],
                         ids=["uniform", "anisotropic"])
def test_cpp_amr_hooks_and_driver_views(hip_context, tmp_path,
                                        time_dependent, inverse_dx):
    """The reference's AMR-hook test through the C++ mirror: hooks and
    driver views equal the level operator (checked in the C++ program),
    and the hook rhs, dt and SSPRK2 state equal the oracle's bit for bit,
    with the inverse_dx_U overloads carrying per-direction spacings."""
    import dataclasses
    assert os.path.exists(HOOKS), "run __graft_entry__.build() first"
    time = 0.3 if time_dependent else 0.0
    box, halves, bg, opts, st = _hook_level(time_dependent, time)
    d = str(tmp_path)
    with open(os.path.join(d, "level.txt"), "w") as f:
        f.write("%d %d %d %r %r %r %r\n" % ((2, 11, 2, 1.0) + box.center))
        f.write("0 0 0 11 7 5\n")
        for h in halves:
            f.write("%d %d %d %d %d %d\n" % (tuple(h.lo) + tuple(h.hi)))
    open(os.path.join(d, "params.bin"), "wb").write(bytes(opts.to_c()))
    open(os.path.join(d, "background.bin"), "wb").write(bytes(bg.to_c()))
    np.concatenate([_cut(st, box, h).ravel() for h in halves]).tofile(
        os.path.join(d, "state.bin"))
    res = subprocess.run([HOOKS, d, str(int(time_dependent)), repr(time)] +
                         [repr(v) for v in inverse_dx],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = open(os.path.join(d, "result.txt")).read().split("\n")
    assert lines[0] == "pass", lines
    hook_dt, dt = (float(v) for v in lines[1].split())

    ob = dataclasses.replace(box, inverse_dx=inverse_dx)
    O = oracle.Oracle(bg, ob, opts, time_dependent=time_dependent)
    rhs_ref = O.compute_rhs(st.copy(), time=time)
    assert O.compute_stable_dt(st.copy(), 0.2, time=time) == hook_dt
    adv_ref = st.copy()
    O.advance_ssprk2(adv_ref, dt, time=time)

    hr = np.fromfile(os.path.join(d, "hook_rhs.bin"))
    ha = np.fromfile(os.path.join(d, "hook_advanced.bin"))
    ro = so = 0
    for h in halves:
        ni = h.num_interior * 11
        got = hr[ro:ro + ni].reshape((11,) + h.interior_shape)
        assert np.array_equal(got, _cut(rhs_ref, box, h, ghosts=False))
        ns = 11 * int(np.prod(h.storage_shape))
        fab = ha[so:so + ns].reshape((11,) + h.storage_shape)
        sl = (slice(None),) + h.interior_slices()
        assert np.array_equal(fab[sl], _cut(adv_ref, box, h)[sl])
        ro += ni
        so += ns
