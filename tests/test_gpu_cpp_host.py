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
    steps, t, m0, m1, finite = open(os.path.join(d, "result.txt")).read(
    ).split()
    assert int(steps) == 5 and int(finite) == 1
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
