"""RK4 against the mathematics, on the CPU oracle.

The reference integrates GRHD only with SSPRK2, so nothing in it pins the
RK4 built on its per-stage seam (FixedBGAMRLevelHooks.hpp:40-86; RK4 per
paper/paper.md:76).  These checks do:
- classical RK4 is 4th order in time: on a smooth flow (tests/smooth_flow.py)
  at fixed dx, halving dt divides the error against a dt/32 solution by 16;
  on the binary-PN background the metric changes in time, so the stage
  times are pinned too (a stage at the wrong time costs the order);
- advance_rk4 equals RK4 assembled here from the oracle's own per-stage
  calls (compute_rhs at t, t + dt/2, t + dt/2, t + dt; update_conserved with
  dt/2, dt/2, dt; the combination dt/6, dt/3, dt/3, dt/6) bit for bit;
- the same assembly with one planted mistake -- a stage at the wrong time,
  or a wrong weight -- loses the order, so the order check would catch it.
"""
import numpy as np
import pytest

import oracle
import smooth_flow

T_END = 1.0
STEPS = [4, 8, 16, 32]
REF_STEPS = 128


def _interior(box, st):
    return st[(slice(0, 5),) + box.interior_slices()].copy()


def _run(O, box, st0, nsteps, driver=None):
    st = st0.copy()
    dt = T_END / nsteps
    if driver is None:
        diag, _ = O.advance_rk4(st, dt, nsteps)
        assert diag["num_conserved_resets"] == 0
        assert diag["num_failed_recoveries"] == 0
    else:
        driver(O, st, dt, nsteps)
    return _interior(box, st)


def rk4_by_stages(O, st, dt, nsteps, t0=0.0, c=(0.5, 0.5, 1.0),
                  b=(1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)):
    """RK4 from per-stage oracle calls: stage s evaluates the rhs at
    t + c[s-1] dt (stage 1 at t); stage states U0 + a_s dt k_{s-1} are
    recovered at their stage time; the step combines the k's with the
    weights b and recovers at t + dt.  The weights are applied as dt/6 and
    dt/3, the way the oracle forms them, so the default is bit-exact."""
    for step in range(nsteps):
        t = t0 + step * dt
        u0 = st.copy()
        ks = []
        stage = st
        times = (t, t + c[0] * dt, t + c[1] * dt, t + c[2] * dt)
        a = (c[0] * dt, c[1] * dt, c[2] * dt)
        for s in range(4):
            k = O.compute_rhs(stage, time=times[s])
            ks.append(k)
            if s < 3:
                stage = u0.copy()
                O.update_conserved(stage, k, a[s], recover=True,
                                   recovery_time=times[s + 1])
        acc = u0.copy()
        w = [dt / (1.0 / bb) for bb in b]
        for s in range(4):
            O.update_conserved(acc, ks[s], w[s], recover=(s == 3),
                               recovery_time=t + dt)
        st[...] = acc


@pytest.fixture(scope="module", params=smooth_flow.cases(),
                ids=lambda c: c[0])
def case(request):
    name, bg, box, td = request.param
    O = oracle.Oracle(bg, box, smooth_flow.options(td), time_dependent=td)
    st0 = smooth_flow.state(box, bg)
    ref = _run(O, box, st0, REF_STEPS)
    return name, O, box, st0, ref


def test_rk4_is_fourth_order_in_time(case):
    name, O, box, st0, ref = case
    res = [_run(O, box, st0, n) for n in STEPS]
    for c in range(5):
        errs, orders = smooth_flow.observed_orders([r[c] for r in res],
                                                   ref[c])
        print(name, "component", c, "errors", errs, "orders", orders)
        assert errs[-1] > 1e-13, "errors too small to resolve an order"
        assert min(orders) >= 3.8, (name, c, errs, orders)


def test_rk4_equals_stagewise_assembly(case):
    name, O, box, st0, _ = case
    a = st0.copy()
    O.advance_rk4(a, T_END / 4, 4)
    b = st0.copy()
    rk4_by_stages(O, b, T_END / 4, 4)
    # interior conserved and primitive components; the ghosts hold
    # whichever stage's outflow fill each driver happened to keep
    inner = (slice(None),) + box.interior_slices()
    assert np.array_equal(a[inner], b[inner])


def test_planted_stage_time_mistake_loses_order():
    """Stage 2 at t instead of t + dt/2: exact on a static metric, but on
    the binary the error is first order at small dt."""
    name, bg, box, td = smooth_flow.cases()[1]
    O = oracle.Oracle(bg, box, smooth_flow.options(td), time_dependent=td)
    st0 = smooth_flow.state(box, bg)
    ref = _run(O, box, st0, REF_STEPS)
    orig = oracle.Oracle.compute_rhs
    calls = {"n": 0}

    def wrong_time_driver(O_, st, dt, nsteps):
        # the second rhs of every step sees the step's start time
        def rhs(state, time=0.0, with_mask=False):
            calls["n"] += 1
            if calls["n"] % 4 == 2:
                time = time - 0.5 * dt
            return orig(O_, state, time=time, with_mask=with_mask)
        O_.compute_rhs = rhs
        try:
            calls["n"] = 0
            rk4_by_stages(O_, st, dt, nsteps)
        finally:
            del O_.compute_rhs

    res = [_run(O, box, st0, n, wrong_time_driver) for n in STEPS]
    errs, orders = smooth_flow.observed_orders(res, ref)
    print("wrong stage time: errors", errs, "orders", orders)
    assert orders[-1] < 2.5


def test_planted_weight_mistake_loses_order():
    """Equal weights 1/4 instead of 1/6, 1/3, 1/3, 1/6: still consistent,
    but only first order."""
    name, bg, box, td = smooth_flow.cases()[0]
    O = oracle.Oracle(bg, box, smooth_flow.options(td), time_dependent=td)
    st0 = smooth_flow.state(box, bg)
    ref = _run(O, box, st0, REF_STEPS)

    def wrong_weights(O_, st, dt, nsteps):
        rk4_by_stages(O_, st, dt, nsteps, b=(0.25, 0.25, 0.25, 0.25))

    res = [_run(O, box, st0, n, wrong_weights) for n in STEPS]
    errs, orders = smooth_flow.observed_orders(res, ref)
    print("wrong weights: errors", errs, "orders", orders)
    assert orders[-1] < 2.5
