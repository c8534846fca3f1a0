"""End-to-end pins of the oracle against numbers the reference logged.

GRHD_PROGRESS.md:180-186  Michel profile (512 radii, Gamma 5/3, r_c 8,
                          rho_c 1e-6): mass-flux err 2.11758e-16, Bernoulli
                          err 4.79408e-14, recovery err 1.55437e-07,
                          max v^2 0.0355675
GRHD_PROGRESS.md:423-428  LevelData Michel (18^3, dx 2, theta 1.2, 5 SSPRK2
                          steps to t = 0.1): mass 0.00172117 -> 0.00172335,
                          rel 0.00126778, max rho 2.06653e-06, max P
                          1.75256e-07, max v^2 0.0357728
GRHD_PROGRESS.md:308-311  LevelData binary-PN ring (18x18x5, dx 1, 2 boxes,
  and :339-343            theta 1.4, 4 SSPRK2 steps to t = 0.2 with static
                          and volume-time sources): mass 0.00296122 ->
                          0.00103393, max rho 1.15766e-05, max v^2
                          0.00223438

The reference printed six significant digits; the bar is agreement to that
precision (relative 5e-6 of the printed value, or the printed rounding).
"""
import numpy as np
import pytest

import michel
import oracle
from grdzhadzha_amd.options import OperatorOptions, RecoveryOptions


def close6(value, printed):
    """`value` rounds to the 6-significant-digit `printed` number."""
    return float("%.6g" % value) == printed


def test_michel_profile_anchor():
    cfg = michel.MichelConfig()
    c = michel.make_constants(cfg)
    opts = OperatorOptions(adiabatic_index=cfg.gamma,
                           recovery=RecoveryOptions(max_iterations=200,
                                                    tolerance=1e-14,
                                                    pressure_floor=1e-16))
    max_flux = max_bern = max_rec = max_v2 = 0.0
    n = 512
    for i in range(n):
        r = 4.0 + (40.0 - 4.0) * (i + 0.5) / n
        pt = michel.make_point(c, cfg, r)
        max_flux = max(max_flux, abs(pt["mass_flux"] - c["mass_flux"]) /
                       max(abs(c["mass_flux"]), 1e-300))
        max_bern = max(max_bern, abs(pt["bernoulli"] - c["bernoulli"]) /
                       max(abs(c["bernoulli"]), 1e-300))
        prim = np.array([pt["rho"], pt["eps"], pt["pressure"], pt["vr"], 0.0,
                         0.0])
        cons = oracle.prim2con(prim, cfg.gamma, pt["g"])
        ok, rec, _, _ = oracle.con2prim(cons, opts, pt["gu"])
        assert ok
        max_rec = max(max_rec, abs(rec[0] - prim[0]) / max(prim[0], 1e-300),
                      abs(rec[2] - prim[2]) / max(prim[2], 1e-300),
                      abs(rec[3] - prim[3]) / max(abs(prim[3]), 1e-300))
        g = pt["g"]
        max_v2 = max(max_v2, g[0][0] * prim[3] * prim[3])
    # sonic closed forms (GRHDMichelAccretion.cpp:384-393)
    assert abs(c["cs2"] - 1.0 / (16.0 - 3.0)) <= 1e-14
    assert abs(c["ur"] ** 2 - 1.0 / 16.0) <= 1e-14
    assert max_flux < 1e-9 and max_bern < 1e-9
    assert close6(max_rec, 1.55437e-07), max_rec
    assert close6(max_v2, 0.0355675), max_v2
    # the flux / Bernoulli errors are round-off; the reference's were
    # 2.11758e-16 and 4.79408e-14
    assert max_flux < 1e-15 and max_bern < 1e-13


@pytest.fixture(scope="module")
def leveldata_michel():
    cfg, box, bg, opts, state = michel.leveldata_setup()
    O = oracle.Oracle(bg, box, opts)
    return cfg, box, bg, opts, state, O


def test_leveldata_michel_anchor(leveldata_michel):
    cfg, box, bg, opts, state, O = leveldata_michel
    st = state.copy()
    geo = O.cell_geometry()
    d0, _ = O.recover_primitives(st)
    assert d0["num_failed_recoveries"] == 0
    assert d0["num_conserved_resets"] == 0
    di = michel.diagnostics(box, st, geo)
    t, steps = 0.0, 0
    while t < 0.1 - 1e-14:
        dt = min(min(O.compute_stable_dt(st, 0.2), 0.02), 0.1 - t)
        O.advance_ssprk2(st, dt)
        t += dt
        steps += 1
    df = michel.diagnostics(box, st, geo)
    rel = (df["mass"] - di["mass"]) / di["mass"]
    assert steps == 5
    assert close6(di["mass"], 0.00172117)
    assert close6(df["mass"], 0.00172335)
    assert close6(rel, 0.00126778)
    assert close6(df["max_rho"], 2.06653e-06)
    assert close6(df["max_pressure"], 1.75256e-07)
    assert close6(df["max_v2"], 0.0357728)


def test_leveldata_michel_two_box_split(leveldata_michel):
    """The reference runs this case as a 2-box x-split whose internal ghosts
    exchange() fills (GRHDLevelDataMichelAccretion.cpp:375-401).  With the
    internal ghosts of each half box copied from the other half (what
    exchange() does), the RHS of each half equals the single-box RHS bit
    for bit, so the one-box anchor above is the reference's 2-box run."""
    from grdzhadzha_amd.options import BoxSpec
    cfg, box, bg, opts, state, O = leveldata_michel
    st = state.copy()
    O.recover_primitives(st)
    rhs_full = O.compute_rhs(st.copy())
    g = box.ghosts
    for lo, hi in [((0, 0, 0), (8, 17, 17)), ((9, 0, 0), (17, 17, 17))]:
        sub = BoxSpec(lo, hi, g, box.dx, box.center, domain_lo=(0, 0, 0),
                      domain_hi=(17, 17, 17))
        x0 = lo[0] - box.glo[0] - g
        fab = np.ascontiguousarray(st[:, :, :, x0:x0 + hi[0] - lo[0] + 1 +
                                      2 * g]).copy()
        rhs = oracle.Oracle(bg, sub, opts).compute_rhs(fab)
        assert np.array_equal(rhs[:5], rhs_full[:5, :, :, lo[0]:hi[0] + 1])


def test_leveldata_binary_pn_ring_anchor():
    """The reference's moving-binary smoke run (tests/binary_ring.py):
    the time-dependent operator (binary metric at each stage time, volume
    sources) through recovery, CFL and SSPRK2 reproduces every printed
    digit.  The run resets cells near the point masses, which is where
    the mass goes."""
    import binary_ring
    box, halves, bg, opts, st = binary_ring.setup()
    O = oracle.Oracle(bg, box, opts, time_dependent=True)
    d0, _ = O.recover_primitives(st, time=0.0)
    assert d0["num_failed_recoveries"] == 0
    assert d0["num_conserved_resets"] == 0
    geo0 = [oracle.Oracle(bg, h, opts).cell_geometry(time=0.0)
            for h in halves]
    di = binary_ring.diagnostics(
        halves, [binary_ring.half_fab(st, box, h) for h in halves], geo0,
        opts)
    t = 0.0
    for _ in range(4):
        dt = min(0.05, 0.25 * O.compute_stable_dt(st, 0.2, time=t))
        O.advance_ssprk2(st, dt, time=t)
        t += dt
    geo1 = [oracle.Oracle(bg, h, opts).cell_geometry(time=t)
            for h in halves]
    df = binary_ring.diagnostics(
        halves, [binary_ring.half_fab(st, box, h) for h in halves], geo1,
        opts)
    a = binary_ring.ANCHOR
    assert close6(t, a["final_time"])
    assert close6(di["mass"], a["initial_mass"]), di
    assert close6(df["mass"], a["final_mass"]), df
    assert close6(df["max_rho"], a["max_rho"]), df
    assert close6(df["max_v2"], a["max_v2"]), df
