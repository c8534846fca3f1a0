"""Unit pins of the oracle's per-cell numerics (CPU only).

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) fails it: closed forms, invariants,
finite differences of the analytic derivatives, the reference tests' own
known-answer states and tolerances, and brute-force branch cases.
"""
import math

import numpy as np
import pytest

import oracle
from grdzhadzha_amd.options import (AtmosphereOptions, CircularBinaryPNParams,
                                    KerrSchildParams, OperatorOptions,
                                    RecoveryOptions)

KS_POINTS = [(6.125, 3.875, 2.625), (3.0, 1.0, -2.0), (-4.2, 2.5, 0.7),
             (0.9, -0.4, 0.25), (10.0, -7.0, 5.5)]
PN = CircularBinaryPNParams(mass_1=0.55, mass_2=0.45, separation=9.0,
                            phase=0.2, time=3.0, softening_radius=0.7)


def adm(bg, x, t=0.0):
    return oracle.adm_point(bg, np.array(x, float), t)


# ------------------------------------------------------------ Kerr-Schild


@pytest.mark.parametrize("spin", [0.0, 0.3, 0.9])
@pytest.mark.parametrize("x", KS_POINTS)
def test_ks_inverse_identity(spin, x):
    """gamma_ij gamma^jk = delta (GRHDFixedBGTest.cpp:58-73, 1e-11)."""
    a = adm(KerrSchildParams(1.0, spin), x)
    assert np.max(np.abs(a["gamma"] @ a["gamma_UU"] - np.eye(3))) < 1e-11


@pytest.mark.parametrize("x", KS_POINTS)
def test_ks_schwarzschild_closed_forms(x):
    """a = 0: alpha = (1+2M/r)^-1/2, gamma = delta + 2M/r n n,
    beta^i = (2M/r)/(1+2M/r) n^i, K = 2 M alpha^3 (1 + 3M/r) / r^2."""
    M = 1.0
    a = adm(KerrSchildParams(M, 0.0), x)
    r = math.sqrt(sum(v * v for v in x))
    n = np.array(x) / r
    H = M / r
    assert abs(a["lapse"] - (1 + 2 * H) ** -0.5) < 1e-15
    assert np.max(np.abs(a["gamma"] - (np.eye(3) + 2 * H * np.outer(n, n)))) \
        < 1e-15
    assert np.max(np.abs(a["shift"] - 2 * H / (1 + 2 * H) * n)) < 1e-15
    K = 2 * M * a["lapse"] ** 3 * (1 + 3 * M / r) / r ** 2
    assert abs(a["K"] - K) <= 1e-14 * abs(K)


@pytest.mark.parametrize("bg", [KerrSchildParams(1.0, 0.3),
                                KerrSchildParams(1.0, 0.9), PN])
@pytest.mark.parametrize("x", KS_POINTS[:3])
def test_analytic_derivatives_match_central_differences(bg, x):
    """d_k gamma_ij, d_i alpha, d_j beta^i against O(h^2) central
    differences of the analytic fields."""
    h = 1e-4
    a0 = adm(bg, x)
    for k in range(3):
        xp = np.array(x, float)
        xm = np.array(x, float)
        xp[k] += h
        xm[k] -= h
        ap, am = adm(bg, xp), adm(bg, xm)
        dg = (ap["gamma"] - am["gamma"]) / (2 * h)
        da = (ap["lapse"] - am["lapse"]) / (2 * h)
        db = (ap["shift"] - am["shift"]) / (2 * h)
        assert np.max(np.abs(a0["d1_gamma"][:, :, k] - dg)) < 1e-7
        assert abs(a0["d1_lapse"][k] - da) < 1e-7
        assert np.max(np.abs(a0["d1_shift"][:, k] - db)) < 1e-7
        # d log sqrt det gamma = 1/2 gamma^ij d gamma_ij
        dl = (ap["log_sqrt_det"] - am["log_sqrt_det"]) / (2 * h)
        assert abs(a0["d_log_sqrt_det"][k] - dl) < 1e-7


def _adm_identity_K(bg, x, t, dt_step=1e-5):
    """K_ij = (D_i beta_j + D_j beta_i - d_t gamma_ij) / (2 alpha), with
    every derivative by central differences of the analytic fields
    (GRHDBinaryPNTest.cpp:155-198)."""
    h = 1e-4
    a0 = adm(bg, x, t)
    g = a0["gamma"]
    dbL = np.zeros((3, 3))       # d_i beta_j (lowered index)
    dg = np.zeros((3, 3, 3))     # [k][i][j] = d_k gamma_ij
    for k in range(3):
        xp = np.array(x, float)
        xm = np.array(x, float)
        xp[k] += h
        xm[k] -= h
        ap, am = adm(bg, xp, t), adm(bg, xm, t)
        dbL[k] = (ap["gamma"] @ ap["shift"] - am["gamma"] @ am["shift"]) / \
            (2 * h)
        dg[k] = (ap["gamma"] - am["gamma"]) / (2 * h)
    gu = np.linalg.inv(g)
    # Gamma^m_ij
    chris = 0.5 * np.einsum("ml,ijl->mij", gu,
                            np.einsum("ilj->ijl", dg) +
                            np.einsum("jli->ijl", dg) -
                            np.einsum("lij->ijl", dg))
    betaL = g @ a0["shift"]
    Di_bj = dbL - np.einsum("mij,m->ij", chris, betaL)
    ap, am = adm(bg, x, t + dt_step), adm(bg, x, t - dt_step)
    dtg = (ap["gamma"] - am["gamma"]) / (2 * dt_step)
    return a0["K_tensor"], (Di_bj + Di_bj.T - dtg) / (2 * a0["lapse"])


@pytest.mark.parametrize("spin", [0.3, 0.9])
def test_ks_adm_identity(spin):
    for x in KS_POINTS[:3]:
        K, expect = _adm_identity_K(KerrSchildParams(1.0, spin), x, 0.0)
        assert np.max(np.abs(K - expect)) < 1e-7


# ------------------------------------------------------------ binary PN


def test_pn_adm_identity_with_time_derivative():
    x = (60 * 0.25 + 0.125, 40 * 0.25 + 0.125, 24 * 0.25 + 0.125)
    for xx in (x, (2.0, -1.5, 0.5)):
        K, expect = _adm_identity_K(PN, xx, 0.0)
        assert np.max(np.abs(K)) > 1e-14
        assert np.max(np.abs(K - expect)) < 1e-8


def test_pn_closed_form_geometry():
    """Conformally flat: gamma = (1+2U) delta, gamma^-1 = delta/(1+2U),
    d ln sqrt gamma = 1.5 d(1+2U)/(1+2U) (AnalyticSpacetime.hpp:185-282),
    U and beta from two softened point masses on the circular orbit."""
    x = np.array([3.0, -2.0, 1.0])
    a = adm(PN, x)
    M = PN.mass_1 + PN.mass_2
    om = math.sqrt(M / PN.separation ** 3)
    ph = PN.phase + om * PN.time
    r1 = PN.mass_2 / M * PN.separation
    r2 = PN.mass_1 / M * PN.separation
    p1 = np.array([r1 * math.cos(ph), r1 * math.sin(ph), 0])
    p2 = -np.array([r2 * math.cos(ph), r2 * math.sin(ph), 0])
    v1 = om * np.array([-p1[1], p1[0], 0])
    v2 = om * np.array([-p2[1], p2[0], 0])
    U = 0.0
    V = np.zeros(3)
    dU = np.zeros(3)
    for m, p, v in ((PN.mass_1, p1, v1), (PN.mass_2, p2, v2)):
        d = x - p
        R = math.sqrt(PN.softening_radius ** 2 + d @ d)
        U += m / R
        V += m * v / R
        dU += -m * d / R ** 3
    cf = 1 + 2 * U
    assert np.max(np.abs(a["gamma"] - cf * np.eye(3))) < 1e-14
    assert np.max(np.abs(a["gamma_UU"] - np.eye(3) / cf)) < 1e-14
    assert abs(a["lapse"] - math.sqrt(1 - 2 * U + 2 * U * U +
                                      16 * (V @ V) / cf)) < 1e-14
    assert np.max(np.abs(a["shift"] + 4 * V / cf)) < 1e-14
    assert np.max(np.abs(a["d_log_sqrt_det"] - 1.5 * 2 * dU / cf)) < 1e-14


def test_log_sqrt_det_time_derivative():
    """Central difference of log sqrt det in time (:616-631) against the
    analytic d_t ln sqrt gamma = 3 dU/dt / (1 + 2U) of the binary."""
    dx, center, iv = 0.25, np.zeros(3), (21, 16, 11)
    t = 0.35
    got = oracle.log_sqrt_det_dt(PN, t, iv, dx, center, 1e-4)
    x = (np.array(iv) + 0.5) * dx
    h = 1e-6
    up = adm(PN, x, t + h)["gamma"][0, 0]
    um = adm(PN, x, t - h)["gamma"][0, 0]
    exact = 1.5 * (up - um) / (2 * h) / adm(PN, x, t)["gamma"][0, 0]
    assert abs(got) > 1e-14
    assert abs(got - exact) < 1e-7 * max(1.0, abs(exact))


# ------------------------------------------------------------ IEEE forms


def test_log_matches_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-40, 40, 20000)),
                         1 + rng.uniform(-1e-3, 1e-3, 2000), [1.0, 2.0, 0.5]])
    worst = 0.0
    for x in xs:
        ref = math.log(x)
        got = oracle.log(float(x))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / math.ulp(ref))
    assert worst <= 2.0, worst


def test_pow_rewrites_within_two_ulp():
    rng = np.random.default_rng(1)
    for x in map(float, rng.uniform(1.0, 40.0, 5000)):
        ref = math.pow(x, -0.5)
        assert abs(1.0 / math.sqrt(x) - ref) <= 2 * math.ulp(ref)
    for a in map(float, rng.uniform(0.05, 1.0, 5000)):
        ref = math.pow(a, 3.0)
        assert abs(a * a * a - ref) <= 2 * math.ulp(ref)
        assert abs(a * a - math.pow(a, 2.0)) <= math.ulp(math.pow(a, 2.0))


# ------------------------------------------------------------ Valencia


def _geom_at(bg, x):
    a = adm(bg, x)
    return a["gamma"], a["gamma_UU"], a["lapse"], a["shift"]


def test_con2prim_round_trip_reference_ks_point():
    """GRHDFixedBGTest.cpp:651-690: KS M = 1, a = 0.3, cell (24, 15, 10),
    dx 0.25; rho 1.1, eps 0.2, Gamma 5/3, v (0.01, -0.015, 0.005); 1e-10."""
    x = (24.5 * 0.25, 15.5 * 0.25, 10.5 * 0.25)
    g, gu, _, _ = _geom_at(KerrSchildParams(1.0, 0.3), x)
    gamma = 5.0 / 3.0
    prim = np.array([1.1, 0.2, (gamma - 1) * 1.1 * 0.2, 0.01, -0.015, 0.005])
    cons = oracle.prim2con(prim, gamma, g)
    ok, rec, iters, res = oracle.con2prim(
        cons, OperatorOptions(adiabatic_index=gamma), gu)
    assert ok and abs(res) < 1e-12
    scale = np.maximum(1.0, np.abs(prim))
    assert np.max(np.abs(rec - prim) / scale) < 1e-10


def test_con2prim_round_trip_reference_pn_point():
    """GRHDBinaryPNTest.cpp:225-255: cell (60, 40, 24), dx 0.25;
    rho 0.9, eps 0.15, v (0.004, -0.006, 0.002); 1e-10."""
    x = (60.5 * 0.25, 40.5 * 0.25, 24.5 * 0.25)
    g, gu, _, _ = _geom_at(PN, x)
    gamma = 5.0 / 3.0
    prim = np.array([0.9, 0.15, (gamma - 1) * 0.9 * 0.15, 0.004, -0.006,
                     0.002])
    cons = oracle.prim2con(prim, gamma, g)
    ok, rec, _, _ = oracle.con2prim(cons, OperatorOptions(
        adiabatic_index=gamma), gu)
    assert ok
    assert np.max(np.abs(rec - prim) / np.maximum(1, np.abs(prim))) < 1e-10


def test_prim2con_closed_form():
    """D = rho W, S_i = rho h W^2 v_i, tau = rho h W^2 - P - D."""
    g, _, _, _ = _geom_at(KerrSchildParams(1.0, 0.9), KS_POINTS[1])
    gamma = 4.0 / 3.0
    rho, eps = 0.3, 0.05
    p = (gamma - 1) * rho * eps
    v = np.array([0.1, -0.2, 0.05])
    c = oracle.prim2con(np.array([rho, eps, p, *v]), gamma, g)
    W = 1 / math.sqrt(1 - v @ g @ v)
    h = 1 + eps + p / rho
    assert abs(c[0] - rho * W) < 1e-15
    assert np.max(np.abs(c[1:4] - rho * h * W * W * (g @ v))) < 1e-15
    assert abs(c[4] - (rho * h * W * W - p - rho * W)) < 1e-15


def test_con2prim_forced_failures():
    opts = OperatorOptions(adiabatic_index=4.0 / 3.0)
    eye = np.eye(3)
    # D <= 0 (Valencia.hpp:123-124)
    ok, *_ = oracle.con2prim(np.array([-1e-3, 0, 0, 0, 1e-3]), opts, eye)
    assert not ok
    # f(p_floor) >= 0 beyond tolerance: tau far below -D
    ok, *_ = oracle.con2prim(np.array([1.0, 0, 0, 0, -0.9]), opts, eye)
    assert not ok
    # superluminal momentum: v^2 at the root >= max_velocity_squared
    ok, *_ = oracle.con2prim(np.array([1.0, 50.0, 0, 0, 1.0]), opts, eye)
    assert not ok
    # bracket exhaustion with max_iterations = 0
    o0 = OperatorOptions(adiabatic_index=4.0 / 3.0,
                         recovery=RecoveryOptions(max_iterations=0))
    ok, *_ = oracle.con2prim(np.array([1.0, 0.1, 0, 0, 5.0]), o0, eye)
    assert not ok


def test_floors_cases():
    atm = AtmosphereOptions(rho_floor=1e-10, pressure_floor=1e-12,
                            max_velocity_squared=0.8)
    opts = OperatorOptions(adiabatic_index=4.0 / 3.0, atmosphere=atm)
    eye = np.eye(3)
    p = oracle.floors(np.array([-1.0, 0.0, -1.0, 0.9, 0.3, 0.0]), opts, eye)
    assert p[0] == 1e-10 and p[2] == 1e-12
    assert abs(p[3] ** 2 + p[4] ** 2 - 0.8) < 1e-15
    assert abs(p[3] / p[4] - 3.0) < 1e-14
    assert p[1] == 1e-12 / ((4.0 / 3.0 - 1.0) * 1e-10)
    q = oracle.floors(np.array([1.0, 0.1, 0.2, 0.1, 0.0, 0.0]), opts, eye)
    assert q[3] == 0.1 and q[0] == 1.0


# ------------------------------------------------------------ reconstruction


def test_minmod_and_limiter():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        l, c, r = rng.normal(size=3)
        # theta = 1 is exactly minmod(c - l, r - c)
        assert oracle.limited_slope(l, c, r, 1.0) == \
            oracle.minmod(c - l, r - c)
        # theta = 2 (MC): minmod(2(c-l), (r-l)/2, 2(r-c)) by brute force
        cand = [2 * (c - l), 0.5 * (r - l), 2 * (r - c)]
        if all(v > 0 for v in cand):
            mc = min(cand)
        elif all(v < 0 for v in cand):
            mc = max(cand)
        else:
            mc = 0.0
        assert abs(oracle.limited_slope(l, c, r, 2.0) - mc) <= 1e-15 * \
            max(1.0, abs(mc))
    assert oracle.minmod(1.0, -1.0) == 0.0
    assert oracle.minmod(0.0, 3.0) == 0.0
    assert oracle.minmod(-2.0, -3.0) == -2.0


# ------------------------------------------------------------ HLLE


def _hlle_case(pl, pr, d=0, bg=KerrSchildParams(1.0, 0.0),
               x=(8.0, 1.0, 0.5)):
    g, gu, lapse, shift = _geom_at(bg, x)
    f, s = oracle.hlle(pl, pr, 5.0 / 3.0, g, gu, lapse, shift, d)
    fl = oracle.flux(pl, 5.0 / 3.0, g, lapse, shift, d)
    fr = oracle.flux(pr, 5.0 / 3.0, g, lapse, shift, d)
    return f, s, fl, fr, (g, gu, lapse, shift)


def test_hlle_consistency():
    """HLLE(U, U) = F(U)."""
    p = np.array([1.0, 0.2, 0.1333, 0.1, -0.05, 0.02])
    for d in range(3):
        f, s, fl, _, _ = _hlle_case(p, p, d)
        assert np.max(np.abs(f - fl)) <= 1e-15 * np.max(np.abs(fl)) * 10


def test_hlle_branches():
    cold_l = np.array([1.0, 1e-4, 6.7e-5, 0.9, 0.0, 0.0])
    cold_r = np.array([0.5, 1e-4, 3.3e-5, 0.9, 0.0, 0.0])
    # everything moves right: s_lower >= 0 -> left flux
    f, s, fl, fr, _ = _hlle_case(cold_l, cold_r, 0, KerrSchildParams(1e-9, 0),
                                 (100.0, 0.0, 0.0))
    assert min(s[0], s[2]) > 0
    assert np.array_equal(f, fl)
    # everything moves left: s_upper <= 0 -> right flux
    cl = cold_l.copy()
    cr = cold_r.copy()
    cl[3] = cr[3] = -0.9
    f, s, fl, fr, _ = _hlle_case(cl, cr, 0, KerrSchildParams(1e-9, 0),
                                 (100.0, 0.0, 0.0))
    assert max(s[1], s[3]) < 0
    assert np.array_equal(f, fr)
    # subsonic: the HLLE average
    pl = np.array([1.0, 0.3, 0.2, 0.05, 0.0, 0.0])
    pr = np.array([0.8, 0.25, 0.13, -0.02, 0.01, 0.0])
    f, s, fl, fr, geo = _hlle_case(pl, pr, 0)
    sl = min(0.0, s[0], s[2])
    su = max(0.0, s[1], s[3])
    ul = oracle.prim2con(pl, 5.0 / 3.0, geo[0])
    ur = oracle.prim2con(pr, 5.0 / 3.0, geo[0])
    expect = (su * fl - sl * fr + sl * su * (ur - ul)) / (su - sl)
    assert np.max(np.abs(f - expect)) < 1e-14


def test_characteristic_speeds_flat_closed_form():
    """Flat metric, v along x: lambda = (v +- c_s) / (1 +- v c_s)."""
    gamma = 5.0 / 3.0
    rho, eps = 1.0, 0.3
    p = (gamma - 1) * rho * eps
    v = 0.4
    h = 1 + eps + p / rho
    cs = math.sqrt(gamma * p / (rho * h))
    eye = np.eye(3)
    _, s = oracle.hlle(np.array([rho, eps, p, v, 0, 0]),
                       np.array([rho, eps, p, v, 0, 0]), gamma, eye, eye, 1.0,
                       np.zeros(3), 0)
    assert abs(s[0] - (v - cs) / (1 - v * cs)) < 1e-14
    assert abs(s[1] - (v + cs) / (1 + v * cs)) < 1e-14


# ------------------------------------------------------------ sources


@pytest.mark.parametrize("bg", [KerrSchildParams(1.0, 0.3), PN])
def test_source_K_contribution_is_tau_only(bg):
    """GRHDFixedBGTest.cpp:704-729 / GRHDBinaryPNTest.cpp:262-287: K_ij only
    changes the tau source, by alpha S^ij K_ij (1e-12); D and S_i are
    unchanged (1e-14)."""
    x = (6.125, 3.875, 2.625)
    gamma = 5.0 / 3.0
    prim = np.array([1.1, 0.2, (gamma - 1) * 0.22, 0.01, -0.015, 0.005])
    g = adm(bg, x)["gamma"]
    cons = oracle.prim2con(prim, gamma, g)
    s = oracle.sources_point(bg, x, prim, cons, gamma, use_K=True)
    s0 = oracle.sources_point(bg, x, prim, cons, gamma, use_K=False)
    assert abs((s[4] - s0[4]) - s[5]) <= 1e-12 * max(1.0, abs(s[5]))
    assert abs(s[5]) > 1e-14
    assert np.max(np.abs(s[:4] - s0[:4])) <= 1e-14


def test_sources_vanish_in_flat_space():
    """With M -> 0 every geometric source vanishes."""
    gamma = 4.0 / 3.0
    prim = np.array([0.5, 0.1, 0.2 / 3, 0.1, 0.2, -0.1])
    cons = oracle.prim2con(prim, gamma, np.eye(3))
    s = oracle.sources_point(KerrSchildParams(0.0, 0.0), (3.0, 1.0, 2.0),
                             prim, cons, gamma)
    assert np.max(np.abs(s[:5])) == 0.0


def test_static_source_momentum_matches_stress_form():
    """S_{S_k} = -(tau + D) d_k alpha + 1/2 alpha S^ij d_k gamma_ij
    + S_j d_k beta^j - sum_d F^d_{S_k} d_d ln sqrt gamma, evaluated
    independently with numpy from the ADM fields."""
    bg = KerrSchildParams(1.0, 0.9)
    x = (3.0, 1.0, -2.0)
    gamma = 4.0 / 3.0
    a = adm(bg, x)
    g, gu, al, be = a["gamma"], a["gamma_UU"], a["lapse"], a["shift"]
    v = np.array([0.05, -0.1, 0.02])
    rho, eps = 0.7, 0.1
    p = (gamma - 1) * rho * eps
    prim = np.array([rho, eps, p, *v])
    cons = oracle.prim2con(prim, gamma, g)
    W = 1 / math.sqrt(1 - v @ g @ v)
    h = 1 + eps + p / rho
    T = rho * h * W * W * np.outer(v, v) + p * gu
    D, S, tau = cons[0], cons[1:4], cons[4]
    F = np.zeros((3, 5))
    for d in range(3):
        tv = al * v[d] - be[d]
        F[d, 0] = D * tv
        F[d, 1:4] = S * tv
        F[d, 1 + d] += al * p
        F[d, 4] = tau * tv + al * p * v[d]
    dl = a["d_log_sqrt_det"]
    Sk = np.array([-(tau + D) * a["d1_lapse"][k] +
                   0.5 * al * np.sum(T * a["d1_gamma"][:, :, k]) +
                   S @ a["d1_shift"][:, k] - F[:, 1 + k] @ dl
                   for k in range(3)])
    Stau = -(F[:, 4] @ dl) - (gu @ S) @ a["d1_lapse"] + \
        al * np.sum(T * a["K_tensor"])
    SD = -(F[:, 0] @ dl)
    s = oracle.sources_point(bg, x, prim, cons, gamma)
    assert abs(s[0] - SD) < 1e-14
    assert np.max(np.abs(s[1:4] - Sk)) < 1e-13
    assert abs(s[4] - Stau) < 1e-13
