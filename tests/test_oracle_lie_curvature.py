"""The mixed-precision sources form K_ij as the Lie derivative of the
spatial metric along the shift,

    K_ij = (beta^l d_l gamma_ij + gamma_lj d_i beta^l + gamma_il d_j beta^l
            - d_t gamma_ij) / (2 alpha),

instead of the reference's Christoffel form (KerrSchild.hpp:146-164,
CircularBinaryPN.hpp).  These tests pin that identity and its index
conventions on the oracle's analytic ADM variables: the Lie form built from
the oracle's gamma, shift, lapse and first derivatives equals the oracle's
K_ij (Christoffel form) to rounding, for Kerr-Schild at three spins and for
the circular binary, whose d_t gamma_ij = 2 dU/dt delta_ij comes from the
closed-form potential."""
import math

import numpy as np
import pytest

import oracle
from grdzhadzha_amd.options import CircularBinaryPNParams, KerrSchildParams

POINTS = [(6.125, 3.875, 2.625), (3.0, 1.0, -2.0), (-4.2, 2.5, 0.7),
          (0.9, -0.4, 0.25), (10.0, -7.0, 5.5)]
PN = CircularBinaryPNParams(mass_1=0.55, mass_2=0.45, separation=9.0,
                            phase=0.2, time=3.0, softening_radius=0.7)


def lie_K(a, dt_gamma):
    g = a["gamma"]
    beta = a["shift"]            # beta^i
    dg = a["d1_gamma"]           # [i][j][k] = d_k gamma_ij
    db = a["d1_shift"]           # [i][j] = d_j beta^i
    K = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            v = -dt_gamma[i, j]
            for l in range(3):
                v += (beta[l] * dg[i][j][l] + g[l][j] * db[l][i] +
                      g[i][l] * db[l][j])
            K[i, j] = v / (2.0 * a["lapse"])
    return K


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.9])
@pytest.mark.parametrize("x", POINTS)
def test_lie_form_equals_christoffel_form_kerr_schild(spin, x):
    a = oracle.adm_point(KerrSchildParams(1.0, spin), np.array(x, float))
    K = a["K_tensor"]
    L = lie_K(a, np.zeros((3, 3)))
    assert np.max(np.abs(K)) > 1e-3
    assert np.max(np.abs(L - K)) <= 1e-13 * np.max(np.abs(K))
    # symmetric by construction, the Christoffel form only to rounding
    assert np.array_equal(L, L.T) or np.max(np.abs(L - L.T)) < 1e-15


def test_lie_form_equals_christoffel_form_binary():
    M = PN.mass_1 + PN.mass_2
    om = math.sqrt(M / PN.separation ** 3)
    ph = PN.phase + om * PN.time
    r1 = PN.mass_2 / M * PN.separation
    r2 = PN.mass_1 / M * PN.separation
    p1 = np.array([r1 * math.cos(ph), r1 * math.sin(ph), 0.0])
    p2 = -np.array([r2 * math.cos(ph), r2 * math.sin(ph), 0.0])
    v1 = om * np.array([-p1[1], p1[0], 0.0])
    v2 = om * np.array([-p2[1], p2[0], 0.0])
    for x in [(3.0, -2.0, 1.0), (2.0, -1.5, 0.5), (15.125, 10.125, 6.125)]:
        x = np.array(x)
        a = oracle.adm_point(PN, x)
        # d_t U of U = sum m / sqrt(s^2 + |x - p|^2) with dp/dt = v
        dUdt = 0.0
        for m, p, v in ((PN.mass_1, p1, v1), (PN.mass_2, p2, v2)):
            d = x - p
            R = math.sqrt(PN.softening_radius ** 2 + d @ d)
            dUdt += m * (d @ v) / R ** 3
        K = a["K_tensor"]
        L = lie_K(a, 2.0 * dUdt * np.eye(3))
        assert np.max(np.abs(K)) > 1e-6
        assert np.max(np.abs(L - K)) <= 1e-12 * np.max(np.abs(K))


def test_lie_form_catches_a_transposed_shift_derivative():
    """The index convention matters: using d_i beta^l as d_l beta^i (the
    transpose of the oracle's d1_shift) breaks the identity."""
    a = oracle.adm_point(KerrSchildParams(1.0, 0.9), np.array(POINTS[1]))
    wrong = dict(a)
    wrong["d1_shift"] = a["d1_shift"].T
    K = a["K_tensor"]
    assert np.max(np.abs(lie_K(wrong, np.zeros((3, 3))) - K)) > \
        1e-3 * np.max(np.abs(K))
