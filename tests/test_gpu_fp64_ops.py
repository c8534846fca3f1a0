"""The kernels' FP64 division and square root against IEEE arithmetic.

Every FP64 kernel divides with qdiv and takes square roots with qsqrt: the
hardware's refined reciprocal and reciprocal square root without the range
scaling of the compiler's general sequences (grhd_physics.hpp).  The
bit-exact parity of the whole path with the CPU oracle rests on these two
being the IEEE operations over the values the path meets.  Here they run
through the C-ABI (grhd_fp64_div_sqrt) on random operands spread over the
stated ranges -- 2^+-380 for the quotients, 2^-760..2^760 for the roots --
on operands of the path's own magnitudes, and on zeros, infinities, NaN and
negative numbers, and must equal numpy's IEEE results bit for bit (NaN
results only need to be NaN)."""
import ctypes

import numpy as np
import pytest
import torch

from grdzhadzha_amd import abi

pytestmark = pytest.mark.gpu


def run(a, b):
    lib = abi.load_library()
    da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    q = torch.empty_like(da)
    r = torch.empty_like(da)
    torch.cuda.synchronize()
    rc = lib.grhd_fp64_div_sqrt(ctypes.c_void_p(da.data_ptr()),
                                ctypes.c_void_p(db.data_ptr()),
                                ctypes.c_void_p(q.data_ptr()),
                                ctypes.c_void_p(r.data_ptr()),
                                ctypes.c_longlong(a.size), None)
    assert rc == 0
    return q.cpu().numpy(), r.cpu().numpy()


def assert_ieee(got, want, what):
    both_nan = np.isnan(got) & np.isnan(want)
    same = (got.view(np.uint64) == want.view(np.uint64)) | both_nan
    bad = np.flatnonzero(~same)
    assert bad.size == 0, (
        f"{what}: {bad.size} of {got.size} differ from IEEE, e.g. index "
        f"{bad[0]}: got {got[bad[0]]!r}, IEEE {want[bad[0]]!r}")


def random_doubles(rng, n, emin, emax):
    """Uniform 52-bit mantissas, uniform exponents in [emin, emax], random
    signs."""
    m = rng.random(n) + 1.0
    e = rng.integers(emin, emax + 1, n)
    s = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    return s * np.ldexp(m, e)


def test_division_and_root_equal_ieee_over_the_stated_ranges():
    rng = np.random.default_rng(20261204)
    n = 1 << 21
    a = random_doubles(rng, n, -380, 380)
    b = random_doubles(rng, n, -380, 380)
    q, r = run(a, b)
    with np.errstate(invalid="ignore"):
        assert_ieee(q, a / b, "a / b")
        root_arg = np.abs(random_doubles(rng, n, -760, 760))
        _, r = run(root_arg, b)
        assert_ieee(r, np.sqrt(root_arg), "sqrt")


def test_division_and_root_equal_ieee_at_the_paths_magnitudes():
    """Operands like the kernels' (floors 1e-13, metric components near 1,
    densities, pressures and momenta from 1e-14 to 1e4), where quotients of
    close numbers (1 / (1 - v^2), residual / derivative) are common."""
    rng = np.random.default_rng(7)
    n = 1 << 20
    a = random_doubles(rng, n, -50, 20)
    b = random_doubles(rng, n, -50, 20)
    # denominators and numerators close to each other and to 1
    k = n // 4
    b[:k] = a[:k] * (1.0 + rng.uniform(-1e-3, 1e-3, k))
    a[k:2 * k] = 1.0
    b[k:2 * k] = 1.0 - np.abs(rng.uniform(0.0, 0.95, k))
    q, r = run(a, b)
    with np.errstate(invalid="ignore"):
        assert_ieee(q, a / b, "a / b")
        assert_ieee(r, np.sqrt(a), "sqrt")


def div_in_range(a, b):
    """Operand pairs whose quotient qdiv promises to be IEEE's: a zero,
    infinite or NaN operand, or |a| >= 2^-969, b normal below 2^1022,
    exponent(a) - exponent(b) < 768 and a finite normal quotient."""
    special = ~np.isfinite(a) | ~np.isfinite(b) | (a == 0.0) | (b == 0.0)
    with np.errstate(all="ignore"):
        ea = np.frexp(a)[1]
        eb = np.frexp(b)[1]
        quo = np.abs(a / b)
    ok = ((np.abs(a) >= 2.0 ** -969) & (np.abs(b) >= 2.0 ** -1022) &
          (np.abs(b) < 2.0 ** 1022) & (ea - eb < 768) &
          (quo >= 2.0 ** -1022) & np.isfinite(quo))
    return special | ok


def test_special_operands():
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0,
                         2.5e-13, 1e300, -3.0e-200])
    a = np.repeat(specials, specials.size)
    b = np.tile(specials, specials.size)
    q, r = run(a, b)
    with np.errstate(all="ignore"):
        want_q = a / b
        want_r = np.sqrt(a)
    keep = div_in_range(a, b)
    assert keep.sum() >= 80
    assert_ieee(q[keep], want_q[keep], "a / b (special operands)")
    assert_ieee(r, want_r, "sqrt (special operands)")
    # the IEEE special results themselves
    assert np.isinf(q[(a == 1.0) & (b == 0.0)]).all()
    assert np.isnan(q[(a == 0.0) & (b == 0.0)]).all()
    assert np.isnan(r[a < 0.0]).all()
    assert (np.signbit(r[a == 0.0]) == np.signbit(a[a == 0.0])).all()
