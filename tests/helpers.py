"""Shared comparison helpers for the parity tests."""
import numpy as np

CONS = ("D", "S1", "S2", "S3", "tau")


def rel_err(a, b):
    """max |a - b| / max |b| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    diff = np.max(np.abs(a - b)) if a.size else 0.0
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)


def conserved_errors(a, b, comp_base=0):
    return {name: rel_err(a[comp_base + c], b[comp_base + c])
            for c, name in enumerate(CONS)}


def assert_conserved_close(a, b, tol=1e-12, comp_base=0, what=""):
    errs = conserved_errors(a, b, comp_base)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, "%s: relative errors above %g: %s" % (what, tol, errs)
    return errs


def assert_bitwise(got, ref, what=""):
    """got == ref bit for bit (NaNs at the same places count as equal).
    On a mismatch the message gives the differing cells, the largest
    difference in units of the last place, and the max-normalised error
    per component."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    if np.array_equal(got, ref, equal_nan=True):
        return
    diff = ~((got == ref) | (np.isnan(got) & np.isnan(ref)))
    g = got[diff].astype(np.float64)
    r = ref[diff].astype(np.float64)
    ulps = np.abs(g.view(np.int64) - r.view(np.int64))
    per = []
    if got.ndim >= 1:
        for c in range(got.shape[0]):
            per.append(rel_err(got[c], ref[c]))
    raise AssertionError(
        "%s: %d of %d values differ; max %d ulp; max-normalised error per "
        "component %s" % (what, int(diff.sum()), diff.size,
                          int(ulps.max()), per))
