"""Oracle pins for the level diagnostics (LevelDataDriver.hpp:22-92):
the sequential sum equals a brute-force Python loop in box / BoxIterator
order bit for bit, lies within the sequential-rounding bound of the exact
sum (math.fsum), and finiteness looks at interior cells only."""
import math

import numpy as np

import oracle
from grdzhadzha_amd.options import BoxSpec


def tiny_level(seed=0):
    rng = np.random.default_rng(seed)
    boxes = [BoxSpec((0, 0, 0), (3, 2, 4), 2, 0.5, (0.0, 0.0, 0.0)),
             BoxSpec((4, 0, 0), (6, 2, 4), 2, 0.5, (0.0, 0.0, 0.0)),
             BoxSpec((0, 3, 0), (6, 5, 1), 3, 0.5, (0.0, 0.0, 0.0))]
    fabs = []
    for b in boxes:
        f = rng.standard_normal((11,) + b.storage_shape) * \
            10.0 ** rng.integers(-8, 8, size=(11,) + b.storage_shape)
        fabs.append(f)
    return boxes, fabs


def brute_sum(boxes, fabs):
    s = [0.0] * 5
    for b, f in zip(boxes, fabs):
        for k in range(b.lo[2], b.hi[2] + 1):
            for j in range(b.lo[1], b.hi[1] + 1):
                for i in range(b.lo[0], b.hi[0] + 1):
                    for m in range(5):
                        s[m] += float(f[m, k - b.glo[2], j - b.glo[1],
                                        i - b.glo[0]])
    return tuple(s)


def interior_values(boxes, fabs, m):
    return np.concatenate([f[m][b.interior_slices()].ravel()
                           for b, f in zip(boxes, fabs)])


def test_sum_is_the_reference_order_sequential_sum():
    boxes, fabs = tiny_level()
    assert oracle.leveldata_conserved_sum(boxes, fabs) == \
        brute_sum(boxes, fabs)


def test_sum_within_sequential_rounding_bound_of_exact():
    boxes, fabs = tiny_level(1)
    got = oracle.leveldata_conserved_sum(boxes, fabs)
    for m in range(5):
        v = interior_values(boxes, fabs, m)
        exact = math.fsum(v)
        bound = (len(v) - 1) * np.finfo(float).eps / 2 * np.abs(v).sum()
        assert abs(got[m] - exact) <= bound


def test_sum_uses_comp_base_and_ignores_ghosts():
    boxes, fabs = tiny_level(2)
    shifted = [np.concatenate([np.full((3,) + f.shape[1:], np.nan), f])
               for f in fabs]
    for b, f in zip(boxes, shifted):
        g = np.ones(f.shape[1:], bool)
        g[b.interior_slices()] = False
        f[:, g] = np.inf  # ghosts do not count
    assert oracle.leveldata_conserved_sum(boxes, shifted, comp_base=3) == \
        oracle.leveldata_conserved_sum(boxes, fabs)
    assert oracle.leveldata_conserved_state_is_finite(boxes, shifted,
                                                      comp_base=3)


def test_finiteness():
    boxes, fabs = tiny_level(3)
    assert oracle.leveldata_conserved_state_is_finite(boxes, fabs)
    for m, bad in ((0, np.nan), (3, np.inf), (4, -np.inf)):
        f = [x.copy() for x in fabs]
        b = boxes[1]
        f[1][m, b.lo[2] - b.glo[2] + 1, b.lo[1] - b.glo[1],
             b.hi[0] - b.glo[0]] = bad
        assert not oracle.leveldata_conserved_state_is_finite(boxes, f)
    f = [x.copy() for x in fabs]
    f[0][5:, ...] = np.nan  # primitives are not part of the check
    assert oracle.leveldata_conserved_state_is_finite(boxes, f)
