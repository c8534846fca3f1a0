"""Python view of the CPU oracle (oracle/libgrhd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, to check and time against; never by the
product.
Arrays are numpy float64 in FArrayBox layout (ncomp, nz, ny, nx).
"""
import ctypes
import os
import subprocess

import numpy as np

from grdzhadzha_amd import abi
from grdzhadzha_amd.options import OperatorOptions

HERE = os.path.dirname(os.path.abspath(__file__))
# GRHD_ORACLE_LIB names another build of the same source (the sanitizer
# build, oracle/_san/libgrhd_oracle.so, made by `make sanitize`)
LIB = os.environ.get("GRHD_ORACLE_LIB") or os.path.join(HERE,
                                                        "libgrhd_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        B, Pa, Bg = P(abi.GrhdBox), P(abi.GrhdParams), P(abi.GrhdBackground)
        E, Dg = P(abi.GrhdError), P(abi.GrhdDiagnostics)
        sig = {
            "grhd_oracle_compute_rhs": (i, [B, Pa, Bg, d, vp, i, i, vp, vp, E]),
            "grhd_oracle_recover_primitives": (i, [B, Pa, Bg, d, vp, i, i, vp,
                                                   Dg]),
            "grhd_oracle_update_conserved": (i, [B, Pa, Bg, vp, i, i, vp, d, i,
                                                 d, vp, Dg]),
            "grhd_oracle_compute_max_inverse_dt": (i, [B, Pa, Bg, d, vp, i, i,
                                                       P(d), E]),
            "grhd_oracle_compute_stable_dt": (i, [B, Pa, Bg, d, vp, i, i, d,
                                                  P(d), E]),
            "grhd_oracle_advance_ssprk2": (i, [B, Pa, Bg, d, d, vp, i, i, vp,
                                               Dg, E]),
            "grhd_oracle_advance_rk4": (i, [B, Pa, Bg, d, d, i, vp, i, i, vp,
                                            Dg, E]),
            "grhd_oracle_adm_point": (i, [Bg, d, vp, vp]),
            "grhd_oracle_cell_geometry_box": (i, [B, Bg, d, vp]),
            "grhd_oracle_prim2con": (None, [vp, d, vp, vp]),
            "grhd_oracle_con2prim": (i, [vp, Pa, vp, vp, vp]),
            "grhd_oracle_floors": (None, [vp, Pa, vp]),
            "grhd_oracle_limited_slope": (d, [d, d, d, d]),
            "grhd_oracle_minmod": (d, [d, d]),
            "grhd_oracle_hlle": (None, [vp, vp, d, vp, vp, d, vp, i, vp, vp]),
            "grhd_oracle_flux": (None, [vp, d, vp, d, vp, i, vp]),
            "grhd_oracle_sources_point": (None, [Bg, d, vp, vp, vp, d, i, vp]),
            "grhd_oracle_log_sqrt_det_dt": (d, [Bg, d, vp, d, vp, d]),
            "grhd_oracle_log": (d, [d]),
            "grhd_oracle_num_threads": (i, []),
            "grhd_oracle_set_num_threads": (None, [i]),
            "grhd_oracle_ks_excised": (i, [d, d, vp, d, d]),
            "grhd_oracle_leveldata_conserved_sum": (None, [B, i, P(vp), i, i,
                                                           P(d)]),
            "grhd_oracle_leveldata_conserved_state_is_finite": (
                i, [B, i, P(vp), i, i]),
            "grhd_oracle_checkpoint_pack": (None, [B, i, P(vp), i, P(i), P(i),
                                                   i, vp]),
            "grhd_oracle_checkpoint_unpack": (None, [B, i, P(vp), i, P(i),
                                                     P(i), i, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    def __init__(self, code, err):
        self.code = code
        self.iv = tuple(err.iv) if err is not None else None
        self.residual = err.residual if err is not None else None
        super().__init__("oracle error %d at %s residual %s"
                         % (code, self.iv, self.residual))


class Oracle:
    """One box, one background, one set of operator options."""

    def __init__(self, background, box, options: OperatorOptions = None,
                 comp_base=0, time_dependent=False):
        self.L = lib()
        self.box = box
        self.options = options or OperatorOptions()
        if not time_dependent:
            o = dict(self.options.__dict__)
            o["use_metric_volume_time_sources"] = False
            self.options = OperatorOptions(**o)
        self.comp_base = comp_base
        self._b = box.to_c()
        self._p = self.options.to_c()
        self._g = background.to_c()

    def _h(self):
        return (ctypes.byref(self._b), ctypes.byref(self._p),
                ctypes.byref(self._g))

    @staticmethod
    def _check(rc, err=None):
        if rc != 0:
            raise OracleError(rc, err)

    def compute_rhs(self, state, time=0.0, with_mask=False):
        ncomp = state.shape[0]
        rhs = np.zeros((ncomp,) + self.box.interior_shape)
        mask = np.zeros(self.box.interior_shape, np.uint8) if with_mask \
            else None
        err = abi.GrhdError()
        rc = self.L.grhd_oracle_compute_rhs(*self._h(), time, _p(state), ncomp,
                                            self.comp_base, _p(rhs), _p(mask),
                                            ctypes.byref(err))
        self._check(rc, err)
        return (rhs, mask) if with_mask else rhs

    def recover_primitives(self, state, time=0.0):
        mask = np.zeros(self.box.interior_shape, np.uint8)
        diag = abi.GrhdDiagnostics()
        rc = self.L.grhd_oracle_recover_primitives(
            *self._h(), time, _p(state), state.shape[0], self.comp_base,
            _p(mask), ctypes.byref(diag))
        self._check(rc)
        return diag.as_dict(), mask

    def update_conserved(self, state, rhs, dt, recover=True,
                         recovery_time=0.0):
        mask = np.zeros(self.box.interior_shape, np.uint8)
        diag = abi.GrhdDiagnostics()
        rc = self.L.grhd_oracle_update_conserved(
            *self._h(), _p(state), state.shape[0], self.comp_base, _p(rhs),
            dt, int(recover), recovery_time, _p(mask), ctypes.byref(diag))
        self._check(rc)
        return diag.as_dict(), mask

    def compute_max_inverse_dt(self, state, time=0.0):
        out = ctypes.c_double()
        err = abi.GrhdError()
        rc = self.L.grhd_oracle_compute_max_inverse_dt(
            *self._h(), time, _p(state), state.shape[0], self.comp_base,
            ctypes.byref(out), ctypes.byref(err))
        self._check(rc, err)
        return out.value

    def compute_stable_dt(self, state, cfl, time=0.0):
        out = ctypes.c_double()
        err = abi.GrhdError()
        rc = self.L.grhd_oracle_compute_stable_dt(
            *self._h(), time, _p(state), state.shape[0], self.comp_base, cfl,
            ctypes.byref(out), ctypes.byref(err))
        self._check(rc, err)
        return out.value

    def advance_ssprk2(self, state, dt, time=0.0):
        mask = np.zeros(self.box.interior_shape, np.uint8)
        diag = abi.GrhdDiagnostics()
        err = abi.GrhdError()
        rc = self.L.grhd_oracle_advance_ssprk2(
            *self._h(), time, dt, _p(state), state.shape[0], self.comp_base,
            _p(mask), ctypes.byref(diag), ctypes.byref(err))
        self._check(rc, err)
        return diag.as_dict(), mask

    def advance_rk4(self, state, dt, nsteps=1, time=0.0):
        mask = np.zeros(self.box.interior_shape, np.uint8)
        diag = abi.GrhdDiagnostics()
        err = abi.GrhdError()
        rc = self.L.grhd_oracle_advance_rk4(
            *self._h(), time, dt, nsteps, _p(state), state.shape[0],
            self.comp_base, _p(mask), ctypes.byref(diag), ctypes.byref(err))
        self._check(rc, err)
        return diag.as_dict(), mask

    def cell_geometry(self, time=0.0):
        out = np.zeros((16,) + self.box.storage_shape)
        self.L.grhd_oracle_cell_geometry_box(ctypes.byref(self._b),
                                             ctypes.byref(self._g), time,
                                             _p(out))
        return out


def adm_point(background, x, time=0.0):
    """Full ADM vars at one point, unpacked into a dict."""
    out = np.zeros(80)
    xa = np.ascontiguousarray(x, dtype=np.float64)
    n = lib().grhd_oracle_adm_point(ctypes.byref(background.to_c()), time,
                                    _p(xa), _p(out))
    o = out[:n]
    res = {}
    k = 0

    def take(name, shape):
        nonlocal k
        size = int(np.prod(shape)) if shape else 1
        res[name] = o[k:k + size].reshape(shape) if shape else o[k]
        k += size

    take("gamma", (3, 3))
    take("K_tensor", (3, 3))
    take("K", ())
    take("lapse", ())
    take("shift", (3,))
    take("d1_gamma", (3, 3, 3))
    take("d1_lapse", (3,))
    take("d1_shift", (3, 3))
    take("gamma_UU", (3, 3))
    take("d_log_sqrt_det", (3,))
    take("log_sqrt_det", ())
    take("det", ())
    return res


def prim2con(prim, gamma, g):
    cons = np.zeros(5)
    lib().grhd_oracle_prim2con(_p(np.ascontiguousarray(prim, float)), gamma,
                               _p(np.ascontiguousarray(g, float)), _p(cons))
    return cons


def con2prim(cons, options: OperatorOptions, gu):
    prim = np.zeros(6)
    info = np.zeros(2)
    p = options.to_c()
    ok = lib().grhd_oracle_con2prim(_p(np.ascontiguousarray(cons, float)),
                                    ctypes.byref(p),
                                    _p(np.ascontiguousarray(gu, float)),
                                    _p(prim), _p(info))
    return bool(ok), prim, int(info[0]), info[1]


def floors(prim, options: OperatorOptions, g):
    q = np.ascontiguousarray(prim, float).copy()
    p = options.to_c()
    lib().grhd_oracle_floors(_p(q), ctypes.byref(p),
                             _p(np.ascontiguousarray(g, float)))
    return q


def hlle(pl, pr, gamma, g, gu, lapse, shift, d):
    f = np.zeros(5)
    s = np.zeros(4)
    c = lambda a: _p(np.ascontiguousarray(a, float))  # noqa: E731
    lib().grhd_oracle_hlle(c(pl), c(pr), gamma, c(g), c(gu), lapse, c(shift),
                           d, _p(f), _p(s))
    return f, s


def flux(p6, gamma, g, lapse, shift, d):
    f = np.zeros(5)
    c = lambda a: _p(np.ascontiguousarray(a, float))  # noqa: E731
    lib().grhd_oracle_flux(c(p6), gamma, c(g), lapse, c(shift), d, _p(f))
    return f


def sources_point(background, x, prim, cons, gamma, use_K=True, time=0.0):
    out = np.zeros(6)
    c = lambda a: _p(np.ascontiguousarray(a, float))  # noqa: E731
    lib().grhd_oracle_sources_point(ctypes.byref(background.to_c()), time,
                                    c(x), c(prim), c(cons), gamma, int(use_K),
                                    _p(out))
    return out


def log_sqrt_det_dt(background, time, iv, dx, center, step):
    iva = np.ascontiguousarray(iv, dtype=np.int32)
    return lib().grhd_oracle_log_sqrt_det_dt(
        ctypes.byref(background.to_c()), time, _p(iva), dx,
        _p(np.ascontiguousarray(center, float)), step)


def log(x):
    return lib().grhd_oracle_log(x)


def limited_slope(l, c, r, theta):
    return lib().grhd_oracle_limited_slope(l, c, r, theta)


def minmod(a, b):
    return lib().grhd_oracle_minmod(a, b)


def num_threads():
    return lib().grhd_oracle_num_threads()


def set_num_threads(n):
    """OpenMP threads of the oracle's parallel loops (omp_set_num_threads)."""
    lib().grhd_oracle_set_num_threads(int(n))


def ks_excised(M, a, x, inner_buffer=1.05, outer_buffer=0.9):
    """KerrSchild::check_if_excised (KerrSchild.hpp:247-280)."""
    xs = (ctypes.c_double * 3)(*[float(c) for c in x])
    return bool(lib().grhd_oracle_ks_excised(M, a, xs, inner_buffer,
                                             outer_buffer))


def _level_args(boxes, fabs):
    cb = (abi.GrhdBox * len(boxes))(*[b.to_c() for b in boxes])
    arrs = [np.ascontiguousarray(f, dtype=np.float64) for f in fabs]
    ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return cb, arrs, ptrs


def leveldata_conserved_sum(boxes, fabs, comp_base=0):
    """Sequential level sums of D, S_1, S_2, S_3, tau (reference order)."""
    cb, arrs, ptrs = _level_args(boxes, fabs)
    out = (ctypes.c_double * 5)()
    ncomp = arrs[0].shape[0]
    lib().grhd_oracle_leveldata_conserved_sum(cb, len(boxes), ptrs, ncomp,
                                              comp_base, out)
    return tuple(out)


def leveldata_conserved_state_is_finite(boxes, fabs, comp_base=0):
    cb, arrs, ptrs = _level_args(boxes, fabs)
    ncomp = arrs[0].shape[0]
    return bool(lib().grhd_oracle_leveldata_conserved_state_is_finite(
        cb, len(boxes), ptrs, ncomp, comp_base))


def checkpoint_pack(boxes, fabs, domain_lo, domain_hi, ghosts):
    """KITPCHK1 payload (float64, point major) of a level of FABs."""
    cb, arrs, ptrs = _level_args(boxes, fabs)
    ncomp = arrs[0].shape[0]
    n = ncomp
    for d in range(3):
        n *= domain_hi[d] - domain_lo[d] + 1 + 2 * ghosts
    out = np.empty(n)
    I3 = ctypes.c_int * 3
    lib().grhd_oracle_checkpoint_pack(cb, len(boxes), ptrs, ncomp,
                                      I3(*domain_lo), I3(*domain_hi), ghosts,
                                      _p(out))
    return out


def checkpoint_unpack(boxes, fabs, values, domain_lo, domain_hi, ghosts):
    """Fill every storage cell of `fabs` (in place) from a payload."""
    for f in fabs:
        assert f.flags["C_CONTIGUOUS"] and f.dtype == np.float64
    cb = (abi.GrhdBox * len(boxes))(*[b.to_c() for b in boxes])
    ptrs = (ctypes.c_void_p * len(fabs))(*[f.ctypes.data for f in fabs])
    I3 = ctypes.c_int * 3
    values = np.ascontiguousarray(values, dtype=np.float64)
    lib().grhd_oracle_checkpoint_unpack(cb, len(boxes), ptrs,
                                        fabs[0].shape[0], I3(*domain_lo),
                                        I3(*domain_hi), ghosts, _p(values))


# ------------------------------------------------------------ op counting
OPCOUNT_LIB = os.path.join(HERE, "libgrhd_opcount.so")
OPCOUNT_PARTS = ("metric", "slopes", "faces", "divergence", "sources",
                 "update", "recovery")
OPCOUNT_KINDS = ("add", "mul", "div", "sqrt", "other")
_opc = None


def _opcount_lib():
    global _opc
    if _opc is None:
        if not os.path.exists(OPCOUNT_LIB):
            build()
        L = ctypes.CDLL(OPCOUNT_LIB)
        P = ctypes.POINTER
        L.grhd_opcount_rk4_stage.restype = ctypes.c_int
        L.grhd_opcount_rk4_stage.argtypes = [
            P(abi.GrhdBox), P(abi.GrhdParams), P(abi.GrhdBackground),
            ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_int, P(ctypes.c_ulonglong)]
        _opc = L
    return _opc


def opcount_rk4_stage(background, box, options, state, dt, time=0.0,
                      comp_base=0):
    """Operation counts of one de-duplicated RK4 stage over the interior of
    `box` (oracle/opcount.cpp): {part: {kind: count}} plus 'cells',
    'newton_iterations', 'source_atmosphere_cells', 'reset_cells'."""
    out = (ctypes.c_ulonglong * (len(OPCOUNT_PARTS) * 5 + 4))()
    st = np.ascontiguousarray(state, dtype=np.float64)
    rc = _opcount_lib().grhd_opcount_rk4_stage(
        ctypes.byref(box.to_c()), ctypes.byref(options.to_c()),
        ctypes.byref(background.to_c()), time, dt, _p(st), st.shape[0],
        comp_base, out)
    if rc != 0:
        raise OracleError(rc, None)
    res = {}
    for p, name in enumerate(OPCOUNT_PARTS):
        res[name] = {k: int(out[p * 5 + q]) for q, k in
                     enumerate(OPCOUNT_KINDS)}
    base = len(OPCOUNT_PARTS) * 5
    res["cells"] = int(out[base])
    res["newton_iterations"] = int(out[base + 1])
    res["source_atmosphere_cells"] = int(out[base + 2])
    res["reset_cells"] = int(out[base + 3])
    return res
