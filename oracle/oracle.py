"""ctypes bindings to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Two libraries are wrapped:

* ``liboracle_mpdata.so``: the C restatement of the reference MPDATA path
  (``oracle/mpdata_oracle_impl.inc``), built by ``oracle/Makefile``.
* ``_ref/libmpdata_ref.so`` and ``_ref/libmpdata_ref32.so``: the reference's
  own Fortran routines compiled from the advectionHPCtester sources by
  ``oracle/build_ref.sh`` with the bind(C) shim ``oracle/ref_shim.f90``, in
  its default FP64 kind and with its -DFLOAT_PRECISION=4 switch (FP32).
  Present only where they were built.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, and only as the checker.

Fields are numpy arrays of shape (lp+2ih, mp+2ih, np+2ih) in C order, which
is the memory layout of the reference's Fortran arrays (np+2ih, mp+2ih,
lp+2ih) with lower bound 1-ih; face arrays add one along their normal.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_mpdata.so")
REF_SO = os.path.join(HERE, "_ref", "libmpdata_ref.so")
REF32_SO = os.path.join(HERE, "_ref", "libmpdata_ref32.so")

UPWIND, GAUGE, STANDARD = 0, 1, 2

_P = ctypes.c_void_p
_I = ctypes.c_int


def shape(np_, mp, lp, ih, kind=0):
    """numpy shape of a field: kind 0 scalar, 1/2/3 faces normal to x/y/z."""
    return (lp + 2 * ih + (kind == 3), mp + 2 * ih + (kind == 2), np_ + 2 * ih + (kind == 1))


def zeros(np_, mp, lp, ih, kind=0, dtype=np.float64):
    return np.zeros(shape(np_, mp, lp, ih, kind), dtype=dtype)


def _ptr(a):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_P)


class Oracle:
    """The C restatement.  ``dtype`` selects the _f64 or _f32 instantiation."""

    def __init__(self, dtype=np.float64, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = ctypes.CDLL(path)
        self.dtype = np.dtype(dtype)
        self.sfx = "_f64" if self.dtype == np.float64 else "_f32"
        self.real = ctypes.c_double if self.sfx == "_f64" else ctypes.c_float

    def _fn(self, name):
        return getattr(self.lib, name + self.sfx)

    def _chk(self, *arrays):
        for a in arrays:
            if a is not None:
                assert a.dtype == self.dtype, (a.dtype, self.dtype)

    def upwind(self, u1, u2, u3, x, xant, rhr, h, bcx, bcy, ibc, dims):
        self._chk(u1, u2, u3, x, xant, rhr, h, bcx, bcy)
        np_, mp, lp, ih = dims
        self._fn("orc_upwind3d_gpubc")(*map(_ptr, (u1, u2, u3, x, xant, rhr, h, bcx, bcy)),
                                       *ibc, np_, mp, lp, ih)

    def scratch(self, dims):
        np_, mp, lp, ih = dims
        z = lambda k: zeros(np_, mp, lp, ih, k, self.dtype)  # noqa: E731
        return {"v1": z(1), "v2": z(2), "v3": z(3), "cp": z(0), "cn": z(0),
                "f1ns": z(1), "f2ns": z(2), "f3ns": z(3), "f1": z(1), "f2": z(2), "f3": z(3)}

    def _scratch_struct(self, s):
        class S(ctypes.Structure):
            _fields_ = [(n, _P) for n in ("v1", "v2", "v3", "cp", "cn", "f1ns", "f2ns",
                                          "f3ns", "f1", "f2", "f3")]
        return S(*[_ptr(s[n]) for n, _ in S._fields_])

    def antidiff_pass(self, variant, nonos, u1, u2, u3, psi, x0, out, rhr, h, ibc, dims,
                      scratch=None, ul=(None, None, None)):
        scratch = scratch if scratch is not None else self.scratch(dims)
        self._chk(u1, u2, u3, psi, x0, out, rhr, h, *ul)
        st = self._scratch_struct(scratch)
        self._fn("orc_antidiff_pass")(variant, nonos, *map(_ptr, (u1, u2, u3, psi, x0, out, rhr, h)),
                                      *ibc, *dims, ctypes.byref(st), *map(_ptr, ul))
        return scratch

    def antidiff(self, variant, u1, u2, u3, xant, x, rhr, h, ibc, dims, scratch=None):
        return self.antidiff_pass(variant, 1, u1, u2, u3, xant, x, x, rhr, h, ibc, dims, scratch)

    def step(self, variant, iters, nonos, x, u1, u2, u3, rhr, h, bcx, bcy, ibc, dims):
        self._chk(x, u1, u2, u3, rhr, h, bcx, bcy)
        rc = self._fn("orc_mpdata_step")(variant, iters, nonos,
                                         *map(_ptr, (x, u1, u2, u3, rhr, h, bcx, bcy)), *ibc, *dims)
        if rc != 0:
            raise MemoryError("orc_mpdata_step")

    def velprdA_to_C(self, ox, oy, oz, u1, u2, u3, ibc, dims):
        self._fn("orc_velprdA_to_C")(*map(_ptr, (ox, oy, oz, u1, u2, u3)), *ibc, *dims)

    def velprd_traj0(self, ox, oy, oz, u1, u2, u3, h, dxi, dyi, dzi, dt, dt_ratio, ibc, dims):
        R = self.real
        self._fn("orc_velprd_traj0")(*map(_ptr, (ox, oy, oz, u1, u2, u3, h)), R(dxi), R(dyi),
                                     R(dzi), R(dt), R(dt_ratio), *ibc, *dims)

    def velprd_rotate(self, u, v, w, ox, oy, oz, h, dt_ratio, dims):
        self._fn("orc_velprd_rotate")(*map(_ptr, (u, v, w, ox, oy, oz, h)), self.real(dt_ratio),
                                      *dims)

    def compute_courants(self, ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt, dims):
        R = self.real
        out = (R * 7)()
        self._fn("orc_compute_courants")(*map(_ptr, (ox, oy, oz, h)), ntimelevel, R(dxi), R(dyi),
                                         R(dzi), R(dt), *dims, out)
        return np.array(list(out), dtype=self.dtype)


    # -- density step with the predictor driver (mpdata3d-rho-gauge-gpubc.f90) --
    RHO_FIELDS = ("ox", "oy", "oz", "u", "v", "w", "u1", "u2", "u3", "x_in", "x_incomp", "h",
                  "rhoadv", "rhr", "pexe", "pexr1", "pexr2", "bcx", "bcy")

    def rho_gauge(self, f, flags, state, crmaxlim, dtmax, liner, dims, part="all"):
        """One call of the density step on the dict of arrays ``f`` (updated in
        place).  flags: dict of lsubstepping, ltimeadapt, lvertsplit,
        lcrdiag_nphlf; state: dict of nsubsteps, nsubsteps_o, dtn, dtnold
        (returned updated with the status).  part: "all", "driver" (the
        predictor driver only) or "core" (the step after the driver, with
        the C-grid velocities already in u1..u3)."""
        R = self.real
        np_, mp, lp, ih = dims
        self._chk(*[f[k] for k in self.RHO_FIELDS])
        ns, nso = ctypes.c_int(state["nsubsteps"]), ctypes.c_int(state["nsubsteps_o"])
        dtn, dtnold = R(state["dtn"]), R(state["dtnold"])
        fl = [int(bool(flags.get(k, 0))) for k in ("lsubstepping", "ltimeadapt", "lvertsplit",
                                                    "lcrdiag_nphlf")]
        sp = [R(f[k]) for k in ("dxi", "dyi", "dzi", "dt")]
        lev = [_ptr(f[k]) for k in ("ox", "oy", "oz", "u", "v", "w", "u1", "u2", "u3")]
        sc = [_ptr(f[k]) for k in ("x_in", "x_incomp", "h", "rhoadv", "rhr", "pexe", "pexr1",
                                   "pexr2", "bcx", "bcy")]
        if part == "core":
            rc = self._fn("orc_rho_gauge_core")(liner, *lev[6:], *sc, 0, 0, 0, *dims)
        elif part == "driver":
            rc = self._fn("orc_velprd_driver")(ctypes.byref(ns), ctypes.byref(nso), fl[0],
                                               ctypes.byref(dtn), ctypes.byref(dtnold), fl[1],
                                               fl[2], fl[3], R(crmaxlim), R(dtmax), *lev,
                                               _ptr(f["h"]), 0, 0, 0, *dims, *sp)
        else:
            rc = self._fn("orc_rho_gauge_step")(ctypes.byref(ns), ctypes.byref(nso), fl[0],
                                                ctypes.byref(dtn), ctypes.byref(dtnold), fl[1],
                                                fl[2], fl[3], R(crmaxlim), R(dtmax), *lev, *sc,
                                                0, 0, 0, *dims, *sp, liner)
        return {"nsubsteps": ns.value, "nsubsteps_o": nso.value, "dtn": dtn.value,
                "dtnold": dtnold.value, "status": rc}


class Reference:
    """The reference's own routines through oracle/ref_shim.f90: the FP64 build
    (``dtype`` float64) or the -DFLOAT_PRECISION=4 build (float32).  Scalar
    arguments cross the shim as C doubles and are converted to the build's
    real kind inside it, as the reference converts its parameters.

    The reference keeps its domain and scratch in module globals, so a
    ``setup`` call fixes the grid for all following calls."""

    def __init__(self, dtype=np.float64, path=None):
        self.dtype = np.dtype(dtype)
        path = path or (REF_SO if self.dtype == np.float64 else REF32_SO)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (oracle/build_ref.sh)")
        # a separate handle per build: both export the same symbol names
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
        kind = self.lib.ref_precision()
        if kind != self.dtype.itemsize:
            raise RuntimeError(f"{path}: real kind {kind}, expected {self.dtype.itemsize}")
        self.dims = None

    def _p(self, a):
        if a is not None:
            assert a.dtype == self.dtype, (a.dtype, self.dtype)
        return _ptr(a)

    @staticmethod
    def available(dtype=np.float64):
        return os.path.exists(REF_SO if np.dtype(dtype) == np.float64 else REF32_SO)

    def setup(self, np_, mp, lp, ih=1, ibc=(0, 0, 0), dx00=100.0, nt=556, dt=0.0):
        self.lib.ref_setup(np_, mp, lp, ih, *ibc, ctypes.c_double(dx00), nt, ctypes.c_double(dt))
        self.dims = (np_, mp, lp, ih)

    def upwind(self, u1, u2, u3, x, xant, rhr, h, bcx, bcy):
        self.lib.ref_upwind(*map(self._p, (u1, u2, u3, x, xant, rhr, h, bcx, bcy)))

    def antidiff(self, variant, u1, u2, u3, xant, x, rhr, h, out=None):
        o = out or {}
        names = ("v1", "v2", "v3", "cp", "cn", "f1", "f2", "f3", "f1ns", "f2ns", "f3ns")
        self.lib.ref_antidiff(variant, *map(self._p, (u1, u2, u3, xant, x, rhr, h)),
                              *[self._p(o.get(n)) for n in names])

    def advec_dwarf(self, legacyoptmode, t_adv, x, bcx, bcy, u1, u2, u3, rhr, h):
        self.lib.ref_advec_dwarf(legacyoptmode, t_adv, *map(self._p, (x, bcx, bcy, u1, u2, u3, rhr, h)))

    def set_initial15(self, u1, u2, u3, x, rho):
        self.lib.ref_set_initial15(*map(self._p, (u1, u2, u3, x, rho)))

    def norms(self, x, x0, icnt):
        self.lib.ref_norms(self._p(x), self._p(x0), icnt)

    def velprdA_to_C(self, ver, ox, oy, oz, u1, u2, u3):
        self.lib.ref_velprdA_to_C(ver, *map(self._p, (ox, oy, oz, u1, u2, u3)))

    def velprd_traj0(self, ox, oy, oz, u1, u2, u3, h, dxi, dyi, dzi, dt, dt_ratio):
        D = ctypes.c_double
        self.lib.ref_velprd_traj0(*map(self._p, (ox, oy, oz, u1, u2, u3, h)), D(dxi), D(dyi), D(dzi),
                                  D(dt), D(dt_ratio))

    def compute_courants(self, ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt):
        D = ctypes.c_double
        out = (D * 7)()
        self.lib.ref_compute_courants(*map(self._p, (ox, oy, oz, h)), ntimelevel, D(dxi), D(dyi),
                                      D(dzi), D(dt), out)
        return np.array(list(out))

    def rho_gauge(self, f, flags, state, crmaxlim, dtmax, liner, only_driver=False):
        """The reference's mpdata3d_rho_gauge_gpubc (or its velprd_driver alone)
        on the dict of FP64 arrays ``f``; see :meth:`Oracle.rho_gauge`."""
        D = ctypes.c_double
        fl = (ctypes.c_int * 5)(int(bool(flags.get("lupdatemulti", 0))),
                                *[int(bool(flags.get(k, 0))) for k in
                                  ("lsubstepping", "ltimeadapt", "lvertsplit", "lcrdiag_nphlf")])
        isub = (ctypes.c_int * 2)(state["nsubsteps"], state["nsubsteps_o"])
        rdt = (D * 4)(state["dtn"], state["dtnold"], crmaxlim, dtmax)
        sp = (D * 4)(*[f[k] for k in ("dxi", "dyi", "dzi", "dt")])
        lin = ctypes.c_int(liner)
        names = ("ox", "oy", "oz", "u", "v", "w", "u1", "u2", "u3", "x_in", "x_incomp", "h",
                 "rhoadv", "rhr", "pexe", "pexr1", "pexr2", "bcx", "bcy")
        self.lib.ref_rho_gauge(int(bool(only_driver)), fl, isub, rdt,
                               *[self._p(f[k]) for k in names], sp, ctypes.byref(lin))
        return {"nsubsteps": isub[0], "nsubsteps_o": isub[1], "dtn": rdt[0], "dtnold": rdt[1],
                "status": 0}
