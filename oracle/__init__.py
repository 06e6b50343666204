"""CPU FP64 oracle for the batched reactor-chemistry path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker / timed CPU baseline.  The product
(zero-rk_amd/) never imports it.

* ``ckparse``   - independent CHEMKIN reader + reference-layout tables.
* ``libzrk_oracle.so`` (zrk_oracle.c) - C restatement of the reference rates,
  thermo, constant-pressure RHS, analytic Jacobian and the CVODE BDF driver.

Parity status per function is stated in the test that pins it
(tests/test_oracle_*.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from . import ckparse

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return os.path.join(HERE, "libzrk_oracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libzrk_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


class _Mech(C.Structure):
    _fields_ = [
        ("nsp", C.c_int), ("nrxn", C.c_int), ("nstep", C.c_int),
        ("mw", C.c_void_p), ("thermo", C.c_void_p),
        ("reac_ptr", C.c_void_p), ("reac_sp", C.c_void_p),
        ("prod_ptr", C.c_void_p), ("prod_sp", C.c_void_p),
        ("n_arr", C.c_int), ("arr_step", C.c_void_p), ("arr_lnA", C.c_void_p),
        ("arr_beta", C.c_void_p), ("arr_Ta", C.c_void_p),
        ("n_keq", C.c_int), ("keq_step", C.c_void_p), ("keq_fwd", C.c_void_p),
        ("keq_ndelta", C.c_void_p),
        ("n_tb", C.c_int), ("tb_fwd", C.c_void_p), ("tb_rev", C.c_void_p),
        ("tb_ptr", C.c_void_p), ("tb_sp", C.c_void_p), ("tb_eff", C.c_void_p),
        ("n_fo", C.c_int), ("fo_fwd", C.c_void_p), ("fo_rev", C.c_void_p),
        ("fo_type", C.c_void_p), ("fo_spc", C.c_void_p), ("fo_param", C.c_void_p),
        ("fo_ptr", C.c_void_p), ("fo_sp", C.c_void_p), ("fo_eff", C.c_void_p),
        ("n_ni", C.c_int), ("step_ni", C.c_void_p), ("ni_step", C.c_void_p),
        ("ni_rop_ptr", C.c_void_p), ("ni_rop_sp", C.c_void_p), ("ni_rop_pow", C.c_void_p),
        ("ni_re_ptr", C.c_void_p), ("ni_re_sp", C.c_void_p), ("ni_re_nu", C.c_void_p),
        ("ni_pr_ptr", C.c_void_p), ("ni_pr_sp", C.c_void_p), ("ni_pr_nu", C.c_void_p),
        ("n_plog", C.c_int), ("plog_step", C.c_void_p), ("plog_ptr", C.c_void_p),
        ("plog_p", C.c_void_p), ("plog_lnp", C.c_void_p), ("plog_lptr", C.c_void_p),
        ("plog_lnA", C.c_void_p), ("plog_sign", C.c_void_p), ("plog_beta", C.c_void_p),
        ("plog_Ta", C.c_void_p), ("ru", C.c_double),
    ]


class Opts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("max_dt", C.c_double),
                ("nlscoef", C.c_double), ("max_steps", C.c_int), ("num_retries", C.c_int),
                ("retry_atol_adj", C.c_double), ("retry_rtol_adj", C.c_double),
                ("min_mass_fraction", C.c_double), ("step_limiter", C.c_double),
                ("Tref", C.c_double), ("constant_volume", C.c_int),
                ("delta_temperature_ignition", C.c_double), ("stop_after_ignition", C.c_int),
                ("abstol_dens", C.c_int), ("solve_temperature", C.c_int), ("dense", C.c_int),
                ("analytic", C.c_int), ("iterative", C.c_int), ("eps_lin", C.c_double),
                ("preconditioner_threshold", C.c_double)]


# the plugin's step callback type (zerork_cfd_plugin.h:44)
STEP_CB = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class Stats(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("nst", C.c_int), ("nfe", C.c_int), ("nje", C.c_int),
                ("nsetups", C.c_int), ("nni", C.c_int), ("ncfn", C.c_int), ("netf", C.c_int),
                ("status", C.c_int)]


def _declare(L):
    P = C.c_void_p
    L.zrko_default_opts.argtypes = [C.POINTER(Opts)]
    L.zrko_thermo.argtypes = [P, C.c_double, P, P, P]
    L.zrko_rate_constants.argtypes = [P, C.c_double, P, P]
    L.zrko_rates.argtypes = [P, C.c_double, P, C.c_double, P, P, P, P]
    L.zrko_rhs.argtypes = [P, P, C.c_double, C.c_double, P, P]
    L.zrko_rhs.restype = C.c_int
    L.zrko_jacobian.argtypes = [P, P, C.c_double, C.c_double, P, P, P]
    L.zrko_integrate_batch.argtypes = [P, P, C.c_int, C.c_double, P, P, P, C.c_int, P, C.c_int]
    L.zrko_integrate_batch.restype = C.c_int
    L.zrko_integrate_batch_aux.argtypes = [P, P, C.c_int, C.c_double, P, P, P, C.c_int, P,
                                           C.c_int, P, P, P, P]
    L.zrko_integrate_one_cb.argtypes = [P, P, C.c_double, P, P, P, P, STEP_CB, C.c_void_p]
    L.zrko_integrate_one_cb.restype = C.c_int
    L.zrko_integrate_batch_aux.restype = C.c_int
    L.zrko_rhs_src.argtypes = [P, P, C.c_double, P, C.c_double, P, C.c_double, C.c_double, P, P]
    L.zrko_rhs_src.restype = C.c_int
    L.zrko_dense_getrf.argtypes = [C.c_int, P, P]
    L.zrko_dense_getrf.restype = C.c_int
    L.zrko_dense_getrs.argtypes = [C.c_int, P, P, P]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().zrko_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Oracle:
    """ctypes view of one mechanism for the C oracle."""

    def __init__(self, mech_path: str, therm_path: str | None = None, ru: float | None = None):
        """ru: gas constant (J/kmol/K) for reproducing outputs of reference
        releases that used another value; default the current one."""
        t = ckparse.build_tables(mech_path, therm_path, ru=ru or ckparse.NIST_RU)
        self.tables = t
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        keep = dict(
            mw=f64(t.mw), thermo=f64(t.thermo), reac_ptr=i32(t.reac_ptr), reac_sp=i32(t.reac_sp),
            prod_ptr=i32(t.prod_ptr), prod_sp=i32(t.prod_sp), arr_step=i32(t.arr_step),
            arr_lnA=f64(t.arr_lnA), arr_beta=f64(t.arr_beta), arr_Ta=f64(t.arr_Ta),
            keq_step=i32(t.keq_step), keq_fwd=i32(t.keq_fwd), keq_ndelta=f64(t.keq_ndelta),
            tb_fwd=i32(t.tb_fwd), tb_rev=i32(t.tb_rev), tb_ptr=i32(t.tb_ptr), tb_sp=i32(t.tb_sp),
            tb_eff=f64(t.tb_eff), fo_fwd=i32(t.fo_fwd), fo_rev=i32(t.fo_rev),
            fo_type=i32(t.fo_type), fo_spc=i32(t.fo_spc), fo_param=f64(t.fo_param),
            fo_ptr=i32(t.fo_ptr), fo_sp=i32(t.fo_sp), fo_eff=f64(t.fo_eff),
            step_ni=i32(t.step_ni), ni_step=i32(t.ni_step), ni_rop_ptr=i32(t.ni_rop_ptr),
            ni_rop_sp=i32(t.ni_rop_sp), ni_rop_pow=f64(t.ni_rop_pow), ni_re_ptr=i32(t.ni_re_ptr),
            ni_re_sp=i32(t.ni_re_sp), ni_re_nu=f64(t.ni_re_nu), ni_pr_ptr=i32(t.ni_pr_ptr),
            ni_pr_sp=i32(t.ni_pr_sp), ni_pr_nu=f64(t.ni_pr_nu), plog_step=i32(t.plog_step),
            plog_ptr=i32(t.plog_ptr), plog_p=f64(t.plog_p), plog_lnp=f64(t.plog_lnp),
            plog_lptr=i32(t.plog_lptr), plog_lnA=f64(t.plog_lnA), plog_sign=f64(t.plog_sign),
            plog_beta=f64(t.plog_beta), plog_Ta=f64(t.plog_Ta))
        self._keep = keep
        m = _Mech()
        m.nsp, m.nrxn, m.nstep = t.nsp, t.nrxn, t.nstep
        for k, v in keep.items():
            setattr(m, k, _p(v))
        m.n_arr, m.n_keq, m.n_tb, m.n_fo = len(t.arr_step), len(t.keq_step), len(t.tb_fwd), len(t.fo_fwd)
        m.n_ni, m.n_plog = len(t.ni_step), len(t.plog_step)
        m.ru = float(ru or 0.0)
        self._m = m
        self.L = lib()

    @property
    def nsp(self):
        return self.tables.nsp

    @property
    def nstep(self):
        return self.tables.nstep

    def ptr(self):
        return C.byref(self._m)

    # -- reference-function restatements -------------------------------------
    def thermo(self, T):
        n = self.nsp
        cp, h, g = np.zeros(n), np.zeros(n), np.zeros(n)
        self.L.zrko_thermo(self.ptr(), T, _p(cp), _p(h), _p(g))
        return cp, h, g

    def conc(self, T, P, Y):
        mw = self.tables.mw
        wmix = 1.0 / np.sum(Y * (1.0 / mw))
        rho = P * wmix / (ckparse.NIST_RU * T)
        return rho * (1.0 / mw) * Y

    def rate_constants(self, T, C):
        K = np.zeros(self.nstep)
        self.L.zrko_rate_constants(self.ptr(), T, _p(np.ascontiguousarray(C, dtype=np.float64)), _p(K))
        return K

    def rates(self, T, C, limiter=1.0e22):
        n, s = self.nsp, self.nstep
        rop, cr, de, net = np.zeros(s), np.zeros(n), np.zeros(n), np.zeros(n)
        self.L.zrko_rates(self.ptr(), T, _p(np.ascontiguousarray(C, dtype=np.float64)), limiter,
                          _p(rop), _p(cr), _p(de), _p(net))
        return rop, cr, de, net

    def rhs(self, P, y, opts: Opts | None = None, t=0.0):
        o = opts or default_opts()
        y = np.ascontiguousarray(y, dtype=np.float64)
        yd = np.zeros_like(y)
        flag = self.L.zrko_rhs(self.ptr(), C.byref(o), P, t, _p(y), _p(yd))
        return yd, flag

    def jacobian(self, P, y, fy=None, opts: Opts | None = None, t=0.0):
        """Dense Jacobian J[row, col] (returned row-major)."""
        o = opts or default_opts()
        y = np.ascontiguousarray(y, dtype=np.float64)
        if fy is None:
            fy, _ = self.rhs(P, y, o, t)
        N = len(y)
        J = np.zeros(N * N)
        self.L.zrko_jacobian(self.ptr(), C.byref(o), P, t, _p(y), _p(np.ascontiguousarray(fy)), _p(J))
        return J.reshape(N, N).T.copy()

    def rhs_src(self, P0, y0, t, y, dpdt=0.0, esrc=0.0, ysrc=None, opts: Opts | None = None):
        """RHS at time t of the reactor initialised at (P0, y0), with the
        plugin's source terms (ysrc: nsp mass-fraction sources or None)."""
        o = opts or default_opts()
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        ys = None if ysrc is None else np.ascontiguousarray(ysrc, dtype=np.float64)
        yd = np.zeros_like(y)
        flag = self.L.zrko_rhs_src(self.ptr(), C.byref(o), P0, _p(y0), t, _p(y), dpdt, esrc,
                                   _p(ys) if ys is not None else None, _p(yd))
        return yd, flag

    def integrate_steps(self, T, P, Y, dt, opts: Opts | None = None, stop_at=None):
        """One reactor with the plugin's step callback (solver_cvode.cpp:247-257):
        returns (T_out, P_out, Y_out, stats, records), records a list of
        (CVode calls, t, t - t_prev, y, ydot) per successful CVode call, y and
        ydot of the solved variables; with stop_at = k the callback returns 1
        on record k, which ends the integration there."""
        o = opts or default_opts()
        nsp = self.nsp
        T = np.array([T], dtype=np.float64).reshape(1)
        P = np.array([P], dtype=np.float64).reshape(1)
        Y = np.array(Y, dtype=np.float64, copy=True).reshape(nsp)
        nvar = nsp + 1 if o.solve_temperature else nsp
        recs = []

        def cb(rid, nsteps, t, dtk, y, ydot, data):
            recs.append((nsteps, t, dtk, np.ctypeslib.as_array(y, (nvar,)).copy(),
                         np.ctypeslib.as_array(ydot, (nvar,)).copy()))
            return 1 if stop_at is not None and len(recs) - 1 == stop_at else 0

        fn = STEP_CB(cb)
        st = Stats()
        self.L.zrko_integrate_one_cb(self.ptr(), C.byref(o), dt, _p(T), _p(P), _p(Y),
                                     C.cast(C.pointer(st), C.c_void_p), fn, None)
        stats = {f: getattr(st, f) for f, _ in Stats._fields_}
        return T[0], P[0], Y, stats, recs

    def integrate(self, T, P, Y, dt, opts: Opts | None = None, nthreads=0, dpdt=None,
                  esrc=None, ysrc=None):
        """Integrate reactors; returns (T_out, Y_out, stats), stats a dict of
        per-reactor counters plus "P" (end pressure) and, with
        opts.delta_temperature_ignition > 0, "ignition_time".  dpdt[n], esrc[n]
        and ysrc[n, nsp] are the plugin's optional source terms."""
        o = opts or default_opts()
        T = np.array(T, dtype=np.float64, copy=True)
        P = np.array(P, dtype=np.float64, copy=True)
        Y = np.array(Y, dtype=np.float64, copy=True, order="C")
        n = len(T)
        st = (Stats * n)()
        arr = lambda a, shape: None if a is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(a, dtype=np.float64), shape))
        dp, es, ys = arr(dpdt, (n,)), arr(esrc, (n,)), arr(ysrc, Y.shape)
        ign = np.zeros(n) if o.delta_temperature_ignition > 0.0 else None
        self.L.zrko_integrate_batch_aux(self.ptr(), C.byref(o), n, dt, _p(T), _p(P), _p(Y),
                                        Y.shape[1], C.cast(st, C.c_void_p), nthreads,
                                        _p(dp) if dp is not None else None,
                                        _p(es) if es is not None else None,
                                        _p(ys) if ys is not None else None,
                                        _p(ign) if ign is not None else None)
        stats = {f: np.array([getattr(s, f) for s in st]) for f, _ in Stats._fields_}
        stats["P"] = P
        if ign is not None:
            stats["ignition_time"] = ign
        return T, Y, stats
