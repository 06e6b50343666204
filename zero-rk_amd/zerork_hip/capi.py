"""ctypes bindings of libzerork_cfd_plugin.so (include/zerork_cfd_plugin.h and
include/zrk_device.h).

The library is the product: a missing or unloadable library raises here; there
is no fallback path.  Device-level calls take device pointers (ints), e.g. from
``torch.Tensor.data_ptr()``, and a HIP stream handle (``stream.cuda_stream``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libzerork_cfd_plugin.so")

_lib = None

# zerork_status_t
STATUS = {
    0: "SUCCESS", 1: "INVALID_HANDLE", 2: "INVALID_FILENAME", 3: "INVALID_FIELD_NAME",
    4: "INVALID_POINTER", 5: "FAILED_MECHANISM_PARSE", 6: "FAILED_OPTIONS_PARSE",
    7: "FAILED_SOLVE", 8: "OPTION_NOT_DEFINED", 99: "UNKNOWN_ERROR",
}
FIELD = {"DPDT": 0, "COST": 1, "GPU": 2, "Y_SRC": 3, "E_SRC": 4, "IGNITION_TIME": 5,
         "TEMPERATURE_DELTA": 6}

# exported symbols, as declared in include/*.h
PLUGIN_SYMBOLS = [
    "zerork_reactor_init", "zerork_reactor_read_options_file",
    "zerork_reactor_set_mechanism_files", "zerork_reactor_load_mechanism",
    "zerork_reactor_solve", "zerork_reactor_set_aux_field_pointer",
    "zerork_reactor_set_int_option", "zerork_reactor_set_double_option",
    "zerork_reactor_set_string_option", "zerork_reactor_get_int_option",
    "zerork_reactor_get_double_option", "zerork_reactor_get_string_option",
    "zerork_reactor_set_callback_fn", "zerork_reactor_set_reactor_ids", "zerork_reactor_free",
]
DEVICE_SYMBOLS = [
    "zrk_bdf_default_options", "zrk_mech_load", "zrk_mech_free", "zrk_mech_info",
    "zrk_mech_species", "zrk_mech_molwt", "zrk_mech_steps", "zrk_mech_reactions",
    "zrk_rates_device",
    "zrk_jacobian_device", "zrk_integrate_device", "zrk_integrate_device_aux",
    "zrk_integrator_resident_waves",
    "zrk_autotune_integrator", "zrk_integrator_set_team_width", "zrk_integrator_set_sparse",
    "zrk_sparse_plan_info", "zrk_integrate_device_noalloc",
]


class BdfOptions(C.Structure):
    _fields_ = [
        ("rel_tol", C.c_double), ("abs_tol", C.c_double), ("max_dt", C.c_double),
        ("nonlinear_convergence_coeff", C.c_double), ("step_limiter", C.c_double),
        ("reference_temperature", C.c_double), ("min_mass_fraction", C.c_double),
        ("cvode_retry_absolute_tolerance_adjustment", C.c_double),
        ("cvode_retry_relative_tolerance_adjustment", C.c_double),
        ("max_steps", C.c_int), ("cvode_num_retries", C.c_int), ("constant_volume", C.c_int),
        ("fp32_preconditioner", C.c_int), ("delta_temperature_ignition", C.c_double),
        ("stop_after_ignition", C.c_int), ("abstol_dens", C.c_int), ("dense", C.c_int),
        ("analytic", C.c_int), ("iterative", C.c_int), ("eps_lin", C.c_double),
        ("preconditioner_threshold", C.c_double),
    ]


def lib():
    """Load the product library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {PKG}` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    P, I, D = C.c_void_p, C.c_int, C.c_double
    L.zerork_reactor_init.restype = P
    L.zerork_reactor_init.argtypes = []
    for name, args in [
        ("zerork_reactor_read_options_file", [C.c_char_p, P]),
        ("zerork_reactor_set_mechanism_files", [C.c_char_p, C.c_char_p, P]),
        ("zerork_reactor_load_mechanism", [P]),
        ("zerork_reactor_solve", [I, D, D, I, P, P, P, P]),
        ("zerork_reactor_set_aux_field_pointer", [I, P, P]),
        ("zerork_reactor_set_int_option", [C.c_char_p, I, P]),
        ("zerork_reactor_set_double_option", [C.c_char_p, D, P]),
        ("zerork_reactor_set_string_option", [C.c_char_p, C.c_char_p, P]),
        ("zerork_reactor_get_int_option", [C.c_char_p, C.POINTER(I), P]),
        ("zerork_reactor_get_double_option", [C.c_char_p, C.POINTER(D), P]),
        ("zerork_reactor_get_string_option", [C.c_char_p, C.POINTER(C.c_char_p), I, P]),
        ("zerork_reactor_set_callback_fn", [P, P, P]),
        ("zerork_reactor_set_reactor_ids", [P, P]),
        ("zerork_reactor_free", [P]),
    ]:
        f = getattr(L, name)
        f.argtypes = args
        f.restype = I
    L.zrk_bdf_default_options.argtypes = [C.POINTER(BdfOptions)]
    L.zrk_bdf_default_options.restype = None
    L.zrk_mech_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(P), C.c_char_p, I]
    L.zrk_mech_load.restype = I
    L.zrk_mech_free.argtypes = [P]
    L.zrk_mech_free.restype = None
    L.zrk_mech_info.argtypes = [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
    L.zrk_mech_info.restype = I
    L.zrk_mech_species.argtypes = [P, I, C.c_char_p, I]
    L.zrk_mech_species.restype = I
    L.zrk_mech_molwt.argtypes = [P, P]
    L.zrk_mech_molwt.restype = I
    L.zrk_mech_steps.argtypes = [P, P, P, P, P, P, P]
    L.zrk_mech_steps.restype = I
    L.zrk_mech_reactions.argtypes = [P, P, P, P, P, P]
    L.zrk_mech_reactions.restype = I
    L.zrk_rates_device.argtypes = [P, I, P, P, P, I, D, P, P, P, P, P, P]
    L.zrk_rates_device.restype = I
    L.zrk_jacobian_device.argtypes = [P, C.POINTER(BdfOptions), I, P, P, P, I, P, P]
    L.zrk_jacobian_device.restype = I
    L.zrk_integrate_device.argtypes = [P, C.POINTER(BdfOptions), I, D, P, P, P, I, P, P, P]
    L.zrk_integrate_device.restype = I
    L.zrk_integrate_device_aux.argtypes = [P, C.POINTER(BdfOptions), I, D, P, P, P, I, P, P, P,
                                           P, P, P, P]
    L.zrk_integrate_device_aux.restype = I
    L.zrk_integrate_device_noalloc.argtypes = [P, C.POINTER(BdfOptions), I, D, P, P, P, I, P,
                                               P, P]
    L.zrk_integrate_device_noalloc.restype = I
    L.zrk_integrator_resident_waves.argtypes = [P, P]
    L.zrk_autotune_integrator.argtypes = [P, C.POINTER(BdfOptions), I, D, P, P, P, I, P, P, P]
    L.zrk_autotune_integrator.restype = I
    L.zrk_integrator_set_team_width.argtypes = [P, I, I]
    L.zrk_integrator_set_team_width.restype = I
    L.zrk_integrator_set_sparse.argtypes = [P, I]
    L.zrk_integrator_set_sparse.restype = I
    L.zrk_sparse_plan_info.argtypes = [P, C.POINTER(I), C.POINTER(I),
                                       C.POINTER(C.c_longlong)]
    L.zrk_sparse_plan_info.restype = I
    L.zrk_integrator_resident_waves.restype = I


def default_options(**kw) -> BdfOptions:
    o = BdfOptions()
    lib().zrk_bdf_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


_F64, _I32 = "float64", "int32"


def _dev(x, kind, numel, name):
    """Device pointer of x after checking on the host that the kernels' reads
    and writes stay inside it: a torch tensor must live on the GPU, be
    contiguous, have the expected element type (kind "float64" or "int32")
    and hold at least numel elements.  None passes as NULL; a plain int is
    taken as a raw device address (the caller vouches for it)."""
    if x is None or isinstance(x, int):
        return x
    if not getattr(x, "is_cuda", False):
        raise ValueError(f"{name}: expected a device (cuda) tensor")
    if str(x.dtype).rsplit(".", 1)[-1] != kind:
        raise ValueError(f"{name}: expected {kind}, got {x.dtype}")
    if not x.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if x.numel() < numel:
        raise ValueError(f"{name}: {x.numel()} elements, at least {numel} needed")
    return x.data_ptr()


class Mech:
    """A mechanism loaded into the library (parsed on the host; uploaded to the
    device on first device call)."""

    def __init__(self, mech_file: str, therm_file: str | None = None):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = L.zrk_mech_load(mech_file.encode(), (therm_file or "").encode(), C.byref(h), err, 512)
        if rc != 0:
            raise RuntimeError(f"zrk_mech_load failed ({rc}): {err.value.decode()}")
        self.h = h
        nsp, nrxn, nstep = C.c_int(), C.c_int(), C.c_int()
        L.zrk_mech_info(h, C.byref(nsp), C.byref(nrxn), C.byref(nstep))
        self.nsp, self.nrxn, self.nstep = nsp.value, nrxn.value, nstep.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().zrk_mech_free(self.h)
                self.h = None
        except Exception:
            pass

    def species(self):
        buf = C.create_string_buffer(64)
        out = []
        for j in range(self.nsp):
            lib().zrk_mech_species(self.h, j, buf, 64)
            out.append(buf.value.decode())
        return out

    def molwt(self):
        import numpy as np
        mw = np.zeros(self.nsp)
        lib().zrk_mech_molwt(self.h, mw.ctypes.data)
        return mw

    def steps(self):
        import numpy as np
        ns = self.nstep
        srx, sdir = np.zeros(ns, np.int32), np.zeros(ns, np.int32)
        rp, pp = np.zeros(ns + 1, np.int32), np.zeros(ns + 1, np.int32)
        lib().zrk_mech_steps(self.h, srx.ctypes.data, sdir.ctypes.data, rp.ctypes.data, None,
                             pp.ctypes.data, None)
        rs, ps = np.zeros(max(int(rp[-1]), 1), np.int32), np.zeros(max(int(pp[-1]), 1), np.int32)
        lib().zrk_mech_steps(self.h, None, None, None, rs.ctypes.data, None, ps.ctypes.data)
        return dict(step_rxn=srx, step_dir=sdir, reac_ptr=rp, reac_sp=rs[:rp[-1]],
                    prod_ptr=pp, prod_sp=ps[:pp[-1]])

    def reactions(self):
        import numpy as np
        n = self.nrxn
        a = {k: np.zeros(n, np.int32) for k in ("fwd_step", "rev_step", "rev_type", "n_enhanced",
                                                   "falloff_type")}
        lib().zrk_mech_reactions(self.h, *(a[k].ctypes.data for k in ("fwd_step", "rev_step",
                                                                      "rev_type", "n_enhanced",
                                                                      "falloff_type")))
        return a

    # -- device entry points ---------------------------------------------------
    def rates(self, n, T, P, Y, stride, step_limiter=1.0e22, ydot=None, rop=None, create=None,
              destroy=None, net=None, stream=0):
        ns = self.nsp
        if stride < ns:
            raise ValueError(f"stride {stride} < {ns} species")
        rc = lib().zrk_rates_device(
            self.h, n, _dev(T, _F64, n, "T"), _dev(P, _F64, n, "P"),
            _dev(Y, _F64, (n - 1) * stride + ns if n else 0, "Y"), stride, step_limiter,
            _dev(ydot, _F64, n * (ns + 1), "ydot"), _dev(rop, _F64, n * self.nstep, "rop"),
            _dev(create, _F64, n * ns, "create"), _dev(destroy, _F64, n * ns, "destroy"),
            _dev(net, _F64, n * ns, "net"), stream)
        if rc != 0:
            raise RuntimeError(f"zrk_rates_device failed ({rc})")

    def jacobian(self, n, T, P, Y, stride, J, opts: BdfOptions | None = None, stream=0):
        o = opts or default_options()
        ns = self.nsp
        if stride < ns:
            raise ValueError(f"stride {stride} < {ns} species")
        rc = lib().zrk_jacobian_device(
            self.h, C.byref(o), n, _dev(T, _F64, n, "T"), _dev(P, _F64, n, "P"),
            _dev(Y, _F64, (n - 1) * stride + ns if n else 0, "Y"), stride,
            _dev(J, _F64, n * (ns + 1) ** 2, "J"), stream)
        if rc != 0:
            raise RuntimeError(f"zrk_jacobian_device failed ({rc})")

    def integrate(self, n, dt, T, P, Y, stride, opts: BdfOptions | None = None, nsteps=None,
                  stats=None, stream=0, dpdt=None, esrc=None, ysrc=None, ignition_time=None):
        """Integrate n reactors (device tensors); dpdt[n], esrc[n], ysrc[n*stride]
        are the optional source terms and ignition_time[n] the optional output
        of ignition times (with opts.delta_temperature_ignition > 0), all
        device tensors."""
        o = opts or default_options()
        ns = self.nsp
        if stride < ns:
            raise ValueError(f"stride {stride} < {ns} species")
        ny = (n - 1) * stride + ns if n else 0
        T, P, Y = _dev(T, _F64, n, "T"), _dev(P, _F64, n, "P"), _dev(Y, _F64, ny, "Y")
        nsteps, stats = _dev(nsteps, _I32, n, "nsteps"), _dev(stats, _I32, 8 * n, "stats")
        if dpdt is None and esrc is None and ysrc is None and ignition_time is None:
            rc = lib().zrk_integrate_device(self.h, C.byref(o), n, dt, T, P, Y, stride, nsteps,
                                            stats, stream)
        else:
            rc = lib().zrk_integrate_device_aux(
                self.h, C.byref(o), n, dt, T, P, Y, stride, _dev(dpdt, _F64, n, "dpdt"),
                _dev(esrc, _F64, n, "esrc"), _dev(ysrc, _F64, ny, "ysrc"),
                _dev(ignition_time, _F64, n, "ignition_time"), nsteps, stats, stream)
        if rc != 0:
            raise RuntimeError(f"zrk_integrate_device failed ({rc})")

    def autotune(self, n, dt, T, P, Y, stride, opts: BdfOptions, stream=0):
        """Time each team size on this batch, keep the fastest; returns
        (wavefronts per reactor, {size: ms})."""
        choice = C.c_int(0)
        ms = (C.c_float * 4)()
        ns = self.nsp
        if stride < ns:
            raise ValueError(f"stride {stride} < {ns} species")
        rc = lib().zrk_autotune_integrator(
            self.h, C.byref(opts), n, dt, _dev(T, _F64, n, "T"), _dev(P, _F64, n, "P"),
            _dev(Y, _F64, (n - 1) * stride + ns if n else 0, "Y"), stride, C.byref(choice), ms,
            stream)
        if rc != 0:
            raise RuntimeError(f"zrk_autotune_integrator failed ({rc})")
        return choice.value, {w: ms[w] for w in (1, 2, 3) if ms[w] >= 0}

    def set_sparse(self, on=True):
        """Route later integrations through the sparse path (any size)."""
        if lib().zrk_integrator_set_sparse(self.h, int(bool(on))) != 0:
            raise RuntimeError("zrk_integrator_set_sparse failed")

    def sparse_plan_info(self):
        """(LU entries, forward levels, multiply-adds per factorisation) of the
        sparse plan (builds it; needs the device)."""
        nnz, lev, fl = C.c_int(0), C.c_int(0), C.c_longlong(0)
        rc = lib().zrk_sparse_plan_info(self.h, C.byref(nnz), C.byref(lev),
                                        C.byref(fl))
        if rc != 0:
            raise RuntimeError(f"zrk_sparse_plan_info failed ({rc})")
        return nnz.value, lev.value, fl.value

    def set_team_width(self, fp32, W):
        rc = lib().zrk_integrator_set_team_width(self.h, int(fp32), int(W))
        if rc != 0:
            raise RuntimeError(f"zrk_integrator_set_team_width failed ({rc})")

    def resident_waves(self, opts: BdfOptions | None = None):
        return lib().zrk_integrator_resident_waves(self.h, C.byref(opts) if opts else None)


# the step callback type (zerork_cfd_plugin.h:44)
STEP_CB = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class Plugin:
    """Thin Python view of the plugin C API (mirrors zerork_cfd_plugin_tester)."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.zerork_reactor_init()
        if not self.h:
            raise RuntimeError("zerork_reactor_init failed")
        self._keep = []

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {STATUS.get(rc, rc)}")
        return rc

    def set_mechanism_files(self, mech, therm):
        return self.L.zerork_reactor_set_mechanism_files(mech.encode(), therm.encode(), self.h)

    def load_mechanism(self):
        return self.L.zerork_reactor_load_mechanism(self.h)

    def read_options_file(self, path):
        return self.L.zerork_reactor_read_options_file(path.encode(), self.h)

    def set_int(self, k, v):
        return self.L.zerork_reactor_set_int_option(k.encode(), int(v), self.h)

    def set_double(self, k, v):
        return self.L.zerork_reactor_set_double_option(k.encode(), float(v), self.h)

    def set_string(self, k, v):
        return self.L.zerork_reactor_set_string_option(k.encode(), v.encode(), self.h)

    def get_int(self, k):
        v = C.c_int()
        rc = self.L.zerork_reactor_get_int_option(k.encode(), C.byref(v), self.h)
        return rc, v.value

    def get_double(self, k):
        v = C.c_double()
        rc = self.L.zerork_reactor_get_double_option(k.encode(), C.byref(v), self.h)
        return rc, v.value

    def get_string(self, k, n=256):
        buf = C.create_string_buffer(n)
        p = C.c_char_p(C.addressof(buf))
        rc = self.L.zerork_reactor_get_string_option(k.encode(), C.byref(p), n, self.h)
        return rc, buf.value.decode()

    def set_callback(self, fn, data=None):
        """zerork_reactor_set_callback_fn with a Python callable
        fn(reactor_id, nsteps, t, dt, y, ydot, data) -> int (y and ydot are
        double pointers), kept alive with the handle."""
        cb = STEP_CB(fn)
        self._keep.append(cb)
        return self.L.zerork_reactor_set_callback_fn(C.cast(cb, C.c_void_p), data, self.h)

    def set_reactor_ids(self, ids):
        self._keep.append(ids)
        return self.L.zerork_reactor_set_reactor_ids(ids.ctypes.data, self.h)

    def set_aux(self, field, arr):
        self._keep.append(arr)
        return self.L.zerork_reactor_set_aux_field_pointer(FIELD[field], arr.ctypes.data, self.h)

    def solve(self, n_cycle, time, dt, T, P, mf):
        n = len(T)
        return self.L.zerork_reactor_solve(n_cycle, time, dt, n, T.ctypes.data, P.ctypes.data,
                                           mf.ctypes.data, self.h)

    def free(self):
        if self.h:
            rc = self.L.zerork_reactor_free(self.h)
            self.h = None
            return rc
        return 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
