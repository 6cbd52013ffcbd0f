"""ctypes binding of libcgli.so (include/cgli.h).

The HIP library is the only compute path: if it is missing or no device is
present, the calls raise.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libcgli.so")

# field ids (include/cgli.h)
U_, U, UQ, J = 0, 5, 10, 15
MET, J_U, MET_U, J_V, MET_V, J_W, MET_W = 16, 25, 26, 35, 36, 45, 46
INFX, INFY, INFZ, VF, MR, NL, NR, RES = 55, 60, 65, 70, 74, 79, 84, 89
EPX, EPY, EPZ, YPOINT, NFIELDS = 94, 95, 96, 97, 98

METRIC_NAMES = ["xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"]

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "device allocation failed",
          -4: "non-finite residual norm", -5: "a required input field was never set", -6: "no HIP device",
          -7: "bounded device wait expired", -8: "restart file could not be opened, read or written",
          -9: "the grid exceeds the fixed scratch of the allocation-free entry point"}


class CgliError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s: %s (%d)" % (what, ERRORS.get(code, "unknown"), code))
        self.code = code


class Config(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nz", "first", "last", "gstart", "gend0", "nx_global",
        "nx_inlet", "nx_outlet", "sponge", "xi_mode", "x_periodic", "precision")] + [(n, C.c_double) for n in (
        "dxi", "deta", "dzeta", "high", "K", "R", "Cv", "mu_L", "Pr_L", "suth_T",
        "suth_S", "rho_ref", "grav", "f_const", "f_dplusgs", "E_ref")]


class Channel(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Lx", "Lz", "beta", "skew", "mach", "re_tau", "re_c", "Pr", "suth_S",
                                          "cfl", "noise", "precond_e")] + [("seed", C.c_ulonglong),
                                                                           ("xi_mode", C.c_int)]


_lib = None


def lib():
    """Load libcgli.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcgli.so is not built (run __graft_entry__.build() or "
                              "python -m cgli_fu_flow_amd.build): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.cgli_config_reference.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.c_int]
        L.cgli_config_reference.restype = None
        L.cgli_create.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(vp)]
        L.cgli_destroy.argtypes = [vp]
        L.cgli_error_string.restype = C.c_char_p
        L.cgli_set_field.argtypes = [vp, C.c_int, dp]
        L.cgli_get_field.argtypes = [vp, C.c_int, dp]
        L.cgli_field_device_ptr.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.cgli_synchronize.argtypes = [vp]
        for n in ("cgli_flux_x", "cgli_flux_y", "cgli_flux_z"):
            getattr(L, n).argtypes = [vp, C.c_double, C.c_double]
        L.cgli_viscous.argtypes = [vp]
        L.cgli_set_dissipation_scaling.argtypes = [vp, C.c_int]
        L.cgli_filter.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.cgli_profiles.argtypes = [vp, dp]
        L.cgli_channel_last_step.argtypes = [vp, C.POINTER(C.c_int)]
        L.cgli_hyperplane.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cgli_dual_time_stepping.argtypes = [vp, C.c_double, C.c_double, dp, C.c_int, C.c_int]
        L.cgli_dplusgs.argtypes = [vp, C.c_double, dp, C.c_double, C.c_int]
        L.cgli_filter_energies.argtypes = [vp, dp, dp]
        L.cgli_dplusgs_noalloc.argtypes = [vp, C.c_double, dp, C.c_double, C.c_int]
        L.cgli_dual_time_stepping_async.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int]
        L.cgli_dual_time_stepping_wait.argtypes = [vp, dp]
        L.cgli_graph_begin.argtypes = [vp]
        L.cgli_graph_end.argtypes = [vp, C.POINTER(vp)]
        L.cgli_graph_launch.argtypes = [vp]
        L.cgli_graph_destroy.argtypes = [vp]
        L.cgli_channel_set_graph.argtypes = [vp, C.c_int]
        L.cgli_bc_y.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int]
        L.cgli_bc_z.argtypes = [vp]
        L.cgli_bc_x_periodic.argtypes = [vp]
        L.cgli_write_restart.argtypes = [vp, C.c_char_p]
        L.cgli_read_restart.argtypes = [vp, C.c_char_p]
        L.cgli_channel_defaults.argtypes = [C.POINTER(Channel), C.c_double]
        L.cgli_channel_defaults.restype = None
        L.cgli_channel_config.argtypes = [C.POINTER(Channel), C.c_int, C.c_int, C.c_int, C.POINTER(Config)]
        L.cgli_channel_init.argtypes = [vp, C.POINTER(Channel), dp]
        L.cgli_channel_ghosts.argtypes = [vp]
        L.cgli_channel_step.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), dp, dp]
        L.cgli_set_lusgs_mode.argtypes = [vp, C.c_int]
        L.cgli_set_sweep_waves.argtypes = [vp, C.c_int]
        L.cgli_set_spin_limit.argtypes = [vp, C.c_int]
        L.cgli_timing_enable.argtypes = [vp, C.c_int]
        L.cgli_timing_reset.argtypes = [vp]
        L.cgli_timing_select.argtypes = [vp, C.c_uint]
        L.cgli_timing_get.argtypes = [vp, C.c_int, dp, C.POINTER(C.c_long)]
        L.cgli_channel_set_rule.argtypes = [vp, C.c_int]
        L.cgli_device_memory.argtypes = [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.cgli_stream_copy.argtypes = [C.c_int, C.c_size_t, C.c_int, dp]
        L.cgli_channel_rule_state.argtypes = [vp, dp, C.c_int]
        L.cgli_channel_set_guess.argtypes = [vp, C.c_int]
        L.cgli_channel_attach_net.argtypes = [vp, vp]
        L.cgli_channel_guess_info.argtypes = [vp, C.POINTER(C.c_int), dp, dp]
        fp = C.POINTER(C.c_float)
        up = C.POINTER(C.c_ubyte)
        L.cgli_net_create.argtypes = [C.c_int, C.c_ulonglong, C.POINTER(vp)]
        L.cgli_net_destroy.argtypes = [vp]
        L.cgli_net_params.argtypes = [vp, fp, C.c_int]
        L.cgli_net_eval.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, up, fp, dp, fp]
        L.cgli_net_adam.argtypes = [vp, C.c_float]
        L.cgli_net_ring.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.cgli_net_push.argtypes = [vp, fp]
        L.cgli_net_records.argtypes = [vp, C.POINTER(C.c_long)]
        L.cgli_net_rewind.argtypes = [vp, C.c_long]
        L.cgli_net_record_scales.argtypes = [vp, C.c_long, fp]
        L.cgli_net_train.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp]
        L.cgli_net_predict.argtypes = [vp, fp, fp]
        L.cgli_net_timing.argtypes = [vp, dp, C.POINTER(C.c_long), dp, C.POINTER(C.c_long)]
        _lib = L
    return _lib


def device_memory(device=0):
    """(free, total) bytes of one device."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(lib().cgli_device_memory(device, C.byref(f), C.byref(t)), "device_memory")
    return f.value, t.value


def stream_copy(nbytes=1 << 31, reps=20, device=0):
    """Measured copy bandwidth in GB/s (2 x nbytes per copy): one double per
    lane, and 16 bytes per lane on a grid-stride loop (cgli_stream_copy)."""
    ms = (C.c_double * 2)()
    _check(lib().cgli_stream_copy(device, nbytes, reps, ms), "stream_copy")
    return {"bytes": nbytes, "reps": reps, "ms": [ms[0], ms[1]],
            "GBps_one_double_per_lane": 2 * nbytes / (ms[0] * 1e-3) / 1e9,
            "GBps_16B_per_lane": 2 * nbytes / (ms[1] * 1e-3) / 1e9}


def reference_config(nx, ny, nz):
    c = Config()
    lib().cgli_config_reference(C.byref(c), nx, ny, nz)
    return c


def channel(re_tau=180.0, **kw):
    """Channel parameters (cgli_channel) with defaults for re_tau, overridden by kw."""
    ch = Channel()
    lib().cgli_channel_defaults(C.byref(ch), re_tau)
    for k, v in kw.items():
        setattr(ch, k, v)
    if ch.precond_e <= 0:
        ch.precond_e = ch.mach * ch.mach   # the library's default, made explicit for callers
    return ch


def channel_config(ch, nx, ny, nz):
    c = Config()
    _check(lib().cgli_channel_config(C.byref(ch), nx, ny, nz, C.byref(c)), "channel_config")
    return c


def _check(code, what):
    if code != 0:
        raise CgliError(code, what)


class Context:
    """One slab resident on one device (cgli_ctx)."""

    def __init__(self, cfg, device=0):
        self.cfg = cfg
        self.shape = (cfg.nx + 6, cfg.ny + 4, cfg.nz + 4)
        self.n = int(np.prod(self.shape))
        h = C.c_void_p()
        _check(lib().cgli_create(C.byref(cfg), device, C.byref(h)), "cgli_create")
        self.h = h

    def close(self):
        if self.h:
            lib().cgli_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- fields
    def set(self, fid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != self.n:
            raise ValueError("field %d: expected %d values, got %d" % (fid, self.n, a.size))
        _check(lib().cgli_set_field(self.h, fid, a.ctypes.data_as(C.POINTER(C.c_double))), "set_field")

    def get(self, fid):
        a = np.empty(self.shape)
        _check(lib().cgli_get_field(self.h, fid, a.ctypes.data_as(C.POINTER(C.c_double))), "get_field")
        return a

    def set_many(self, fid, arrs):
        for q, a in enumerate(arrs):
            self.set(fid + q, a)

    def get_many(self, fid, count):
        return [self.get(fid + q) for q in range(count)]

    def sync(self):
        _check(lib().cgli_synchronize(self.h), "synchronize")

    # -- hot path
    def flux_x(self, Ep, e):
        _check(lib().cgli_flux_x(self.h, Ep, e), "flux_x")

    def flux_y(self, Ep, e):
        _check(lib().cgli_flux_y(self.h, Ep, e), "flux_y")

    def flux_z(self, Ep, e):
        _check(lib().cgli_flux_z(self.h, Ep, e), "flux_z")

    DISSIPATION = {"none": 0, "ep": 1, "sensor": 2}

    def set_dissipation_scaling(self, mode):
        """Roe dissipation scaling: 'none' (the live reference), 'ep' (the flux
        calls' Ep argument, Flux_X~~.cpp:552-556) or 'sensor' (EpX/EpY/EpZ)."""
        _check(lib().cgli_set_dissipation_scaling(self.h, self.DISSIPATION.get(mode, mode)),
               "set_dissipation_scaling")

    def viscous(self):
        _check(lib().cgli_viscous(self.h), "viscous")

    def last_step_fp32(self):
        """FP32-residual inner iterations of the last channel step."""
        n = C.c_int()
        _check(lib().cgli_channel_last_step(self.h, C.byref(n)), "channel_last_step")
        return n.value

    PROFILE_NAMES = ["U", "V", "W", "T", "UU", "VV", "WW", "TT", "UV", "UT", "VT"]

    def profiles(self):
        """Statistic.cpp:130-164 quantities averaged over x and z: (11, ny)."""
        out = np.zeros(11 * self.cfg.ny)
        _check(lib().cgli_profiles(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "profiles")
        return out.reshape(11, self.cfg.ny)

    def filter(self, Roe_criterion=0.007, E_high=1.0, E_low=0.1):
        """Filter.cpp:12-1720: EpX/EpY/EpZ from U_, J and Y_point."""
        _check(lib().cgli_filter(self.h, C.c_double(Roe_criterion), C.c_double(E_high), C.c_double(E_low)),
               "filter")

    def dual_time_stepping(self, deltaT, e, sweeps=1, write_residual=False):
        er = (C.c_double * 5)()
        _check(lib().cgli_dual_time_stepping(self.h, deltaT, e, er, sweeps, int(write_residual)),
               "dual_time_stepping")
        return np.array(list(er))

    def dual_time_stepping_async(self, deltaT, e, sweeps=1, write_residual=False):
        """Queue one dual-time iteration; dual_time_stepping_wait returns its norms."""
        _check(lib().cgli_dual_time_stepping_async(self.h, deltaT, e, sweeps, 1 if write_residual else 0),
               "dual_time_stepping_async")

    def dual_time_stepping_wait(self):
        er = (C.c_double * 5)()
        _check(lib().cgli_dual_time_stepping_wait(self.h, er), "dual_time_stepping_wait")
        return np.array(list(er))

    # -- HIP graphs of a sequence of calls on this context
    def graph_begin(self):
        _check(lib().cgli_graph_begin(self.h), "graph_begin")

    def graph_end(self):
        g = C.c_void_p()
        _check(lib().cgli_graph_end(self.h, C.byref(g)), "graph_end")
        return Graph(g, self)

    def set_graph(self, on=True):
        """Replay channel_step's inner iterations from captured graphs."""
        _check(lib().cgli_channel_set_graph(self.h, 1 if on else 0), "channel_set_graph")

    def filter_energies(self):
        """E_h and E_2h of the last filter call (field layout)."""
        Eh, E2h = np.zeros(self.shape), np.zeros(self.shape)
        _check(lib().cgli_filter_energies(self.h, Eh.ctypes.data_as(C.POINTER(C.c_double)),
                                          E2h.ctypes.data_as(C.POINTER(C.c_double))), "filter_energies")
        return Eh, E2h

    def dplusgs(self, deltaT, e, sweeps=4):
        er = (C.c_double * 5)()
        _check(lib().cgli_dplusgs(self.h, deltaT, er, e, sweeps), "dplusgs")
        return np.array(list(er))

    def dplusgs_noalloc(self, deltaT, e, sweeps=4):
        """cgli_dplusgs_noalloc: the same, allocating nothing (fields must exist)"""
        er = (C.c_double * 5)()
        _check(lib().cgli_dplusgs_noalloc(self.h, deltaT, er, e, sweeps), "dplusgs_noalloc")
        return np.array(list(er))

    def hyperplane(self):
        c = self.cfg
        rows = c.nx + 2 + c.ny + 1 + c.nz + 1 + 8
        cells = c.nx * c.ny * c.nz
        N = np.zeros((rows, 2), np.int32)
        I = np.zeros((cells + 1, 3), np.int32)
        _check(lib().cgli_hyperplane(self.h, N.ctypes.data_as(C.POINTER(C.c_int)),
                                     I.ctypes.data_as(C.POINTER(C.c_int))), "hyperplane")
        return N, I

    def bc_y(self, T_low, T_high, heated_lo=-(2 ** 31), heated_hi=2 ** 31 - 1):
        _check(lib().cgli_bc_y(self.h, T_low, T_high, heated_lo, heated_hi), "bc_y")

    def bc_z(self):
        _check(lib().cgli_bc_z(self.h), "bc_z")

    def bc_x_periodic(self):
        _check(lib().cgli_bc_x_periodic(self.h), "bc_x_periodic")

    # -- restart file (Output.cpp:167-284, Initial_condition.cpp:112-219)
    def write_restart(self, path):
        """Write U*J and Uq*J of the interior in the reference's LES%05d.bin layout."""
        _check(lib().cgli_write_restart(self.h, os.fsencode(path)), "write_restart")

    def read_restart(self, path):
        """Read such a file: U = Q/J, Uq = Q_old/J, U_ = U on the interior."""
        _check(lib().cgli_read_restart(self.h, os.fsencode(path)), "read_restart")

    # -- channel workload
    def channel_init(self, ch):
        dt = C.c_double()
        _check(lib().cgli_channel_init(self.h, C.byref(ch), C.byref(dt)), "channel_init")
        return dt.value

    def channel_ghosts(self):
        _check(lib().cgli_channel_ghosts(self.h), "channel_ghosts")

    def channel_step(self, deltaT, tol=1e-6, max_inner=100, sweeps=1):
        it = C.c_int()
        e1 = (C.c_double * 5)()
        en = (C.c_double * 5)()
        _check(lib().cgli_channel_step(self.h, deltaT, tol, max_inner, sweeps, C.byref(it), e1, en),
               "channel_step")
        return it.value, np.array(list(e1)), np.array(list(en))

    # -- stopping rule and initial guess of channel_step
    def set_rule(self, rule):
        """0: per-step drop er/er(1); 1: running maximum er/E (main.cpp:287-304)."""
        _check(lib().cgli_channel_set_rule(self.h, rule), "channel_set_rule")

    def rule_state(self, E=None):
        """Return the running maxima E[5]; overwrite them first when E is given."""
        buf = (C.c_double * 5)(*(E if E is not None else [0.0] * 5))
        _check(lib().cgli_channel_rule_state(self.h, buf, 1 if E is not None else 0), "channel_rule_state")
        return np.array(list(buf))

    GUESSES = {"previous": 0, "extrapolate": 1, "network": 2, "quadratic": 3}

    def set_guess(self, guess):
        """'previous' (U^n), 'extrapolate' (2U^n - U^{n-1}), 'quadratic' or 'network'."""
        _check(lib().cgli_channel_set_guess(self.h, self.GUESSES.get(guess, guess)), "channel_set_guess")

    def attach_net(self, net):
        _check(lib().cgli_channel_attach_net(self.h, net.h if net is not None else None), "channel_attach_net")
        self._net = net   # keep it alive while attached

    def guess_info(self):
        used = C.c_int()
        rn = (C.c_double * 5)()
        re = (C.c_double * 5)()
        _check(lib().cgli_channel_guess_info(self.h, C.byref(used), rn, re), "channel_guess_info")
        return bool(used.value), np.array(list(rn)), np.array(list(re))

    # -- timing (HIP events around each launch class)
    TIMING_CLASSES = ["flux_x", "flux_y", "flux_z", "viscous", "rhs_diag", "lusgs_fwd", "lusgs_bwd",
                      "update_norm", "split_coef", "filter", "march_scatter", "march_gather"]

    def set_lusgs_mode(self, mode):
        """0: one launch per hyperplane, 1: tiled wavefront, 3: the wavefront
        on march-ordered coefficient records, 4: the wavefront on
        march-ordered state computing its own coefficients (default)."""
        _check(lib().cgli_set_lusgs_mode(self.h, mode), "set_lusgs_mode")

    def set_sweep_waves(self, waves):
        """Schedule 4's waves per tile: 1 (one wave), 2-4 (a chain wave with
        1-3 helper waves computing the coefficients ahead), 0 (by tile
        count, the default)."""
        _check(lib().cgli_set_sweep_waves(self.h, waves), "set_sweep_waves")

    def set_spin_limit(self, n):
        _check(lib().cgli_set_spin_limit(self.h, n), "set_spin_limit")

    def timing(self, on=True, classes=None):
        """Time kernel classes with HIP events (all, or the named ones), reset."""
        mask = 0xFFFFFFFF if classes is None else sum(1 << self.TIMING_CLASSES.index(c) for c in classes)
        _check(lib().cgli_timing_select(self.h, mask), "timing_select")
        _check(lib().cgli_timing_enable(self.h, int(on)), "timing_enable")
        _check(lib().cgli_timing_reset(self.h), "timing_reset")

    def timing_report(self):
        out = {}
        for i, name in enumerate(self.TIMING_CLASSES):
            t = C.c_double()
            n = C.c_long()
            _check(lib().cgli_timing_get(self.h, i, C.byref(t), C.byref(n)), "timing_get")
            out[name] = (t.value, n.value)
        return out


class Graph:
    """A captured sequence of calls on one context (cgli_graph)."""

    def __init__(self, handle, ctx):
        self.h = handle
        self._ctx = ctx   # keep the context alive while the graph exists

    def launch(self):
        _check(lib().cgli_graph_launch(self.h), "graph_launch")

    def close(self):
        if self.h:
            lib().cgli_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Net:
    """The learned initial guess's network (cgli_net)."""

    NPARAM = 83056

    def __init__(self, seed=1, device=0):
        h = C.c_void_p()
        _check(lib().cgli_net_create(device, seed, C.byref(h)), "cgli_net_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().cgli_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, new=None):
        """FP32 master parameters (padded layout); set them first when new is given."""
        if new is not None:
            a = np.ascontiguousarray(new, dtype=np.float32)
            assert a.size == self.NPARAM
            _check(lib().cgli_net_params(self.h, _fp(a), 1), "net_params")
            return a.copy()
        a = np.zeros(self.NPARAM, np.float32)
        _check(lib().cgli_net_params(self.h, _fp(a), 0), "net_params")
        return a

    def eval(self, x, t=None, mask=None, periodic=False, want_grad=True):
        """x[B,D,H,W,15] -> y[B,D,H,W,5]; with t and mask also (loss, grad)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B, D, H, W, c = x.shape
        assert c == 15
        y = np.zeros((B, D, H, W, 5), np.float32)
        loss = C.c_double()
        g = np.zeros(self.NPARAM, np.float32) if (t is not None and want_grad) else None
        tt = np.ascontiguousarray(t, dtype=np.float32) if t is not None else None
        mm = np.ascontiguousarray(mask, dtype=np.uint8) if mask is not None else None
        _check(lib().cgli_net_eval(self.h, B, D, H, W, int(periodic), _fp(x),
                                   _fp(tt) if tt is not None else None,
                                   mm.ctypes.data_as(C.POINTER(C.c_ubyte)) if mm is not None else None,
                                   _fp(y), C.byref(loss) if t is not None else None,
                                   _fp(g) if g is not None else None), "net_eval")
        if t is None:
            return y
        return y, loss.value, g

    def adam(self, lr=1e-3):
        _check(lib().cgli_net_adam(self.h, lr), "net_adam")

    def ring(self, capacity, nx, ny, nz):
        _check(lib().cgli_net_ring(self.h, capacity, nx, ny, nz), "net_ring")
        self.grid = (nx, ny, nz)

    def push(self, dq):
        a = np.ascontiguousarray(dq, dtype=np.float32)
        _check(lib().cgli_net_push(self.h, _fp(a)), "net_push")

    def rewind(self, nrec):
        _check(lib().cgli_net_rewind(self.h, nrec), "net_rewind")

    def records(self):
        n = C.c_long()
        _check(lib().cgli_net_records(self.h, C.byref(n)), "net_records")
        return n.value

    def record_scales(self, rec):
        """inverse scales of record `rec`: (1/s1, 1/sd, 1/s2), five each."""
        a = np.zeros(15, np.float32)
        _check(lib().cgli_net_record_scales(self.h, rec, _fp(a)), "net_record_scales")
        return a[:5], a[5:10], a[10:]

    def train(self, steps, batch=8, crop=(32, 32, 32), lr=1e-3, seed=1):
        loss = np.zeros(steps, np.float32)
        _check(lib().cgli_net_train(self.h, steps, batch, crop[0], crop[1], crop[2], lr, seed, _fp(loss)),
               "net_train")
        return loss

    def predict(self):
        """(x, y) of a whole-grid pass on the latest three records."""
        nx, ny, nz = self.grid
        x = np.zeros((nx, ny, nz, 15), np.float32)
        y = np.zeros((nx, ny, nz, 5), np.float32)
        _check(lib().cgli_net_predict(self.h, _fp(x), _fp(y)), "net_predict")
        return x, y

    def timing(self):
        pm, tm = C.c_double(), C.c_double()
        pn, tn = C.c_long(), C.c_long()
        _check(lib().cgli_net_timing(self.h, C.byref(pm), C.byref(pn), C.byref(tm), C.byref(tn)), "net_timing")
        return {"predict_ms": pm.value, "predicts": pn.value, "train_ms": tm.value, "train_steps": tn.value}
