"""Test-side access to the CPU oracle and, where built, the reference library.

TEST INFRASTRUCTURE ONLY.  Loads ``oracle/liboracle.so`` (the C restatement
of the reference routines) and ``oracle/_ref/libref_*.so`` (the reference's
own routines, built from its sources by ``oracle/ref/build_ref.sh`` where the
reference tree is available) through ctypes, and lays out fields exactly like
the reference's ``double (*)[Y_m][Z_m]`` arrays.
"""
import ctypes as C
import glob
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class OrcCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nz", "first", "last", "gstart", "gend0", "nx_global",
        "nx_inlet", "nx_outlet", "sponge", "xi_mode", "x_periodic", "precision")] + [(n, C.c_double) for n in (
        "dxi", "deta", "dzeta", "high", "K", "R", "Cv", "mu_L", "Pr_L", "suth_T",
        "suth_S", "rho_ref", "grav", "f_const", "f_dplusgs", "E_ref")]


def load_oracle():
    path = os.path.join(ORACLE_DIR, "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    lib = C.CDLL(path)
    lib.orc_fill.argtypes = [DP, C.c_long, C.c_uint64, C.c_double, C.c_double]
    lib.orc_fill_state.argtypes = [C.c_long, C.c_uint64, DP, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.POINTER(DP)]
    lib.orc_cfg_reference.argtypes = [C.POINTER(OrcCfg), C.c_int, C.c_int, C.c_int]
    return lib


def ref_libraries():
    return sorted(glob.glob(os.path.join(REF_DIR, "libref_*.so")))


def ref_cfg_from(lib):
    """(X_out, Y_out, Z_out, Y_m, Z_m, np, nx_inlet, nx_outlet), deltas."""
    dims = (C.c_int * 8)()
    lib.ref_dims(dims)
    d = (C.c_double * 3)()
    lib.ref_deltas(d)
    return tuple(dims), tuple(d)


def ptrs(arrs):
    """double** from a list of contiguous float64 arrays."""
    return (DP * len(arrs))(*[a.ctypes.data_as(DP) for a in arrs])


def dp(a):
    return a.ctypes.data_as(DP)


class Slab:
    """All fields of one slab in the reference layout [X_np][Y_m][Z_m]."""

    def __init__(self, nx, ny, nz):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nx + 6, ny + 4, nz + 4)
        self.n = int(np.prod(self.shape))

    def zeros(self, count=None):
        if count is None:
            return np.zeros(self.shape)
        return [np.zeros(self.shape) for _ in range(count)]

    # reference index ranges (Resolution.h:15-29)
    @property
    def gend(self):
        return self.nx + 2

    @property
    def NY(self):
        return self.ny + 1

    @property
    def NZ(self):
        return self.nz + 1


def fill(orc, shape, seed, lo, hi):
    a = np.empty(shape)
    orc.orc_fill(dp(a), a.size, seed, lo, hi)
    return a


def make_inputs(orc, slab, seed, K=1.4, dimensional=True):
    """Seeded valid state, metrics and time levels for one slab.

    dimensional=True draws air at reference conditions (the reference's SI
    units); otherwise a nondimensional state (rho~1, p~1/(gamma M^2))."""
    sh = slab.shape
    f = {}
    f["J"] = fill(orc, sh, seed + 1, 0.7, 1.3)
    # cell metrics: xi_x, eta_y, zeta_z dominant, cross terms small
    names = ["xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"]
    diag = {"xidx", "etdy", "ztdz"}
    for tag, base in (("", 10), ("_u", 30), ("_v", 50), ("_w", 70)):
        for q, nme in enumerate(names):
            lo, hi = (0.6, 1.4) if nme in diag else (-0.25, 0.25)
            f[nme + tag] = fill(orc, sh, seed + base + q, lo, hi)
        if tag:
            f["J" + tag] = fill(orc, sh, seed + base + 9, 0.7, 1.3)
    # grid y coordinate (Filter input): strictly increasing in j with a
    # seeded jitter, so one-sided spacings differ from cell to cell
    jj = np.arange(sh[1], dtype=float)[None, :, None]
    f["Y_point"] = 0.1 * jj + 0.03 * fill(orc, sh, seed + 90, -1.0, 1.0)
    if dimensional:
        r0, r1, ua, p0, p1 = 1.05, 1.35, 30.0, 95000.0, 107000.0
    else:
        r0, r1, ua, p0, p1 = 0.9, 1.1, 0.5, 2.6, 3.1
    for lvl, base in (("U_", 100), ("U", 200), ("Uq", 300)):
        arrs = [np.empty(sh) for _ in range(5)]
        orc.orc_fill_state(slab.n, seed + base, dp(f["J"]), K, r0, r1, ua, p0, p1, ptrs(arrs))
        f[lvl] = arrs
    return f


def met9(f, tag=""):
    names = ["xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"]
    return [f[n + tag] for n in names]


def ref_cfg(orc, nx, ny, nz, deltas=None):
    c = OrcCfg()
    orc.orc_cfg_reference(C.byref(c), nx, ny, nz)
    if deltas is not None:
        c.dxi, c.deta, c.dzeta = deltas
    return c


def hyperplane_sizes(slab):
    rows = slab.gend + slab.NY + slab.NZ + 8
    cells = slab.nx * slab.ny * slab.nz
    return rows, cells


def rel_maxnorm(a, b):
    """max|a-b| / max|b| over the arrays (the parity metric)."""
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0:
        return float(num)
    return float(num / den)


# ---------------------------------------------------------------------------
# one inner iteration of the reference call sequence (main.cpp:213-274),
# run either through the reference library or through the oracle
# ---------------------------------------------------------------------------
class Case:
    """A slab configuration matching one reference build."""

    def __init__(self, X, Y, Z, NP, nin, nout, myid, deltas):
        self.X, self.Y, self.Z, self.NP, self.nin, self.nout, self.myid = X, Y, Z, NP, nin, nout, myid
        self.gc = X // NP
        self.slab = Slab(self.gc, Y, Z)
        self.deltas = deltas
        self.xnp = self.gc + 6
        self.overrides = {}

    @property
    def tag(self):
        return "%dx%dx%d_np%d_in%d_out%d" % (self.X, self.Y, self.Z, self.NP, self.nin, self.nout)

    def cfg(self, orc, xi_mode=0):
        c = ref_cfg(orc, self.gc, self.Y, self.Z, self.deltas)
        c.nx_inlet, c.nx_outlet = self.nin, self.nout
        c.first = int(self.myid == 0)
        c.last = int(self.myid == self.NP - 1)
        c.gstart = self.myid * self.gc
        c.gend0 = c.gstart + self.gc - 1
        c.nx_global = self.X
        c.xi_mode = xi_mode
        for k, v in self.overrides.items():
            setattr(c, k, v)
        return c


def run_chain(backend, lib, case, f, orc=None, Ep=1.0, e=1e-4, deltaT=2e-4, dts=True, sweeps=2):
    """Fluxes, viscous terms and (optionally) one Dual_time_stepping call.
    Returns a dict of output arrays."""
    slab = case.slab
    out = {}
    U_ = [a.copy() for a in f["U_"]]
    MR, ML = slab.zeros(5), slab.zeros(5)
    for d, tag in enumerate(("_u", "_v", "_w")):
        F = slab.zeros(5)
        if backend == "ref":
            lib.ref_flux(d, case.xnp, case.myid, C.c_double(Ep), C.c_double(e), ptrs(U_), ptrs(MR), ptrs(ML),
                         dp(f["J"]), dp(f["J" + tag]), ptrs(met9(f, tag)), ptrs(F), dp(slab.zeros()))
        else:
            cfg = case.cfg(lib)
            fn = (lib.orc_flux_x, lib.orc_flux_y, lib.orc_flux_z)[d]
            fn(C.byref(cfg), C.c_double(Ep), C.c_double(e), ptrs(U_), ptrs(MR), ptrs(ML), dp(f["J"]),
               dp(f["J" + tag]), ptrs(met9(f, tag)), ptrs(F))
        out["inF" + "xyz"[d]] = F
    vF = slab.zeros(4)
    S = [slab.zeros(5) for _ in range(6)]
    if backend == "ref":
        lib.ref_viscous(case.xnp, case.myid, ptrs(U_), ptrs(vF), dp(f["J"]), ptrs(met9(f)), dp(f["J_u"]),
                        ptrs(met9(f, "_u")), dp(f["J_v"]), ptrs(met9(f, "_v")), dp(f["J_w"]), ptrs(met9(f, "_w")),
                        *[ptrs(s) for s in S])
    else:
        cfg = case.cfg(lib)
        lib.orc_viscous(C.byref(cfg), ptrs(U_), ptrs(vF), dp(f["J"]), dp(f["J_u"]), ptrs(met9(f, "_u")),
                        dp(f["J_v"]), ptrs(met9(f, "_v")), dp(f["J_w"]), ptrs(met9(f, "_w")),
                        ptrs(S[1]), ptrs(S[3]), ptrs(S[5]))
    out["vF"] = vF
    if not dts:
        return out
    rows, cells = hyperplane_sizes(slab)
    N = np.zeros((rows, 2), np.int32)
    I = np.zeros((3 * cells + 1, 3), np.int32)
    NR, NL, Res = slab.zeros(5), slab.zeros(5), slab.zeros(5)
    MR = slab.zeros(5)
    er = (C.c_double * 5)()
    if backend == "ref":
        lib.ref_hyperplane(case.xnp, case.myid, N.ctypes.data_as(IP), I.ctypes.data_as(IP))
        lib.ref_dts(case.xnp, case.myid, C.c_double(deltaT), C.c_double(e), er, N.ctypes.data_as(IP),
                    I.ctypes.data_as(IP), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]), dp(f["J"]), ptrs(met9(f)),
                    ptrs(out["inFx"]), ptrs(out["inFy"]), ptrs(out["inFz"]), ptrs(vF), ptrs(MR), ptrs(NR),
                    ptrs(NL), ptrs(Res), dp(f["etdy_v"]), dp(f["J_v"]))
    else:
        cfg = case.cfg(lib)
        lib.orc_hyperplane(C.byref(cfg), N.ctypes.data_as(IP), I.ctypes.data_as(IP))
        lib.orc_dual_time_stepping(C.byref(cfg), C.c_double(deltaT), C.c_double(e), er, N.ctypes.data_as(IP),
                                   I.ctypes.data_as(IP), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]), dp(f["J"]),
                                   ptrs(met9(f)), ptrs(out["inFx"]), ptrs(out["inFy"]), ptrs(out["inFz"]),
                                   ptrs(vF), ptrs(MR), ptrs(NR), ptrs(NL), ptrs(Res), dp(f["etdy_v"]),
                                   dp(f["J_v"]), sweeps)
    out["Nijk"], out["ijk"] = N, I
    out["U_"], out["NR"], out["MR"] = U_, NR, MR
    out["er"] = np.array(list(er))
    return out


FILTER_CRITERIA = (0.007, 0.024)   # the reference default, and the median E_h/E_2h of the seeded data


def run_filter(backend, lib, case, f, criterion, E_high=1.0, E_low=0.1):
    """EpX, EpY, EpZ of Filter.cpp on the seeded inputs."""
    slab = case.slab
    Ep = [np.full(slab.shape, -7.0) for _ in range(3)]
    if backend == "ref":
        lib.ref_filter(case.xnp, case.myid, C.c_double(criterion), C.c_double(E_high), C.c_double(E_low),
                       ptrs(f["U_"]), dp(f["Y_point"]), dp(Ep[0]), dp(Ep[1]), dp(Ep[2]), dp(f["J"]))
    else:
        cfg = case.cfg(lib)
        lib.orc_filter(C.byref(cfg), C.c_double(criterion), C.c_double(E_high), C.c_double(E_low), ptrs(f["U_"]),
                       dp(f["Y_point"]), dp(Ep[0]), dp(Ep[1]), dp(Ep[2]), dp(f["J"]))
    return Ep


def regions(slab):
    """Index windows each output is defined on (reference loop ranges)."""
    g, NY, NZ = slab.gend, slab.NY, slab.NZ
    return {
        "EpX": (slice(2, g + 1), slice(1, NY), slice(1, NZ)),         # Filter.cpp:1564-1583
        "EpY": (slice(2, g), slice(1, NY + 1), slice(1, NZ)),         # Filter.cpp:1593-1614
        "EpZ": (slice(2, g), slice(1, NY), slice(1, NZ + 1)),         # Filter.cpp:1624-1643
        "inFx": (slice(2, g + 1), slice(1, NY), slice(1, NZ)),        # Flux_X.cpp:229-245
        "inFy": (slice(2, g), slice(1, NY + 1), slice(1, NZ)),        # Flux_Y.cpp:272-289
        "inFz": (slice(2, g), slice(1, NY), slice(1, NZ + 1)),        # Flux_Z.cpp:230-248
        "vF": (slice(3, g + 1), slice(2, NY + 1), slice(2, NZ + 1)),  # Viscous_terms.cpp:875-878
        "cell": (slice(3, g + 1), slice(2, NY + 1), slice(2, NZ + 1)),
    }
