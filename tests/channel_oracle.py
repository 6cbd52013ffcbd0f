"""The channel workload on the CPU: independent numpy setup + oracle steps.

TEST INFRASTRUCTURE ONLY.  Builds the grid metrics and the initial state of
the compressible channel from the mapping's closed form (numpy), and runs
physical steps through the oracle's routines in the library driver's order
(ghost fills y, z, x; fluxes; viscous terms; dual-time LU-SGS; time shift).
"""
import ctypes as C

import numpy as np

from oracle_harness import OrcCfg, dp, load_oracle, ptrs

MET = ["xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"]


def splitmix_stream(seed, n):
    """First n outputs of splitmix64 from state `seed`, as uniforms in [0,1)."""
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + g * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def channel_cfg(nx, ny, nz, Lx=4 * np.pi, Lz=2 * np.pi, mach=0.5, re_tau=180.0, re_c=3300.0, Pr=0.72,
                suth_S=110.4 / 288.15, xi_mode=1):
    orc = load_oracle()
    c = OrcCfg()
    orc.orc_cfg_reference(C.byref(c), nx, ny, nz)
    c.first = c.last = 0
    c.x_periodic = 1
    c.sponge = 0
    c.xi_mode = xi_mode
    c.dxi, c.deta, c.dzeta = Lx / nx, 1.0 / ny, Lz / nz
    c.high = 2.0
    c.K = 1.4
    c.R = 1.0 / (c.K * mach * mach)
    c.Cv = c.R / (c.K - 1)
    c.mu_L = 1.0 / re_c
    c.Pr_L = Pr
    c.suth_T = 1.0
    c.suth_S = suth_S
    c.rho_ref = 1.0
    c.grav = 0.0
    c.f_const = (re_tau / re_c) ** 2
    c.f_dplusgs = c.f_const
    c.E_ref = 0.0
    return c


def metrics(c, beta, skew, Lz=2 * np.pi):
    """J and the 9 metrics of the cell, x-face, y-face and z-face sets."""
    Ym, Zm = c.ny + 4, c.nz + 4
    shape = (c.nx + 6, Ym, Zm)
    j = np.arange(Ym, dtype=float)[:, None]
    k = np.arange(Zm, dtype=float)[None, :]
    out = {}
    for tag, eo, zo in (("", 1.5, 1.5), ("_u", 0.5, 0.5), ("_v", 1.0, 0.5), ("_w", 0.5, 1.0)):
        eta = (j - eo) * c.deta + 0 * k
        zeta = (k - zo) * c.dzeta + 0 * j
        t = np.tanh(beta * (1 - 2 * eta))
        y = -t / np.tanh(beta)
        ye = 2 * beta * (1 - t * t) / np.tanh(beta)
        xe = skew * 2 * np.pi * np.cos(2 * np.pi * y) * ye * np.sin(2 * np.pi * zeta / Lz)
        xz = skew * np.sin(2 * np.pi * y) * (2 * np.pi / Lz) * np.cos(2 * np.pi * zeta / Lz)
        # M = [[1, xe, xz], [0, ye, 0], [0, 0, 1]]; inverse in closed form
        m = {"xidx": 1.0 + 0 * ye, "xidy": -xe / ye, "xidz": -xz, "etdx": 0 * ye, "etdy": 1.0 / ye,
             "etdz": 0 * ye, "ztdx": 0 * ye, "ztdy": 0 * ye, "ztdz": 1.0 + 0 * ye}
        for n in MET:
            out[n + tag] = np.broadcast_to(m[n], shape).copy()
        out["J" + tag] = np.broadcast_to(1.0 / ye, shape).copy()
        if tag == "":
            out["y"] = y
    # Filter input: cell-centre y, ghost rows mirrored about the walls y = -1
    # and y = 1 (Grid.cpp:88-89 mirrors about its walls y = 0 and y = high)
    yc = out["y"][:, 0].copy()
    NY = c.ny + 1
    yc[1], yc[0] = -2.0 - yc[2], -2.0 - yc[3]
    yc[NY + 1], yc[NY + 2] = 2.0 - yc[NY], 2.0 - yc[NY - 1]
    out["Y_point"] = np.broadcast_to(yc[None, :, None], shape).copy()
    return out


def delta_t(c, f, U, cfl):
    """Physical step at acoustic CFL number cfl: cfl over the largest
    sum_d (|U_d| + c |grad d|) / delta_d of the interior cells, with U_d the
    contravariant velocity and c the speed of sound of the state U (Q/J)."""
    sl = (slice(3, c.nx + 3), slice(2, c.ny + 2), slice(2, c.nz + 2))
    J = f["J"][sl]
    rho = U[0][sl] * J
    u, v, w = U[1][sl] / U[0][sl], U[2][sl] / U[0][sl], U[3][sl] / U[0][sl]
    p = (U[4][sl] * J - 0.5 * rho * (u * u + v * v + w * w)) * (c.K - 1)
    cs = np.sqrt(c.K * p / rho)
    lam = 0.0
    for d, dd in enumerate((c.dxi, c.deta, c.dzeta)):
        m = [f[MET[3 * d + t]][sl] for t in range(3)]
        Ud = m[0] * u + m[1] * v + m[2] * w
        lam = lam + (np.abs(Ud) + cs * np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2)) / dd
    return cfl / float(np.max(lam))


def initial_state(c, f, seed=20261203, noise=0.1):
    """Laminar u = 1 - y^2 plus uniform noise; rho = T = 1, p = R."""
    nx, ny, nz = c.nx, c.ny, c.nz
    n = nx * ny * nz
    shape = f["J"].shape
    nzs = [noise * (2 * splitmix_stream(seed * 3 + q, n) - 1) for q in range(3)]
    # interior cells drawn with i fastest, then j, then k
    nzs = [a.reshape(nz, ny, nx).transpose(2, 1, 0) for a in nzs]
    sl = (slice(3, nx + 3), slice(2, ny + 2), slice(2, nz + 2))
    y = f["y"][2:ny + 2, 2:nz + 2][None, :, :]
    J = f["J"][sl]
    u = (1 - y * y) + nzs[0]
    v, w = nzs[1], nzs[2]
    rho = 1.0
    p0 = c.R
    U = [np.zeros(shape) for _ in range(5)]
    U[0][sl] = rho / J
    U[1][sl] = rho * u / J
    U[2][sl] = rho * v / J
    U[3][sl] = rho * w / J
    U[4][sl] = (p0 / (c.K - 1) + 0.5 * rho * (u * u + v * v + w * w)) / J
    return U


class CpuChannel:
    def __init__(self, nx, ny, nz, beta=2.2, skew=0.05, e=None, **kw):
        self.orc = load_oracle()
        # preconditioning floor passed to the flux and dual-time routines
        self.e = e if e is not None else kw.get("mach", 0.5) ** 2
        self.c = channel_cfg(nx, ny, nz, **kw)
        self.f = metrics(self.c, beta, skew)
        U = initial_state(self.c, self.f)
        self.f["U_"] = U
        self.f["U"] = [a.copy() for a in U]
        self.f["Uq"] = [a.copy() for a in U]
        self.shape = U[0].shape
        rows = nx + 2 + ny + 1 + nz + 1 + 8
        self.N = np.zeros((rows, 2), np.int32)
        self.I = np.zeros((nx * ny * nz + 1, 3), np.int32)
        self.orc.orc_hyperplane(C.byref(self.c), self.N.ctypes.data_as(C.POINTER(C.c_int)),
                                self.I.ctypes.data_as(C.POINTER(C.c_int)))
        z = lambda n=5: [np.zeros(self.shape) for _ in range(n)]
        self.MR, self.ML, self.Fx, self.Fy, self.Fz, self.NL, self.NR, self.Res = (z() for _ in range(8))
        self.LL, self.MLv, self.NLv = z(), z(), z()
        self.vF = z(4)

    def inner(self, dt, sweeps=1):
        o, c, f = self.orc, self.c, self.f
        U_ = f["U_"]
        o.orc_bc_y(C.byref(c), ptrs(U_), dp(f["J"]), C.c_double(1.0), C.c_double(1.0), -(2 ** 31), 2 ** 31 - 1)
        o.orc_bc_z(C.byref(c), ptrs(U_))
        o.orc_bc_x_periodic(C.byref(c), ptrs(U_))
        for fn, tag, F in ((o.orc_flux_x, "_u", self.Fx), (o.orc_flux_y, "_v", self.Fy), (o.orc_flux_z, "_w", self.Fz)):
            fn(C.byref(c), C.c_double(1.0), C.c_double(self.e), ptrs(U_), ptrs(self.MR), ptrs(self.ML), dp(f["J"]),
               dp(f["J" + tag]), ptrs([f[n + tag] for n in MET]), ptrs(F))
        o.orc_viscous(C.byref(c), ptrs(U_), ptrs(self.vF), dp(f["J"]), dp(f["J_u"]), ptrs([f[n + "_u"] for n in MET]),
                      dp(f["J_v"]), ptrs([f[n + "_v"] for n in MET]), dp(f["J_w"]), ptrs([f[n + "_w"] for n in MET]),
                      ptrs(self.LL), ptrs(self.MLv), ptrs(self.NLv))
        er = (C.c_double * 5)()
        o.orc_dual_time_stepping(C.byref(c), C.c_double(dt), C.c_double(self.e), er,
                                 self.N.ctypes.data_as(C.POINTER(C.c_int)), self.I.ctypes.data_as(C.POINTER(C.c_int)),
                                 ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]), dp(f["J"]), ptrs([f[n] for n in MET]),
                                 ptrs(self.Fx), ptrs(self.Fy), ptrs(self.Fz), ptrs(self.vF), ptrs(self.MR),
                                 ptrs(self.NR), ptrs(self.NL), ptrs(self.Res), dp(f["etdy_v"]), dp(f["J_v"]), sweeps)
        return np.array(list(er))

    def shift(self):
        for q in range(5):
            self.f["Uq"][q][:] = self.f["U"][q]
            self.f["U"][q][:] = self.f["U_"][q]
