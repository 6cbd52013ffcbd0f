"""The oracle's implicit update against the mathematics of its operator.

The LU-SGS sweeps of orc_dual_time_stepping are bit-identical to the
reference's Dual_time_stepping (test_oracle.py).  The reference's Jacobi
variant DPLUSGS reads the never-assigned j_/k_ (DPLUSGS.cpp:997-1052), so no
reference output exists for it; orc_dplusgs is pinned here instead to the
operator the pinned LU-SGS factorises:

* the dense 5x5-block operator (tests/implicit_dense.py, restated from the
  split formulas DPLUSGS.cpp:585-990) satisfies the LU-SGS factorisation
  identity (D^-1 + L) D (D^-1 + U) x = MR for the increments x of the
  pinned orc_dual_time_stepping, which checks the restated blocks;
* every sweep of orc_dplusgs is one Jacobi step x' = D (MR - (L + U) x)
  of that operator from x0 = D MR, and many sweeps converge to the dense
  solve of (D^-1 + L + U) x = MR where the iteration contracts;
* its right-hand side is the pinned Dual_time_stepping's with the constant
  force of prm.h:11 in place of the buoyancy (DPLUSGS.cpp:417-420).

Also here: the constant driving force (no reference golden data holds a
nonzero one) and the single-slab periodic x ghost fill.
"""
import ctypes as C

import numpy as np
import pytest

from implicit_dense import MET, DenseOperator
from oracle_harness import IP, Case, dp, hyperplane_sizes, load_oracle, make_inputs, met9, ptrs

SEED = 20261203


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


def box(orc, n=4, deltas=(0.1, 0.08, 0.12), seed=SEED):
    """A 4^3 single slab (both x ends: zero x ghost increments) with a seeded
    nondimensional state, random metrics and no sponge or buoyancy."""
    case = Case(n, n, n, 1, 0, 0, 0, deltas)
    case.overrides = {"sponge": 0, "grav": 0.0, "f_const": 0.0}
    f = make_inputs(orc, case.slab, seed, dimensional=False)
    return case, f


def residual_inputs(case):
    sl = case.slab
    rng = np.random.default_rng(5)
    F = {k: [rng.uniform(-0.2, 0.2, sl.shape) for _ in range(5)] for k in ("inFx", "inFy", "inFz")}
    F["vF"] = [rng.uniform(-0.1, 0.1, sl.shape) for _ in range(4)]
    return F


def run_dts(orc, case, f, F, deltaT, e, sweeps=1):
    sl = case.slab
    cfg = case.cfg(orc)
    rows, cells = hyperplane_sizes(sl)
    N = np.zeros((rows, 2), np.int32)
    I = np.zeros((cells + 1, 3), np.int32)
    orc.orc_hyperplane(C.byref(cfg), N.ctypes.data_as(IP), I.ctypes.data_as(IP))
    U_ = [a.copy() for a in f["U_"]]
    MR, NR, NL, Res = sl.zeros(5), sl.zeros(5), sl.zeros(5), sl.zeros(5)
    er = (C.c_double * 5)()
    orc.orc_dual_time_stepping(C.byref(cfg), C.c_double(deltaT), C.c_double(e), er, N.ctypes.data_as(IP),
                               I.ctypes.data_as(IP), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]), dp(f["J"]),
                               ptrs(met9(f)), ptrs(F["inFx"]), ptrs(F["inFy"]), ptrs(F["inFz"]), ptrs(F["vF"]),
                               ptrs(MR), ptrs(NR), ptrs(NL), ptrs(Res), dp(f["etdy_v"]), dp(f["J_v"]), sweeps)
    return {"MR": MR, "NR": NR, "U_": U_, "er": np.array(list(er))}


def run_dplusgs(orc, case, f, F, deltaT, e, sweeps):
    sl = case.slab
    cfg = case.cfg(orc)
    U_ = [a.copy() for a in f["U_"]]
    MR, NR, NL, Res = sl.zeros(5), sl.zeros(5), sl.zeros(5), sl.zeros(5)
    er = (C.c_double * 5)()
    orc.orc_dplusgs(C.byref(cfg), C.c_double(deltaT), er, C.c_double(e), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]),
                    dp(f["J"]), ptrs(met9(f)), ptrs(F["inFx"]), ptrs(F["inFy"]), ptrs(F["inFz"]), ptrs(F["vF"]),
                    ptrs(MR), ptrs(NR), ptrs(NL), ptrs(Res), sweeps)
    return {"MR": MR, "NR": NR, "U_": U_, "er": np.array(list(er))}


def operator(case, f, deltaT, e):
    c = case.deltas
    return DenseOperator(f["U_"], f["J"], [f[n] for n in MET], deltaT, e, c[0], c[1], c[2])


def vrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("deltaT,e", [(2e-3, 1e-4), (5e-2, 0.25), (1.0, 0.25)])
def test_lusgs_factorisation_identity(orc, deltaT, e):
    """The symmetric sweep pair of the pinned oracle solves the factored
    system (D^-1 + L) D (D^-1 + U) x = MR of the dense operator."""
    case, f = box(orc)
    F = residual_inputs(case)
    out = run_dts(orc, case, f, F, deltaT, e)
    op = operator(case, f, deltaT, e)
    x = op.gather(out["NR"])
    MR = op.gather(out["MR"])
    r = vrel(op.lusgs_operator() @ x, MR)
    print("deltaT %g e %g: identity residual %.2e" % (deltaT, e, r))
    assert r <= 1e-12
    # a dropped or transposed block breaks it
    bad = (op.Dinv + op.L.T) @ op.D @ (op.Dinv + op.U) @ x
    assert vrel(bad, MR) > 1e-6


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_dplusgs_sweeps_are_jacobi_steps(orc, sweeps):
    deltaT, e = 5e-2, 0.25
    case, f = box(orc)
    F = residual_inputs(case)
    out = run_dplusgs(orc, case, f, F, deltaT, e, sweeps)
    op = operator(case, f, deltaT, e)
    MR = op.gather(out["MR"])
    want = op.jacobi(MR, sweeps)
    got = op.gather(out["NR"])
    r = vrel(got, want)
    print("sweeps %d: %.2e" % (sweeps, r))
    assert r <= 1e-12
    # Gauss-Seidel in place of Jacobi, or the sweep count off by one, fails
    assert vrel(got, op.jacobi(MR, sweeps + 1)) > 1e-6


def test_dplusgs_converges_to_the_dense_solve(orc):
    """Where the Jacobi iteration contracts (small physical step), the
    sweeps approach the solution of (D^-1 + L + U) x = MR geometrically."""
    deltaT, e = 2e-3, 0.25
    case, f = box(orc)
    F = residual_inputs(case)
    op = operator(case, f, deltaT, e)
    rho = max(abs(np.linalg.eigvals(op.D @ (op.L + op.U))))
    assert rho < 0.5, rho
    errs = []
    for sweeps in (2, 6, 12):
        out = run_dplusgs(orc, case, f, F, deltaT, e, sweeps)
        MR = op.gather(out["MR"])
        exact = np.linalg.solve(op.Dinv + op.L + op.U, MR)
        errs.append(vrel(op.gather(out["NR"]), exact))
    print("spectral radius %.3f, errors %s" % (rho, errs))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] <= 10 * rho ** 12


def test_dplusgs_rhs_is_the_dual_time_rhs_with_its_force(orc):
    """MR of DPLUSGS equals MR of the pinned Dual_time_stepping (no sponge,
    no buoyancy) with the constant force f_dplusgs as the body force."""
    case, f = box(orc)
    F = residual_inputs(case)
    case.overrides["f_dplusgs"] = 0.37
    a = run_dplusgs(orc, case, f, F, 5e-2, 0.25, 1)
    case.overrides["f_const"] = 0.37
    b = run_dts(orc, case, f, F, 5e-2, 0.25)
    sl = (slice(3, 7), slice(2, 6), slice(2, 6))
    for q in range(5):
        assert np.array_equal(a["MR"][q][sl], b["MR"][q][sl])


def test_constant_force_enters_as_gamma_inverse_of_its_source(orc):
    """On a uniform state (no flux, viscous or time-derivative residual) the
    driving force f gives MR = Gamma^-1 (0, f/J, 0, 0, f u/J).  The columns
    of Gamma^-1 are read off the pinned BDF2 term: with U = Uq = U_ - d e_q
    the residual is -3 d e_q / (2 deltaT)."""
    case, f = box(orc)
    sl = case.slab
    K, J0 = 1.4, 0.8
    rho, u, v, w, p = 1.05, 0.35, -0.1, 0.05, 2.9
    E = p / (K - 1) + 0.5 * rho * (u * u + v * v + w * w)
    q0 = [rho / J0, rho * u / J0, rho * v / J0, rho * w / J0, E / J0]
    f["J"][:] = J0
    zero = {k: [np.zeros(sl.shape) for _ in range(5)] for k in ("inFx", "inFy", "inFz")}
    zero["vF"] = [np.zeros(sl.shape) for _ in range(4)]
    dt, e, cell = 5e-2, 0.25, (4, 3, 4)

    def mr(fconst, lvl_shift=None, d=0.0):
        case.overrides["f_const"] = fconst
        g = dict(f)
        g["U_"] = [np.full(sl.shape, x) for x in q0]
        g["U"] = [np.full(sl.shape, x) for x in q0]
        if lvl_shift is not None:
            g["U"][lvl_shift] = g["U"][lvl_shift] - d
        g["Uq"] = [a.copy() for a in g["U"]]
        return np.array([a[cell] for a in run_dts(orc, case, g, zero, dt, e)["MR"]])

    d = 1e-6
    col = {qq: mr(0.0, qq, d) / (-3 * d / (2 * dt)) for qq in (1, 4)}
    fc = 0.0137
    got = mr(fc)
    want = col[1] * fc / J0 + col[4] * fc * u / J0
    print("force MR", got, "expected", want)
    assert np.max(np.abs(got - want)) <= 1e-7 * np.max(np.abs(want))
    assert np.max(np.abs(mr(0.0))) <= 1e-12 * np.max(np.abs(want))


def test_periodic_x_ghosts_are_the_periodic_images(orc):
    """Three ghost planes at each x end hold the interior planes one period
    away: ghost i takes i + nx (low end) or i - nx (high end), on a field
    whose values encode their own (i, j, k)."""
    case = Case(7, 5, 4, 1, 0, 0, 0, (0.1, 0.1, 0.1))
    sl = case.slab
    cfg = case.cfg(orc)
    cfg.x_periodic = 1
    i, j, k = np.meshgrid(*[np.arange(n, dtype=float) for n in sl.shape], indexing="ij")
    U = [1e4 * i + 1e2 * j + k + 0.1 * q for q in range(5)]
    interior = [a.copy() for a in U]
    orc.orc_bc_x_periodic(C.byref(cfg), ptrs(U))
    nx = sl.nx
    for q in range(5):
        for g in range(3):
            assert np.array_equal(U[q][g], interior[q][g + nx])
            assert np.array_equal(U[q][nx + 3 + g], interior[q][3 + g])
        assert np.array_equal(U[q][3:nx + 3], interior[q][3:nx + 3])
