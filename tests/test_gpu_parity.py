"""GPU (gfx950) kernels against the CPU oracle, called through the C ABI.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12 per conserved variable (FP64).
The oracle is itself bit-identical to the reference build (test_oracle.py).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle_harness import (FILTER_CRITERIA, Case, Slab, dp, load_oracle, make_inputs, ptrs, regions, run_chain,
                            run_filter, IP)
from gpu_harness import config_from, rel, upload_inputs

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
SEED = MANIFEST["seed"]

CASES = {
    "ref_8x10x6_both_ends": (8, 10, 6, 1, 3, 3, 0, (0.001, 0.005, 0.002)),
    "ref_24x10x6_interior": (24, 10, 6, 3, 3, 3, 1, (0.001, 0.005, 0.002)),
    "ref_12x16x10": (12, 16, 10, 1, 4, 4, 0, (0.001, 0.005, 0.002)),
    "ragged_20x24x17": (20, 24, 17, 1, 5, 6, 0, (0.002, 0.004, 0.003)),
    "wide_40x12x70": (40, 12, 70, 1, 8, 9, 0, (0.001, 0.005, 0.002)),
}


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


@pytest.fixture(scope="module")
def nv():
    from cgli_fu_flow_amd import native
    native.lib()
    return native


def make_case(name):
    X, Y, Z, NP, nin, nout, myid, d = CASES[name]
    return Case(X, Y, Z, NP, nin, nout, myid, d)


def open_ctx(nv, orc, case, xi_mode=0, **over):
    cfg = case.cfg(orc, xi_mode=xi_mode)
    for k, v in over.items():
        setattr(cfg, k, v)
    case.overrides = dict(over)
    return nv.Context(config_from(cfg))


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("xi_mode", [0, 1])
def test_fluxes_match_oracle(nv, orc, name, xi_mode):
    print("case", name, "xi_mode", xi_mode, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED)
    case_cfg = case.cfg(orc, xi_mode=xi_mode)
    want = {}
    sl = case.slab
    for d, tag in enumerate(("_u", "_v", "_w")):
        U_ = [a.copy() for a in f["U_"]]
        MR, ML, F = sl.zeros(5), sl.zeros(5), sl.zeros(5)
        fn = (orc.orc_flux_x, orc.orc_flux_y, orc.orc_flux_z)[d]
        from oracle_harness import met9
        fn(C.byref(case_cfg), C.c_double(1.0), C.c_double(1e-4), ptrs(U_), ptrs(MR), ptrs(ML), dp(f["J"]),
           dp(f["J" + tag]), ptrs(met9(f, tag)), ptrs(F))
        want["inF" + "xyz"[d]] = F
    with open_ctx(nv, orc, case, xi_mode=xi_mode) as ctx:
        upload_inputs(ctx, f, with_levels=False)
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        reg = regions(sl)
        for d, fid in enumerate((nv.INFX, nv.INFY, nv.INFZ)):
            key = "inF" + "xyz"[d]
            got = ctx.get_many(fid, 5)
            for v in range(5):
                g, w = got[v][reg[key]], want[key][v][reg[key]]
                assert np.all(np.isfinite(g))
                assert rel(g, w) <= TOL, (key, v, rel(g, w))


@pytest.mark.parametrize("name", ["ref_24x10x6_interior", "ragged_20x24x17"])
@pytest.mark.parametrize("mode", ["ep", "sensor"])
def test_scaled_dissipation_matches_oracle(nv, orc, name, mode):
    """The optional Roe dissipation scaling (Flux_X~~.cpp:552-556 semantics)
    by the call's Ep or by the filter sensor's face values."""
    print("case", name, "dissipation scaling", mode, flush=True)
    case = make_case(name)
    sl = case.slab
    f = make_inputs(orc, sl, SEED + 5)
    mi = {"ep": 1, "sensor": 2}[mode]
    Ep = 0.37
    rng = np.random.default_rng(13)
    EpF = [np.where(rng.random(sl.shape) < 0.5, 1.0, 0.1) for _ in range(3)]
    from oracle_harness import met9
    want = {}
    for d, tag in enumerate(("_u", "_v", "_w")):
        U_ = [a.copy() for a in f["U_"]]
        MR, ML, F = sl.zeros(5), sl.zeros(5), sl.zeros(5)
        fn = (orc.orc_flux_x_scaled, orc.orc_flux_y_scaled, orc.orc_flux_z_scaled)[d]
        fn(C.byref(case.cfg(orc, xi_mode=1)), mi, C.c_double(Ep), C.c_double(1e-4), ptrs(U_), ptrs(MR), ptrs(ML),
           dp(f["J"]), dp(f["J" + tag]), ptrs(met9(f, tag)), ptrs(F), dp(EpF[d]))
        want["inF" + "xyz"[d]] = F
    with open_ctx(nv, orc, case, xi_mode=1) as ctx:
        upload_inputs(ctx, f, with_levels=False)
        for d in range(3):
            ctx.set(nv.EPX + d, EpF[d])
        ctx.flux_x(Ep, 1e-4)   # unscaled first: the scaled fluxes must differ
        plain = ctx.get_many(nv.INFX, 5)
        ctx.set_dissipation_scaling(mode)
        ctx.flux_x(Ep, 1e-4)
        ctx.flux_y(Ep, 1e-4)
        ctx.flux_z(Ep, 1e-4)
        reg = regions(sl)
        for d, fid in enumerate((nv.INFX, nv.INFY, nv.INFZ)):
            key = "inF" + "xyz"[d]
            got = ctx.get_many(fid, 5)
            for v in range(5):
                g, w = got[v][reg[key]], want[key][v][reg[key]]
                assert np.all(np.isfinite(g))
                assert rel(g, w) <= TOL, (key, v, rel(g, w))
        assert not np.array_equal(plain[0][reg["inFx"]], ctx.get(nv.INFX)[reg["inFx"]])
        ctx.set_dissipation_scaling("none")
        ctx.flux_x(Ep, 1e-4)
        assert np.array_equal(plain[1], ctx.get(nv.INFX + 1))


@pytest.mark.parametrize("name", sorted(CASES))
def test_viscous_matches_oracle(nv, orc, name):
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 1)
    want = run_chain("orc", orc, case, f, dts=False)["vF"]
    with open_ctx(nv, orc, case) as ctx:
        upload_inputs(ctx, f, with_levels=False)
        ctx.viscous()
        got = ctx.get_many(nv.VF, 4)
    reg = regions(case.slab)["vF"]
    for v in range(4):
        assert rel(got[v][reg], want[v][reg]) <= TOL, (v, rel(got[v][reg], want[v][reg]))


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n][3] == 1])
@pytest.mark.parametrize("sponge", [1, 0])
def test_dual_time_stepping_matches_oracle(nv, orc, name, sponge):
    print("case", name, "sponge", sponge, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 2)
    case.overrides = {"sponge": sponge}
    want = run_chain("orc", orc, case, f, dts=True, sweeps=2)
    with open_ctx(nv, orc, case, sponge=sponge) as ctx:
        upload_inputs(ctx, f)
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        ctx.viscous()
        er = ctx.dual_time_stepping(2e-4, 1e-4, sweeps=2, write_residual=True)
        U_ = ctx.get_many(nv.U_, 5)
        NR = ctx.get_many(nv.NR, 5)
        MR = ctx.get_many(nv.MR, 5)
    reg = regions(case.slab)["cell"]
    for v in range(5):
        assert rel(MR[v][reg], want["MR"][v][reg]) <= TOL, ("MR", v)
        assert rel(NR[v][reg], want["NR"][v][reg]) <= TOL, ("NR", v, rel(NR[v][reg], want["NR"][v][reg]))
        assert rel(U_[v][reg], want["U_"][v][reg]) <= TOL, ("U_", v)
        assert abs(er[v] - want["er"][v]) <= TOL * abs(want["er"][v]), ("er", v, er[v], want["er"][v])


def orc_dplusgs(orc, case, f, chain, sweeps=4):
    sl = case.slab
    cfg = case.cfg(orc)
    U_ = [a.copy() for a in f["U_"]]
    MR, NR, NL, Res = sl.zeros(5), sl.zeros(5), sl.zeros(5), sl.zeros(5)
    er = (C.c_double * 5)()
    from oracle_harness import met9
    orc.orc_dplusgs(C.byref(cfg), C.c_double(2e-4), er, C.c_double(1e-4), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]),
                    dp(f["J"]), ptrs(met9(f)), ptrs(chain["inFx"]), ptrs(chain["inFy"]), ptrs(chain["inFz"]),
                    ptrs(chain["vF"]), ptrs(MR), ptrs(NR), ptrs(NL), ptrs(Res), sweeps)
    return U_, NR, np.array(list(er))


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n][3] == 1])
def test_dplusgs_matches_oracle(nv, orc, name):
    """cgli_dplusgs on the oracle's own face fluxes and viscous terms: the
    same inputs on both sides, so the bar is the plain 1e-12."""
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 3)
    chain = run_chain("orc", orc, case, f, dts=False)
    U_, NR, er = orc_dplusgs(orc, case, f, chain)
    with open_ctx(nv, orc, case) as ctx:
        upload_inputs(ctx, f)
        ctx.set_many(nv.INFX, chain["inFx"])
        ctx.set_many(nv.INFY, chain["inFy"])
        ctx.set_many(nv.INFZ, chain["inFz"])
        ctx.set_many(nv.VF, chain["vF"])
        ger = ctx.dplusgs(2e-4, 1e-4, sweeps=4)
        gU = ctx.get_many(nv.U_, 5)
        gNR = ctx.get_many(nv.NR, 5)
    reg = regions(case.slab)["cell"]
    for v in range(5):
        assert rel(gNR[v][reg], NR[v][reg]) <= TOL, ("NR", v)
        assert rel(gU[v][reg], U_[v][reg]) <= TOL, ("U_", v, rel(gU[v][reg], U_[v][reg]))
        assert abs(ger[v] - er[v]) <= TOL * abs(er[v])


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n][3] == 1])
def test_dplusgs_chain_tracks_oracle(nv, orc, name):
    """The whole chain on the GPU (its own fluxes and viscous terms, which
    agree with the oracle's to rounding) against the oracle's chain.  Four
    D+LU-GS sweeps amplify rounding differences of their inputs, on some of
    these random states a thousandfold (the update is as large as U itself),
    and the GPU's face fluxes differ from the oracle's by a few roundings (up
    to tens of ulps where a flux is small): the bar is 1e-12 or a hundred
    times the change that one rounding of every input (each face flux and
    viscous term multiplied by 1 +- 2^-52, signs at random) makes in the
    oracle's own result, whichever is larger."""
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 3)
    chain = run_chain("orc", orc, case, f, dts=False)
    U_, NR, er = orc_dplusgs(orc, case, f, chain)
    rng = np.random.default_rng(SEED)
    bumped = {k: [a * (1 + np.where(rng.random(a.shape) < 0.5, -1.0, 1.0) * 2.0 ** -52) for a in v]
              for k, v in chain.items()}
    U1, _, _ = orc_dplusgs(orc, case, f, bumped)
    reg = regions(case.slab)["cell"]
    with open_ctx(nv, orc, case) as ctx:
        upload_inputs(ctx, f)
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        ctx.viscous()
        ctx.dplusgs(2e-4, 1e-4, sweeps=4)
        gU = ctx.get_many(nv.U_, 5)
        gNR = ctx.get_many(nv.NR, 5)
    for v in range(5):
        assert rel(gNR[v][reg], NR[v][reg]) <= TOL, ("NR", v)
        tol = max(TOL, 100 * rel(U1[v][reg], U_[v][reg]))
        assert rel(gU[v][reg], U_[v][reg]) <= tol, ("U_", v, rel(gU[v][reg], U_[v][reg]), tol)


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n][3] == 1])
def test_ghost_fills_match_oracle(nv, orc, name):
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 4)
    cfg = case.cfg(orc)
    lo, hi = case.nin + 2, case.X - case.nout + 3
    U_ = [a.copy() for a in f["U_"]]
    orc.orc_bc_y(C.byref(cfg), ptrs(U_), dp(f["J"]), C.c_double(343.0592), C.c_double(298.0592), lo, hi)
    orc.orc_bc_z(C.byref(cfg), ptrs(U_))
    orc.orc_bc_x_periodic(C.byref(cfg), ptrs(U_))
    with open_ctx(nv, orc, case) as ctx:
        ctx.set_many(nv.U_, f["U_"])
        ctx.set(nv.J, f["J"])
        ctx.bc_y(343.0592, 298.0592, lo, hi)
        ctx.bc_z()
        ctx.bc_x_periodic()
        got = ctx.get_many(nv.U_, 5)
    for v in range(5):
        assert rel(got[v], U_[v]) <= TOL


def test_hyperplane_matches_oracle(nv, orc):
    case = make_case("ragged_20x24x17")
    cfg = case.cfg(orc)
    sl = case.slab
    rows = sl.gend + sl.NY + sl.NZ + 8
    cells = sl.nx * sl.ny * sl.nz
    N = np.zeros((rows, 2), np.int32)
    I = np.zeros((cells + 1, 3), np.int32)
    orc.orc_hyperplane(C.byref(cfg), N.ctypes.data_as(IP), I.ctypes.data_as(IP))
    with open_ctx(nv, orc, case) as ctx:
        gN, gI = ctx.hyperplane()
    assert np.array_equal(gN, N) and np.array_equal(gI, I)


def test_dual_time_stepping_is_deterministic(nv, orc):
    case = make_case("ragged_20x24x17")
    f = make_inputs(orc, case.slab, SEED + 5)
    outs = []
    for _ in range(2):
        with open_ctx(nv, orc, case) as ctx:
            upload_inputs(ctx, f)
            ctx.flux_x(1.0, 1e-4)
            ctx.flux_y(1.0, 1e-4)
            ctx.flux_z(1.0, 1e-4)
            ctx.viscous()
            er = ctx.dual_time_stepping(2e-4, 1e-4)
            outs.append((er, ctx.get_many(nv.U_, 5)))
    assert np.array_equal(outs[0][0], outs[1][0])
    for v in range(5):
        assert np.array_equal(outs[0][1][v], outs[1][1][v])


def test_missing_input_is_an_error(nv, orc):
    case = make_case("ref_8x10x6_both_ends")
    with open_ctx(nv, orc, case) as ctx:
        with pytest.raises(nv.CgliError) as ei:
            ctx.flux_x(1.0, 1e-4)
        assert ei.value.code == -5
    with open_ctx(nv, orc, case) as ctx:
        upload_inputs(ctx, make_inputs(orc, case.slab, SEED), with_levels=False)
        ctx.set_dissipation_scaling("sensor")   # no sensor values yet
        with pytest.raises(nv.CgliError) as ei:
            ctx.flux_z(1.0, 1e-4)
        assert ei.value.code == -5
        with pytest.raises(nv.CgliError) as ei:
            ctx.set_dissipation_scaling(3)
        assert ei.value.code == -1


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n][3] == 1])
def test_wavefront_and_plane_schedules_agree(nv, orc, name):
    """The tiled wavefronts visit cells in another topological order of the
    same dependencies as the reference's plane order: results agree to
    rounding.  The wavefront on coefficient records in field blocks (mode 1)
    and in march order (mode 3) do the same arithmetic on the same values:
    identical bits.  The march that computes its own coefficients (mode 4,
    the default) multiplies by reciprocals where the others divide: it
    agrees with both to rounding."""
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 6)
    outs = []
    for mode in (0, 1, 3, 4):
        with open_ctx(nv, orc, case) as ctx:
            ctx.set_lusgs_mode(mode)
            upload_inputs(ctx, f)
            ctx.flux_x(1.0, 1e-4)
            ctx.flux_y(1.0, 1e-4)
            ctx.flux_z(1.0, 1e-4)
            ctx.viscous()
            er = ctx.dual_time_stepping(2e-4, 1e-4, sweeps=2)
            outs.append((er, ctx.get_many(nv.NL, 5), ctx.get_many(nv.NR, 5), ctx.get_many(nv.U_, 5)))
    reg = regions(case.slab)["cell"]
    for v in range(5):
        assert outs[1][0][v] == outs[2][0][v]
        for m in (1, 3):
            assert abs(outs[0][0][v] - outs[m][0][v]) <= 1e-13 * abs(outs[0][0][v]), (m, v)
        for q in (1, 2, 3):
            a = outs[0][q][v][reg]
            for m in (1, 3):
                b = outs[m][q][v][reg]
                assert np.all(np.isfinite(b))
                assert rel(b, a) <= 1e-13, (m, q, v, rel(b, a))
            assert np.array_equal(outs[1][q][v], outs[2][q][v])


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n][3] == 1])
def test_sweep_helper_waves_give_the_same_bits(nv, orc, name):
    """Schedule 4 with helper waves computing the coefficients ahead into
    LDS slots does the same arithmetic on the same values as the one-wave
    march: identical bits for every wave count (these slabs' metrics vary
    along i, so the slots carry the metrics too)."""
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 6)
    outs = []
    for waves in (1, 2, 3, 4):
        with open_ctx(nv, orc, case) as ctx:
            ctx.set_sweep_waves(waves)
            upload_inputs(ctx, f)
            ctx.flux_x(1.0, 1e-4)
            ctx.flux_y(1.0, 1e-4)
            ctx.flux_z(1.0, 1e-4)
            ctx.viscous()
            er = ctx.dual_time_stepping(2e-4, 1e-4, sweeps=2)
            outs.append((er, ctx.get_many(nv.NL, 5), ctx.get_many(nv.NR, 5), ctx.get_many(nv.U_, 5)))
    for w in (1, 2, 3):
        assert np.array_equal(outs[0][0], outs[w][0]), w
        for q in (1, 2, 3):
            for v in range(5):
                assert np.array_equal(outs[0][q][v], outs[w][q][v]), (w, q, v)


def test_unknown_sweep_wave_count_is_an_error(nv, orc):
    case = make_case("ref_8x10x6_both_ends")
    with open_ctx(nv, orc, case) as ctx:
        for waves in (-1, 5):
            with pytest.raises(nv.CgliError) as ei:
                ctx.set_sweep_waves(waves)
            assert ei.value.code == -1
        ctx.set_sweep_waves(0)


def test_unknown_schedule_is_an_error(nv, orc):
    case = make_case("ref_8x10x6_both_ends")
    with open_ctx(nv, orc, case) as ctx:
        for mode in (-1, 2, 5):
            with pytest.raises(nv.CgliError) as ei:
                ctx.set_lusgs_mode(mode)
            assert ei.value.code == -1


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("periodic", [0, 1])
def test_filter_matches_oracle(nv, orc, name, periodic):
    """Filter sensor: identical E_high/E_low decisions on every face of the
    reference's windows, for both thresholds (the second splits the faces
    roughly in half); and the filtered energies E_h, E_2h of every cell a
    face compares, finite and equal to the oracle's at 1e-12 (a work cell
    read before it is written holds NaN, which the decisions alone would
    hide where the oracle picks E_low)."""
    print("case", name, "periodic", periodic, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 7)
    over = {"x_periodic": 1, "first": 0, "last": 0} if periodic else {}
    case.overrides = dict(over)
    reg = regions(case.slab)
    sl = case.slab
    g, NY, NZ = sl.gend, sl.NY, sl.NZ
    # the cells the x, y and z faces compare (Filter.cpp:1564-1643)
    cells = [(slice(2, g + 2), slice(2, NY + 1), slice(2, NZ + 1)), (slice(3, g + 1), slice(1, NY + 2), slice(2, NZ + 1)),
             (slice(3, g + 1), slice(2, NY + 1), slice(1, NZ + 2))]
    Eh_o, E2h_o = sl.zeros(), sl.zeros()
    orc.orc_filter_energy(C.byref(case.cfg(orc)), ptrs(f["U_"]), dp(f["Y_point"]), dp(f["J"]), dp(Eh_o), dp(E2h_o))
    with open_ctx(nv, orc, case, **over) as ctx:
        ctx.set_many(nv.U_, f["U_"])
        ctx.set(nv.J, f["J"])
        ctx.set(nv.YPOINT, f["Y_point"])
        for crit in FILTER_CRITERIA:
            want = run_filter("orc", orc, case, f, crit)
            ctx.filter(crit, 1.0, 0.1)
            got = ctx.get_many(nv.EPX, 3)
            for d, nm in enumerate(("EpX", "EpY", "EpZ")):
                gg, w = got[d][reg[nm]], want[d][reg[nm]]
                assert np.array_equal(gg, w), (nm, crit, int(np.sum(gg != w)))
        Eh, E2h = ctx.filter_energies()
        for win in cells:
            for a, b in ((Eh, Eh_o), (E2h, E2h_o)):
                assert np.all(np.isfinite(a[win]))
                assert rel(a[win], b[win]) <= TOL, rel(a[win], b[win])


def test_filter_needs_its_inputs(nv, orc):
    case = make_case("ref_12x16x10")
    with open_ctx(nv, orc, case) as ctx:
        ctx.set_many(nv.U_, make_inputs(orc, case.slab, SEED)["U_"])
        with pytest.raises(nv.CgliError) as ei:
            ctx.filter()
        assert ei.value.code == -5


@pytest.mark.parametrize("name", ["ref_12x16x10", "ragged_20x24x17"])
def test_profiles_match_oracle(nv, orc, name):
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 8)
    cfg = case.cfg(orc)
    want = np.zeros(11 * case.slab.ny)
    orc.orc_profiles(C.byref(cfg), ptrs(f["U_"]), dp(f["J"]), dp(want))
    want = want.reshape(11, case.slab.ny)
    with open_ctx(nv, orc, case) as ctx:
        ctx.set_many(nv.U_, f["U_"])
        ctx.set(nv.J, f["J"])
        got = ctx.profiles()
        again = ctx.profiles()
    assert np.array_equal(got, again)                 # fixed-order reduction
    for q in range(11):
        assert rel(got[q], want[q]) <= TOL, (q, rel(got[q], want[q]))


# FP32 residual kernels (cfg.precision = 1) against the FP64 oracle: the
# bar is single-precision rounding of the stencils, max-norm relative per
# variable
FP32_TOL = {"inF": 5e-5, "vF": 1e-4}


@pytest.mark.parametrize("name", ["ref_12x16x10", "ragged_20x24x17", "wide_40x12x70"])
def test_fp32_residuals_track_the_fp64_oracle(nv, orc, name):
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 9)
    want = run_chain("orc", orc, case, f, dts=False)
    reg = regions(case.slab)
    with open_ctx(nv, orc, case, precision=1) as ctx:
        upload_inputs(ctx, f, with_levels=False)
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        ctx.viscous()
        worst = {}
        for d, fid in enumerate((nv.INFX, nv.INFY, nv.INFZ)):
            key = "inF" + "xyz"[d]
            got = ctx.get_many(fid, 5)
            for v in range(5):
                r = rel(got[v][reg[key]], want[key][v][reg[key]])
                worst[(key, v)] = r
                assert r <= FP32_TOL["inF"], (key, v, r)
                assert r > 0.0   # really single precision
        got = ctx.get_many(nv.VF, 4)
        for v in range(4):
            r = rel(got[v][reg["vF"]], want["vF"][v][reg["vF"]])
            worst[("vF", v)] = r
            assert r <= FP32_TOL["vF"], ("vF", v, r)
    print("worst FP32 relative error", max(worst.values()), flush=True)


def test_fp32_scaled_dissipation_tracks_the_fp64_oracle(nv, orc):
    """The FP32 fluxes with the sensor-scaled Roe dissipation against the
    FP64 oracle's, at the FP32 residual tolerance."""
    name = "ragged_20x24x17"
    print("case", name, "FP32 sensor-scaled dissipation", flush=True)
    case = make_case(name)
    sl = case.slab
    f = make_inputs(orc, sl, SEED + 10)
    rng = np.random.default_rng(17)
    EpF = [np.where(rng.random(sl.shape) < 0.5, 1.0, 0.1) for _ in range(3)]
    from oracle_harness import met9
    want = {}
    for d, tag in enumerate(("_u", "_v", "_w")):
        U_ = [a.copy() for a in f["U_"]]
        MR, ML, F = sl.zeros(5), sl.zeros(5), sl.zeros(5)
        fn = (orc.orc_flux_x_scaled, orc.orc_flux_y_scaled, orc.orc_flux_z_scaled)[d]
        fn(C.byref(case.cfg(orc, xi_mode=1)), 2, C.c_double(1.0), C.c_double(1e-4), ptrs(U_), ptrs(MR), ptrs(ML),
           dp(f["J"]), dp(f["J" + tag]), ptrs(met9(f, tag)), ptrs(F), dp(EpF[d]))
        want["inF" + "xyz"[d]] = F
    reg = regions(sl)
    with open_ctx(nv, orc, case, xi_mode=1, precision=1) as ctx:
        upload_inputs(ctx, f, with_levels=False)
        for d in range(3):
            ctx.set(nv.EPX + d, EpF[d])
        ctx.set_dissipation_scaling("sensor")
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        for d, fid in enumerate((nv.INFX, nv.INFY, nv.INFZ)):
            key = "inF" + "xyz"[d]
            got = ctx.get_many(fid, 5)
            for v in range(5):
                r = rel(got[v][reg[key]], want[key][v][reg[key]])
                assert r <= FP32_TOL["inF"], (key, v, r)
                assert r > 0.0   # really single precision


@pytest.mark.parametrize("name", ["ref_12x16x10", "ragged_20x24x17"])
def test_mixed_precision_iteration_tracks_the_fp64_oracle(nv, orc, name):
    """One inner iteration with cfg.precision = 1 (FP32 residual kernels and
    an FP32 off-diagonal sweep around the FP64 diagonal, update and norm):
    the increments agree with the FP64 oracle to single-precision rounding.
    The wavefront on march-ordered records gives the same bits."""
    print("case", name, flush=True)
    case = make_case(name)
    f = make_inputs(orc, case.slab, SEED + 10)
    case.overrides = {"sponge": 0}
    want = run_chain("orc", orc, case, f, dts=True, sweeps=1)
    got = []
    for mode in (1, 3, 4):
        with open_ctx(nv, orc, case, sponge=0, precision=1) as ctx:
            ctx.set_lusgs_mode(mode)
            upload_inputs(ctx, f)
            ctx.flux_x(1.0, 1e-4)
            ctx.flux_y(1.0, 1e-4)
            ctx.flux_z(1.0, 1e-4)
            ctx.viscous()
            got.append((ctx.dual_time_stepping(2e-4, 1e-4, sweeps=1), ctx.get_many(nv.NR, 5)))
    (er, NR), (er3, NR3), (er4, NR4) = got
    assert list(er) == list(er3)
    for v in range(5):
        assert np.array_equal(NR[v], NR3[v])
    reg = regions(case.slab)["cell"]
    worst = 0.0
    for v in range(5):
        for e_, N_ in ((er, NR), (er4, NR4)):   # mode 4 sweeps in FP64 on FP32 residuals
            r = rel(N_[v][reg], want["NR"][v][reg])
            worst = max(worst, r)
            assert r <= 1e-4, ("NR", v, r)
            assert abs(e_[v] - want["er"][v]) <= 1e-4 * abs(want["er"][v])
    print("worst mixed-precision increment error", worst, flush=True)
    assert worst > 1e-12   # the FP32 kernels really ran
