"""The reference-signature drop-ins (include/cgli_reference_api.hpp) on the GPU.

A C++ probe compiled like a reference translation unit (double (*)[Y_m][Z_m]
arguments) calls Flux_X/Y/Z, Viscous_terms, Dual_time_stepping, DPLUSGS,
Filter and Hyperplane with host arrays; the results must match the CPU
oracle to the same bar as the C-ABI parity tests.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle_harness import (FILTER_CRITERIA, IP, Case, dp, load_oracle, make_inputs, met9, ptrs, regions,
                            run_chain, run_filter)
from gpu_harness import config_from, rel

pytestmark = pytest.mark.gpu
TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))
MANIFEST = json.load(open(os.path.join(HERE, "golden", "manifest.json")))
SEED = MANIFEST["seed"] + 11
CASES = {"12x16x10": (12, 16, 10, 1, 4, 4, 0, (0.001, 0.005, 0.002)),
         "8x10x6": (8, 10, 6, 1, 3, 3, 0, (0.001, 0.005, 0.002))}


def probe_lib(ny, nz):
    from cgli_fu_flow_amd import native as nv
    nv.lib()   # the probe links the same libcgli
    path = os.path.join(HERE, "ref_api", "_build", "libshim_%d_%d.so" % (ny, nz))
    if not os.path.exists(path):
        pytest.fail("reference-signature probe not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for n in ("probe_flux", "probe_viscous", "probe_implicit", "probe_filter", "probe_hyperplane"):
        getattr(lib, n).argtypes = None
    return lib


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_signatures_match_oracle(name):
    from cgli_fu_flow_amd import native as nv
    print("case", name, flush=True)
    X, Y, Z, NP, nin, nout, myid, d = CASES[name]
    case = Case(X, Y, Z, NP, nin, nout, myid, d)
    orc = load_oracle()
    sl = case.slab
    f = make_inputs(orc, sl, SEED)
    lib = probe_lib(Y, Z)
    want = run_chain("orc", orc, case, f, dts=True, sweeps=2)
    with nv.Context(config_from(case.cfg(orc))) as ctx:
        h = ctx.h
        # fluxes, exactly as main.cpp calls them (same U_ on every call)
        U_ = [a.copy() for a in f["U_"]]
        scratch = sl.zeros(10)
        F = {}
        for dd, tag in enumerate(("_u", "_v", "_w")):
            F[dd] = sl.zeros(5)
            assert lib.probe_flux(h, dd, C.c_double(1.0), C.c_double(1e-4), ptrs(U_), ptrs(scratch), dp(f["J"]),
                                  dp(f["J" + tag]), ptrs(met9(f, tag)), ptrs(F[dd])) == 0
        reg = regions(sl)
        for dd, key in enumerate(("inFx", "inFy", "inFz")):
            for v in range(5):
                assert rel(F[dd][v][reg[key]], want[key][v][reg[key]]) <= TOL, (key, v)
        vF = sl.zeros(4)
        assert lib.probe_viscous(h, ptrs(U_), ptrs(vF), dp(f["J"]), ptrs(met9(f)), dp(f["J_u"]), ptrs(met9(f, "_u")),
                                 dp(f["J_v"]), ptrs(met9(f, "_v")), dp(f["J_w"]), ptrs(met9(f, "_w")),
                                 ptrs(sl.zeros(30))) == 0
        for v in range(4):
            assert rel(vF[v][reg["vF"]], want["vF"][v][reg["vF"]]) <= TOL, ("vF", v)
        # Hyperplane writes the reference's lists
        rows, cells = sl.gend + sl.NY + sl.NZ + 8, sl.nx * sl.ny * sl.nz
        N = np.zeros((rows, 2), np.int32)
        I = np.zeros((3 * cells + 1, 3), np.int32)
        assert lib.probe_hyperplane(h, N.ctypes.data_as(IP), I.ctypes.data_as(IP)) == 0
        assert np.array_equal(N, want["Nijk"]) and np.array_equal(I[: cells + 1], want["ijk"][: cells + 1])
        # Dual_time_stepping (2 sweeps, residual written)
        MR, NR, NL, Res = sl.zeros(5), sl.zeros(5), sl.zeros(5), sl.zeros(5)
        er = (C.c_double * 5)()
        assert lib.probe_implicit(h, 0, C.c_double(2e-4), C.c_double(1e-4), er, N.ctypes.data_as(IP),
                                  I.ctypes.data_as(IP), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]), dp(f["J"]),
                                  ptrs(met9(f)), ptrs(F[0]), ptrs(F[1]), ptrs(F[2]), ptrs(vF), ptrs(MR), ptrs(NR),
                                  ptrs(NL), ptrs(Res), dp(f["etdy_v"]), dp(f["J_v"])) == 0
        cell = reg["cell"]
        for v in range(5):
            assert rel(U_[v][cell], want["U_"][v][cell]) <= TOL, ("U_", v)
            assert rel(NR[v][cell], want["NR"][v][cell]) <= TOL, ("NR", v)
            assert rel(MR[v][cell], want["MR"][v][cell]) <= TOL, ("MR", v)
            assert abs(er[v] - want["er"][v]) <= TOL * abs(want["er"][v]), ("er", v)
        # Filter: both thresholds
        for crit in FILTER_CRITERIA:
            w = run_filter("orc", orc, case, f, crit)
            Ep = [np.full(sl.shape, -7.0) for _ in range(3)]
            assert lib.probe_filter(h, C.c_double(crit), C.c_double(1.0), C.c_double(0.1), ptrs(f["U_"]),
                                    dp(f["Y_point"]), dp(Ep[0]), dp(Ep[1]), dp(Ep[2]), dp(f["J"])) == 0
            for q in range(3):
                assert np.array_equal(Ep[q], w[q]), ("Ep", q, crit)   # windows and untouched entries


def test_dplusgs_signature_matches_oracle():
    from cgli_fu_flow_amd import native as nv
    X, Y, Z, NP, nin, nout, myid, d = CASES["12x16x10"]
    case = Case(X, Y, Z, NP, nin, nout, myid, d)
    orc = load_oracle()
    sl = case.slab
    f = make_inputs(orc, sl, SEED + 1)
    lib = probe_lib(Y, Z)
    chain = run_chain("orc", orc, case, f, dts=False)
    cfg = case.cfg(orc)
    U_ = [a.copy() for a in f["U_"]]
    MR, NR, NL, Res = sl.zeros(5), sl.zeros(5), sl.zeros(5), sl.zeros(5)
    er = (C.c_double * 5)()
    orc.orc_dplusgs(C.byref(cfg), C.c_double(2e-4), er, C.c_double(1e-4), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]),
                    dp(f["J"]), ptrs(met9(f)), ptrs(chain["inFx"]), ptrs(chain["inFy"]), ptrs(chain["inFz"]),
                    ptrs(chain["vF"]), ptrs(MR), ptrs(NR), ptrs(NL), ptrs(Res), 4)
    with nv.Context(config_from(cfg)) as ctx:
        gU = [a.copy() for a in f["U_"]]
        gMR, gNR, gNL, gRes = sl.zeros(5), sl.zeros(5), sl.zeros(5), sl.zeros(5)
        ger = (C.c_double * 5)()
        assert lib.probe_implicit(ctx.h, 1, C.c_double(2e-4), C.c_double(1e-4), ger, None, None,
                                  ptrs(gU), ptrs(f["U"]), ptrs(f["Uq"]), dp(f["J"]), ptrs(met9(f)),
                                  ptrs(chain["inFx"]), ptrs(chain["inFy"]), ptrs(chain["inFz"]), ptrs(chain["vF"]),
                                  ptrs(gMR), ptrs(gNR), ptrs(gNL), ptrs(gRes), None, None) == 0
    cell = regions(sl)["cell"]
    for v in range(5):
        assert rel(gU[v][cell], U_[v][cell]) <= TOL, ("U_", v)
        assert rel(gNR[v][cell], NR[v][cell]) <= TOL, ("NR", v)
        assert abs(ger[v] - er[v]) <= TOL * abs(er[v])
