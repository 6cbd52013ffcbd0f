"""Write the golden fixtures of the hot path from the reference's own code.

Run in a container that has the reference tree:  python tests/golden/make_golden.py

It builds (if needed) oracle/_ref/libref_<size>.so from the reference sources
with oracle/ref/build_ref.sh, fills seeded inputs with the oracle library's
splitmix64 generator, runs the reference routines (Flux_X/Y/Z,
Viscous_terms, Hyperplane, Dual_time_stepping, Jacobian*,
Y_/Z_boundary_condition, Filter) and stores their outputs on the index windows they
define.  Only data is written: golden_<size>.npz plus manifest.json with the
inputs' seed and SHA-256 of every fixture.
"""
import ctypes as C
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_harness import (Case, FILTER_CRITERIA, REF_DIR, ORACLE_DIR, dp, load_oracle, make_inputs,  # noqa: E402
                            met9, ptrs, ref_cfg_from, regions, run_chain, run_filter)

SEED = 20261203
# (X_out, Y_out, Z_out, np, nx_inlet, nx_outlet, myid, run Dual_time_stepping)
CONFIGS = [
    (8, 10, 6, 1, 3, 3, 0, True),     # one slab touching both x ends, sponges inside
    (24, 10, 6, 3, 3, 3, 1, False),   # interior slab: 5th order everywhere in x
    (12, 16, 10, 1, 4, 4, 0, True),   # larger single slab
]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    orc = load_oracle()
    manifest = {"seed": SEED, "generator": "splitmix64 via oracle orc_fill/orc_fill_state", "cases": {}}
    for X, Y, Z, NP, nin, nout, myid, dts in CONFIGS:
        tag = "%dx%dx%d_np%d_in%d_out%d" % (X, Y, Z, NP, nin, nout)
        path = os.path.join(REF_DIR, "libref_%s.so" % tag)
        if not os.path.exists(path):
            subprocess.run([os.path.join(ORACLE_DIR, "ref", "build_ref.sh"), str(X), str(Y), str(Z), str(NP),
                            str(nin), str(nout)], check=True)
        ref = C.CDLL(path)
        ref.ref_init()
        dims, deltas = ref_cfg_from(ref)
        case = Case(X, Y, Z, NP, nin, nout, myid, deltas)
        f = make_inputs(orc, case.slab, SEED)
        out = run_chain("ref", ref, case, f, dts=dts)
        reg = regions(case.slab)
        fx = {}
        for name in ("inFx", "inFy", "inFz"):
            fx[name] = np.stack([a[reg[name]] for a in out[name]])
        fx["vF"] = np.stack([a[reg["vF"]] for a in out["vF"]])
        if dts:
            fx["U_"] = np.stack([a[reg["cell"]] for a in out["U_"]])
            fx["NR"] = np.stack([a[reg["cell"]] for a in out["NR"]])
            fx["er"] = out["er"]
            fx["Nijk"] = out["Nijk"]
            fx["ijk"] = out["ijk"][: case.slab.nx * case.slab.ny * case.slab.nz + 1]
        # metrics and ghost fills of the reference's analytic grid
        if NP == 1:
            sl = case.slab
            J, m = sl.zeros(), sl.zeros(9)
            ref.ref_jacobian(0, case.xnp, myid, dp(J), ptrs(m), None, None)
            Ju, mu = sl.zeros(), sl.zeros(9)
            ref.ref_jacobian(1, case.xnp, myid, dp(Ju), ptrs(mu), None, None)
            Jv, mv = sl.zeros(), sl.zeros(9)
            ref.ref_jacobian(2, case.xnp, myid, dp(Jv), ptrs(mv), dp(J), ptrs(m))
            Jw, mw = sl.zeros(), sl.zeros(9)
            ref.ref_jacobian(3, case.xnp, myid, dp(Jw), ptrs(mw), None, None)
            fx["jac_cell"] = np.stack([J] + m)
            fx["jac_u"] = np.stack([Ju] + mu)
            fx["jac_v"] = np.stack([Jv] + mv)
            fx["jac_w"] = np.stack([Jw] + mw)
            U_ = [a.copy() for a in f["U_"]]
            ref.ref_bc_yz(0, case.xnp, myid, C.c_double(2e-4), C.c_double(1e-4), ptrs(U_), ptrs(f["U"]),
                          ptrs(f["Uq"]), dp(f["etdy"]), dp(f["J"]))
            ref.ref_bc_yz(1, case.xnp, myid, C.c_double(2e-4), C.c_double(1e-4), ptrs(U_), ptrs(f["U"]),
                          ptrs(f["Uq"]), dp(f["etdy"]), dp(f["J"]))
            fx["bc_U_"] = np.stack(U_)
            # the filter sensor (Filter.cpp): single-slab builds only, an
            # interior slab's halo exchange needs its MPI neighbours
            for q, crit in enumerate(FILTER_CRITERIA):
                Ep = run_filter("ref", ref, case, f, crit)
                for d, nm in enumerate(("EpX", "EpY", "EpZ")):
                    fx["%s_%d" % (nm, q)] = Ep[d][reg[nm]]
        # full outputs are pinned by SHA-256; the npz keeps values on windows
        # small enough to commit (all of them for the small slabs)
        digests = {k: sha(v) for k, v in fx.items()}
        small = case.slab.nx * case.slab.ny * case.slab.nz <= 500
        keep = {}
        for k, v in fx.items():
            if k.startswith("jac_"):
                keep[k] = v[:, :, :, 2]                      # one k plane
            elif k == "bc_U_":
                sl = case.slab
                keep["bc_U_jplanes"] = v[:, :, [1, 2, sl.NY, sl.NY + 1], :]
                keep["bc_U_kplanes"] = v[:, :, :, [1, 2, sl.NZ, sl.NZ + 1]]
            elif small or k in ("er", "Nijk") or k.startswith("Ep"):
                keep[k] = v
            elif k != "ijk":
                keep[k] = v[:, :2]                           # first two i planes
        fname = "golden_%s.npz" % tag
        np.savez_compressed(os.path.join(HERE, fname), **keep)
        manifest["cases"][tag] = {
            "file": fname, "myid": myid, "dims": list(dims), "deltas": list(deltas),
            "sha256": digests,
        }
        print(tag, {k: v.shape for k, v in fx.items()})
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
