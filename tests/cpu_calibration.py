"""CPU calibration of the oracle against the reference's own routines --
TEST INFRASTRUCTURE, run by hand (not collected by pytest):

    OMP_NUM_THREADS=1 python tests/cpu_calibration.py [nx ny nz]
    OMP_NUM_THREADS=8 python tests/cpu_calibration.py nx ny nz --omp

(--omp: the reference's routines compiled with -fopenmp, as in its tuned
build, so that both run on the same OpenMP thread count)

Times one inner iteration (the three fluxes, the viscous terms and one
Dual_time_stepping call with its two hyperplane sweeps) of the C oracle and
of the reference's routines compiled from its sources
(oracle/_ref/libref_<size>_np1_in3_out3.so, built by oracle/ref/build_ref.sh
where the reference tree is present), on the same seeded inputs and the
same cores, and prints their ratio as one JSON line.  bench.py's
cpu_baseline times the oracle; this ratio says how the oracle's time
relates to the reference's.
"""
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from oracle_harness import REF_DIR, Case, load_oracle, make_inputs, ref_cfg_from, run_chain  # noqa: E402


def main():
    omp = "--omp" in sys.argv
    args = [a for a in sys.argv[1:] if a != "--omp"]
    nx, ny, nz = (int(a) for a in args[:3]) if len(args) >= 3 else (128, 128, 128)
    tag = "%dx%dx%d_np1_in3_out3%s" % (nx, ny, nz, "_omp" if omp else "")
    path = os.path.join(REF_DIR, "libref_%s.so" % tag)
    if not os.path.exists(path):
        env = dict(os.environ, CGLI_REF_OPENMP="1" if omp else "0")
        subprocess.run([os.path.join(os.path.dirname(HERE), "oracle", "ref", "build_ref.sh"), str(nx), str(ny),
                        str(nz), "1", "3", "3"], check=True, stdout=subprocess.DEVNULL, env=env)
    ref = C.CDLL(path)
    ref.ref_init()
    orc = load_oracle()
    _, deltas = ref_cfg_from(ref)
    case = Case(nx, ny, nz, 1, 3, 3, 0, deltas)
    f = make_inputs(orc, case.slab, 20261203)
    times = {}
    for backend, lib in (("reference", ref), ("oracle", orc)):
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            run_chain("ref" if backend == "reference" else "orc", lib, case, f, dts=True, sweeps=2)
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        times[backend] = best
        print("%s: %.3f s per inner iteration" % (backend, best), file=sys.stderr, flush=True)
    print(json.dumps({"grid": [nx, ny, nz], "threads": int(os.environ.get("OMP_NUM_THREADS", "0")) or None,
                      "reference_openmp": omp,
                      "reference_s_per_iteration": times["reference"], "oracle_s_per_iteration": times["oracle"],
                      "oracle_over_reference": times["oracle"] / times["reference"]}))


if __name__ == "__main__":
    main()
