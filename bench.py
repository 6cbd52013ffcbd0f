"""Benchmark: physical steps/s of the compressible channel at a 1e-6 inner drop.

Workload (the plain FP64 GPU path): channel 256x192x256, periodic x (4 pi)
and z (2 pi), isothermal walls, tanh stretching beta=2.4, Re_tau=395,
Mach 0.5, acoustic CFL 0.5, preconditioning floor e = Mach^2, laminar
Poiseuille + seeded +-10% noise (--skew adds the sinusoidal mesh skew of the
curvilinear configurations).  Every physical step iterates the dual-time
LU-SGS inner loop until max_v er_v/er_v(first) <= 1e-6 (or --max-inner),
after one evaluation of the filter sensor.  State is resident in HBM before
the timed region starts.  `value` is this path's physical steps/s.

`variants` then compares, from one saved state, plain FP64 (each step
starting from U^n, as the reference does) with FP32 residuals plus the
quadratic-extrapolation and learned initial guesses, under the reference's
stopping rule (max_v er_v / E_v <= 1e-6 with E_v the running maximum of er_v
since the rule was set) and under the per-step rule above.  The network is
first trained on the GPU (Adam, batches of random crops) on increments the
solver records while stepping; that training and its snapshot run are set-up,
outside the timed regions, and their costs are reported.

Prints ONE JSON line.  `kernels` times every kernel class with HIP events
over one step before the timed region; that names the dominant class, and
the timed region records events around that class only (two event records
per timed launch).  `roofline` is the dominant class measured so, on the
library's stream over the timed region;
`graph_replay` runs the same steps from one saved state eagerly and with
each inner iteration replayed from a captured HIP graph;
`traffic` is the HBM bytes per launch of that kernel measured by two
rocprofv3 counter passes (profiles/channel_<size>_traffic.md) for the same
workload; `hbm_copy` is the measured copy bandwidth of the box (two copy
kernels, 2 GiB); `cpu_baseline` is the oracle (C restatement, OpenMP) on
this host for a bounded number of inner iterations at the same size.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "physical steps/s and Mcell-updates/s at 1e-6 inner-residual drop; plain vs mixed+learned"
HBM_PEAK_GBS = 8000.0
# algorithmic FP64 bytes per interior cell per launch of each kernel class
# (each input element read once, each output written once), in the forms the
# default schedule runs: rhs_diag reads U_, U, Uq, J, the face fluxes and vF
# and writes MR; update_norm (k_update4) reads U_'s march-ordered copy, the
# increments and J and writes U_ to its field and its march-ordered copy;
# march_gather in an inner iteration after a step's first moves MR only
# (the first also U_: 160)
ALGO_BYTES = {"flux_x": 120, "flux_y": 120, "flux_z": 120, "viscous": 320, "rhs_diag": 320,
              "lusgs_fwd": 240, "lusgs_bwd": 240, "update_norm": 168, "split_coef": 560, "filter": 72, "march_gather": 80,
              "march_scatter": 80}


def traffic_file(nx, ny, nz):
    return os.path.join(ROOT, "profiles", "channel_%dx%dx%d_traffic.json" % (nx, ny, nz))


def workload(a, ch):
    return ("compressible channel %dx%dx%d, beta=%.1f, skew=%.2f, Re_tau=%g, Mach %.1f, acoustic CFL %.2f, "
            "preconditioning floor e=%.3g, %s dual-time LU-SGS to a %.0e inner drop"
            % (a.nx, a.ny, a.nz, a.beta, a.skew, a.re_tau, ch.mach, ch.cfl, ch.precond_e,
               "plain FP64" if a.precision == "fp64" else "FP32-residual", a.tol))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--beta", type=float, default=2.4)
    ap.add_argument("--skew", type=float, default=0.0)
    ap.add_argument("--re-tau", type=float, default=395.0)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-inner", type=int, default=200)
    ap.add_argument("--cfl", type=float, default=0.5, help="acoustic CFL of the physical step")
    ap.add_argument("--precond-e", type=float, default=0.0,
                    help="preconditioning floor e (0: Mach^2, the channel default; the reference's own air "
                         "case uses 1e-4)")
    ap.add_argument("--no-copy", action="store_true", help="skip the measured copy bandwidth")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-filter", action="store_true", help="skip the per-step filter sensor")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 residual kernels until the FP32 floor, FP32 off-diagonal sweep")
    ap.add_argument("--no-variants", action="store_true", help="skip the initial-guess / precision comparison")
    ap.add_argument("--no-graph", action="store_true", help="skip the eager / graph-replay comparison")
    ap.add_argument("--lusgs-mode", type=int, default=None,
                    help="LU-SGS schedule (0 plane order, 1 tiled wavefront, 3 march-ordered records, "
                         "4 march computing its coefficients; default: the library's, 4)")
    ap.add_argument("--sweep-waves", type=int, default=0,
                    help="schedule 4's waves per tile (0: the library's choice by tile count)")
    ap.add_argument("--variant-steps", type=int, default=20, help="timed steps of each variant")
    ap.add_argument("--gen", type=int, default=200, help="recorded steps the network trains on")
    ap.add_argument("--train", type=int, default=1000, help="Adam steps")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=32)
    return ap.parse_args()


def cpu_baseline(ctx, nv, cfg, dt, e, inner_per_step, budget_s, with_filter):
    """Oracle inner iterations on the host, same grid and state as the GPU run
    (plus one filter evaluation per physical step when the GPU run has one)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_harness import OrcCfg, dp, load_oracle, ptrs
    orc = load_oracle()
    oc = OrcCfg()
    for name, _ in OrcCfg._fields_:
        setattr(oc, name, getattr(cfg, name))
    f = {}
    for base, n, key in ((nv.U_, 5, "U_"), (nv.U, 5, "U"), (nv.UQ, 5, "Uq"), (nv.MET, 9, "met"),
                         (nv.MET_U, 9, "met_u"), (nv.MET_V, 9, "met_v"), (nv.MET_W, 9, "met_w")):
        f[key] = ctx.get_many(base, n)
    for fid, key in ((nv.J, "J"), (nv.J_U, "J_u"), (nv.J_V, "J_v"), (nv.J_W, "J_w")):
        f[key] = ctx.get(fid)
    N, I = ctx.hyperplane()
    z = lambda: [np.zeros(ctx.shape) for _ in range(5)]
    MR, ML, Fx, Fy, Fz, NL, NR, Res = z(), z(), z(), z(), z(), z(), z(), z()
    vF = [np.zeros(ctx.shape) for _ in range(4)]
    LL, MLv, NLv = z(), z(), z()
    er = (C.c_double * 5)()
    U_ = f["U_"]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    iters, t_total = 0, 0.0
    while True:
        t0 = time.perf_counter()
        orc.orc_bc_y(C.byref(oc), ptrs(U_), dp(f["J"]), C.c_double(1.0), C.c_double(1.0), -(2 ** 31), 2 ** 31 - 1)
        orc.orc_bc_z(C.byref(oc), ptrs(U_))
        orc.orc_bc_x_periodic(C.byref(oc), ptrs(U_))
        orc.orc_flux_x(C.byref(oc), C.c_double(1.0), C.c_double(e), ptrs(U_), ptrs(MR), ptrs(ML), dp(f["J"]),
                       dp(f["J_u"]), ptrs(f["met_u"]), ptrs(Fx))
        orc.orc_flux_y(C.byref(oc), C.c_double(1.0), C.c_double(e), ptrs(U_), ptrs(MR), ptrs(ML), dp(f["J"]),
                       dp(f["J_v"]), ptrs(f["met_v"]), ptrs(Fy))
        orc.orc_flux_z(C.byref(oc), C.c_double(1.0), C.c_double(e), ptrs(U_), ptrs(MR), ptrs(ML), dp(f["J"]),
                       dp(f["J_w"]), ptrs(f["met_w"]), ptrs(Fz))
        orc.orc_viscous(C.byref(oc), ptrs(U_), ptrs(vF), dp(f["J"]), dp(f["J_u"]), ptrs(f["met_u"]), dp(f["J_v"]),
                        ptrs(f["met_v"]), dp(f["J_w"]), ptrs(f["met_w"]), ptrs(LL), ptrs(MLv), ptrs(NLv))
        orc.orc_dual_time_stepping(C.byref(oc), C.c_double(dt), C.c_double(e), er, N.ctypes.data_as(C.POINTER(C.c_int)),
                                   I.ctypes.data_as(C.POINTER(C.c_int)), ptrs(U_), ptrs(f["U"]), ptrs(f["Uq"]),
                                   dp(f["J"]), ptrs(f["met"]), ptrs(Fx), ptrs(Fy), ptrs(Fz), ptrs(vF), ptrs(MR),
                                   ptrs(NR), ptrs(NL), ptrs(Res), dp(f["met_v"][4]), dp(f["J_v"]), 1)
        t_total += time.perf_counter() - t0
        iters += 1
        if t_total >= budget_s or iters >= 50:
            break
    t_inner = t_total / iters
    t_filter = 0.0
    if with_filter:
        Y = ctx.get(nv.YPOINT)
        Ep = [np.zeros(ctx.shape) for _ in range(3)]
        t0 = time.perf_counter()
        orc.orc_filter(C.byref(oc), C.c_double(0.007), C.c_double(1.0), C.c_double(0.1), ptrs(U_), dp(Y),
                       dp(Ep[0]), dp(Ep[1]), dp(Ep[2]), dp(f["J"]))
        t_filter = time.perf_counter() - t0
    steps_per_s = 1.0 / (t_inner * inner_per_step + t_filter)
    ncell = cfg.nx * cfg.ny * cfg.nz
    return {"value": steps_per_s, "unit": "physical steps/s", "cores": threads, "kind": "port",
            "mcell_updates_per_s": ncell * steps_per_s / 1e6,
            "sample": "%d inner iterations (ghosts, fluxes, viscous, dual-time LU-SGS) of the oracle at %dx%dx%d "
                      "on %d OpenMP threads, %.3f s each, scaled by the GPU run's %.2f inner iterations per "
                      "physical step, plus %.3f s for one filter evaluation"
                      % (iters, cfg.nx, cfg.ny, cfg.nz, threads, t_inner, inner_per_step, t_filter)}


# forward flop per voxel of the network: 2 x 27 x (15x32 + 32x32 + 32x32 + 32x5)
NET_FLOP_PER_VOXEL = 2 * 27 * (15 * 32 + 32 * 32 + 32 * 32 + 32 * 5)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def state_of(ctx, nv):
    return [ctx.get_many(f, 5) for f in (nv.U_, nv.U, nv.UQ)]


def restore(ctx, nv, state, E):
    for f, arrs in zip((nv.U_, nv.U, nv.UQ), state):
        ctx.set_many(f, arrs)
    ctx.rule_state(E)


def timed_steps(ctx, dt, a, steps, warm, with_filter, label):
    for s in range(warm):
        if with_filter:
            ctx.filter()
        ctx.channel_step(dt, a.tol, a.max_inner)
    ctx.sync()
    t0 = time.perf_counter()
    inner = []
    for s in range(steps):
        if with_filter:
            ctx.filter()
        it, _, _ = ctx.channel_step(dt, a.tol, a.max_inner)
        inner.append(it)
    ctx.sync()
    el = time.perf_counter() - t0
    print("%s: %d steps, %.2f inner/step, %.2f ms/step" % (label, steps, np.mean(inner), 1e3 * el / steps),
          file=sys.stderr, flush=True)
    return {"steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "inner_iterations_per_step": float(np.mean(inner))}


def variants(a, nv, ch, ctx64, dt, ncell):
    """plain FP64 vs FP32 residuals + quadratic / learned guesses from one state"""
    cfg32 = nv.channel_config(ch, a.nx, a.ny, a.nz)
    cfg32.precision = 1
    ctx32 = nv.Context(cfg32, device=0)
    if a.lusgs_mode is not None:
        ctx32.set_lusgs_mode(a.lusgs_mode)
    ctx32.channel_init(ch)
    restore(ctx32, nv, state_of(ctx64, nv), np.zeros(5))
    net = nv.Net(seed=20261203)
    net.ring(a.gen + 8, a.nx, a.ny, a.nz)
    ctx32.attach_net(net)
    ctx32.set_rule(1)
    ctx32.set_guess("quadratic")
    with_filter = not a.no_filter
    t0 = time.perf_counter()
    for s in range(a.gen):
        if with_filter:
            ctx32.filter()
        ctx32.channel_step(dt, a.tol, a.max_inner)
    ctx32.sync()
    t_gen = time.perf_counter() - t0
    print("recorded %d steps in %.1f s" % (a.gen, t_gen), file=sys.stderr, flush=True)
    loss = net.train(a.train, batch=a.batch, crop=(a.crop, min(a.crop, a.ny), a.crop), lr=1e-3, seed=7)
    tm = net.timing()
    print("trained %d steps in %.2f s: loss %.3e -> %.3e" % (a.train, tm["train_ms"] / 1e3, float(np.mean(loss[:20])),
                                                           float(np.mean(loss[-20:]))), file=sys.stderr, flush=True)
    state = state_of(ctx32, nv)
    E = ctx32.rule_state()
    nrec = net.records()
    out = {"rule_running_max": {}, "rule_per_step": {}}
    runs = [("mixed_network", ctx32, "network", 1), ("mixed_quadratic", ctx32, "quadratic", 1),
            ("mixed_extrapolate", ctx32, "extrapolate", 1), ("plain_fp64", ctx64, "previous", 1),
            ("mixed_network", ctx32, "network", 0), ("plain_fp64", ctx64, "previous", 0)]
    for name, ctx, guess, rule in runs:
        ctx.set_rule(rule)
        restore(ctx, nv, state, E)
        net.rewind(nrec)
        ctx.attach_net(net if ctx is ctx32 else None)
        ctx.set_guess(guess)
        p0 = net.timing()
        r = timed_steps(ctx, dt, a, a.variant_steps, 1, with_filter, "%s (%s rule)" % (name, "running-max" if rule else "per-step"))
        p1 = net.timing()
        if guess == "network":
            npred = p1["predicts"] - p0["predicts"]
            if npred:
                ms = (p1["predict_ms"] - p0["predict_ms"]) / npred
                r["predict_ms"] = ms
                r["predict_tflops"] = NET_FLOP_PER_VOXEL * ncell / (ms * 1e-3) / 1e12
            used = ctx.guess_info()[0]
            r["network_guess_taken_last_step"] = bool(used)
        out["rule_running_max" if rule else "rule_per_step"][name] = r
    rm = out["rule_running_max"]
    out["speedup_mixed_network_vs_plain"] = rm["mixed_network"]["steps_per_s"] / rm["plain_fp64"]["steps_per_s"]
    out["speedup_mixed_network_vs_extrapolate"] = (rm["mixed_network"]["steps_per_s"] /
                                                   rm["mixed_extrapolate"]["steps_per_s"])
    out["network"] = {"recorded_steps": a.gen, "record_seconds": t_gen, "train_steps": a.train,
                      "train_ms_per_step": tm["train_ms"] / max(1, tm["train_steps"]),
                      "batch": a.batch, "crop": a.crop, "loss_first20": float(np.mean(loss[:20])),
                      "loss_last20": float(np.mean(loss[-20:])),
                      "train_tflops": 3 * NET_FLOP_PER_VOXEL * a.batch * a.crop * min(a.crop, a.ny) * a.crop /
                                      (tm["train_ms"] / max(1, tm["train_steps"]) * 1e-3) / 1e12}
    ctx32.attach_net(None)
    ctx32.close()
    net.close()
    return out


def graph_replay(a, nv, ctx, dt):
    """The same steps from one saved state, eager and with each inner
    iteration replayed from a captured HIP graph (kernel timing off for both:
    events do not see graph launches); the results are the same bits."""
    saved = state_of(ctx, nv)
    eager = timed_steps(ctx, dt, a, a.steps, 0, not a.no_filter, "eager, untimed")
    restore(ctx, nv, saved, ctx.rule_state())
    ctx.set_graph(True)
    ctx.channel_step(dt, a.tol, 3)   # captures the iteration (its first occurrence ran above)
    restore(ctx, nv, saved, ctx.rule_state())
    graph = timed_steps(ctx, dt, a, a.steps, 0, not a.no_filter, "graph replay")
    ctx.set_graph(False)
    return {"eager_ms_per_step": eager["ms_per_step"], "graph_ms_per_step": graph["ms_per_step"],
            "inner_iterations_per_step": graph["inner_iterations_per_step"],
            "same_inner_counts": eager["inner_iterations_per_step"] == graph["inner_iterations_per_step"],
            "speedup": eager["ms_per_step"] / graph["ms_per_step"]}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no inter-GPU decomposition "
              "(requested --gpus %d)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from cgli_fu_flow_amd import native as nv
    ch = nv.channel(a.re_tau, beta=a.beta, skew=a.skew, cfl=a.cfl, precond_e=a.precond_e)
    cfg = nv.channel_config(ch, a.nx, a.ny, a.nz)
    cfg.precision = 1 if a.precision == "mixed" else 0
    copy = None
    if not a.no_copy:
        copy = nv.stream_copy(1 << 31, 20)
        print("measured copy: %.0f / %.0f GB/s" % (copy["GBps_one_double_per_lane"], copy["GBps_16B_per_lane"]),
              file=sys.stderr, flush=True)
    ctx = nv.Context(cfg, device=0)
    if a.lusgs_mode is not None:
        ctx.set_lusgs_mode(a.lusgs_mode)
    if a.sweep_waves:
        ctx.set_sweep_waves(a.sweep_waves)
    dt = ctx.channel_init(ch)
    print("setup done: dt=%.4e" % dt, file=sys.stderr, flush=True)
    for s in range(a.warmup):
        if not a.no_filter:
            ctx.filter()
        it, _, _ = ctx.channel_step(dt, a.tol, a.max_inner)
        print("warmup step %d: %d inner iterations" % (s + 1, it), file=sys.stderr, flush=True)
    # one step with every kernel class timed: the per-class table and the
    # dominant class; the timed region then records events around that class
    # only (each timed launch adds two event records to the stream)
    ctx.timing(True)
    if not a.no_filter:
        ctx.filter()
    ctx.channel_step(dt, a.tol, a.max_inner)
    ctx.sync()
    ncell = a.nx * a.ny * a.nz
    kernels = {}
    for name, (tot, cnt) in ctx.timing_report().items():
        if cnt:
            avg = tot / cnt
            kernels[name] = {"avg_ms": avg, "launches": cnt, "total_ms": tot,
                             "GBps": ALGO_BYTES[name] * ncell / (avg * 1e-3) / 1e9}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    ctx.timing(True, classes=[dom])
    ctx.sync()
    t0 = time.perf_counter()
    inner = []
    for s in range(a.steps):
        if not a.no_filter:
            ctx.filter()
        it, e1, en = ctx.channel_step(dt, a.tol, a.max_inner)
        inner.append(it)
        print("step %d: %d inner iterations, drop %.2e" % (s + 1, it, float(np.max(en / e1))), file=sys.stderr,
              flush=True)
    ctx.sync()
    t1 = time.perf_counter()
    tot, cnt = ctx.timing_report()[dom]
    ctx.timing(False)
    elapsed = t1 - t0
    steps_per_s = a.steps / elapsed
    dom_avg_ms = tot / cnt
    achieved = ALGO_BYTES[dom] * ncell / (dom_avg_ms * 1e-3) / 1e9
    # the better of the two copy forms (one double per lane over the whole grid,
    # 16 bytes per lane grid-stride) is the measured HBM peak
    copy_peak = max(copy["GBps_one_double_per_lane"], copy["GBps_16B_per_lane"]) if copy else None
    traffic = None
    tf = traffic_file(a.nx, a.ny, a.nz)
    if os.path.exists(tf):
        tj = json.load(open(tf))
        if tj.get("workload") == workload(a, ch):
            traffic = tj["bytes_per_launch"].get(dom)
    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "physical steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64 (f32 residuals and sweep records)",
        "data": "synthetic: laminar Poiseuille + seeded +-10%% velocity noise (splitmix64 seed %d)" % ch.seed,
        "config": {"workload": workload(a, ch),
                   "cells": ncell, "deltaT": dt, "max_inner": a.max_inner, "filter_every_step": not a.no_filter},
        "mcell_updates_per_s": ncell * steps_per_s / 1e6,
        "inner_iterations_per_step": float(np.mean(inner)),
        "mcell_inner_iterations_per_s": ncell * float(np.sum(inner)) / elapsed / 1e6,
        "kernels": kernels,   # one fully timed step before the timed region
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "measured_copy_peak": copy_peak,
                     "frac_of_measured_copy": achieved / copy_peak if copy_peak else None,
                     "algorithmic_bytes_per_launch": ALGO_BYTES[dom] * ncell,
                     "avg_ms_in_timed_region": dom_avg_ms, "launches_in_timed_region": cnt},
        "cpu_baseline": None,
        "hbm_copy": copy,
    }
    if not a.no_graph:
        out["graph_replay"] = graph_replay(a, nv, ctx, dt)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ctx, nv, cfg, dt, ch.precond_e, float(np.mean(inner)), a.cpu_seconds,
                                           not a.no_filter)
    if not a.no_variants and a.precision == "fp64":
        out["variants"] = variants(a, nv, ch, ctx, dt, ncell)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
