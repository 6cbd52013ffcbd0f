"""Reports for the larger channel configurations.

  python -m cgli_fu_flow_amd.reports precision [--nx 256 --ny 192 --nz 256 ...]
      FP64 and FP32-residual runs of the same channel from the same state:
      physical steps/s of each at the same inner drop, and the error of the
      mixed run against the FP64 one on the conserved variables and on the
      wall-normal mean and second-moment profiles (the Statistic.cpp:130-164
      quantities).
  python -m cgli_fu_flow_amd.reports footprint [--nx 512 --ny 384 --nz 512 ...]
      device memory held by one channel context (every field and work array
      the step, the filter and, with --net, the learned guess allocate), and
      the time of its steps, against the device's HBM.
  python -m cgli_fu_flow_amd.reports learned [--gen 3000 --train 2000 --held 500 ...]
      the learned initial guess's schedule: record --gen consecutive steps of
      the 128^3 curvilinear channel (every increment goes to the network's
      ring), train the network --train Adam steps on random crops of them,
      then run the same --held steps that follow, from the same state, with
      the network's guess and with the extrapolations it is measured
      against.  The runs of one tolerance stop on the same absolute
      tolerance (--held-tols, --held-rule): per step, tol times the first
      residual the quadratic extrapolation has at that step (its 1e-6 drop
      is then everyone's target), or tol times the maxima E the recording
      run reached (the reference's running-maximum rule, which the start-up
      transient makes loose).  Per guess: inner
      iterations and wall time per step, the first inner residual relative
      to E and to the quadratic extrapolation's, and whether every step met
      the tolerance.

Each prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import native as nv


def _channel(a):
    ch = nv.channel(a.re_tau, beta=a.beta, skew=a.skew, cfl=a.cfl)
    return ch


def _run(ctx, dt, a, steps, with_filter=True):
    inner = []
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        if with_filter:
            ctx.filter()
        it, _, _ = ctx.channel_step(dt, a.tol, a.max_inner)
        inner.append(it)
        print("  step %d: %d inner iterations" % (s + 1, it), file=sys.stderr, flush=True)
    ctx.sync()
    el = time.perf_counter() - t0
    return {"steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps,
            "inner_iterations_per_step": float(np.mean(inner))}


def precision(a):
    ch = _channel(a)
    out = {"workload": "channel %dx%dx%d, Re_tau=%g, beta=%.1f, skew=%.2f, acoustic CFL %.2f, %.0e inner drop "
                       "(per-step rule), %d physical steps from the seeded initial state"
                       % (a.nx, a.ny, a.nz, a.re_tau, a.beta, a.skew, a.cfl, a.tol, a.steps)}
    res = {}
    for prec in (0, 1):
        cfg = nv.channel_config(ch, a.nx, a.ny, a.nz)
        cfg.precision = prec
        with nv.Context(cfg) as ctx:
            dt = ctx.channel_init(ch)
            name = "mixed" if prec else "fp64"
            print("%s:" % name, file=sys.stderr, flush=True)
            r = _run(ctx, dt, a, a.steps)
            r["profiles"] = ctx.profiles()
            r["U"] = ctx.get_many(nv.U, 5)
            res[name] = r
    P64, P32 = res["fp64"].pop("profiles"), res["mixed"].pop("profiles")
    U64, U32 = res["fp64"].pop("U"), res["mixed"].pop("U")
    interior = (slice(3, a.nx + 3), slice(2, a.ny + 2), slice(2, a.nz + 2))
    out["fp64"] = res["fp64"]
    out["mixed"] = res["mixed"]
    out["speedup_mixed"] = res["mixed"]["steps_per_s"] / res["fp64"]["steps_per_s"]
    out["profile_error"] = {}
    for q, name in enumerate(nv.Context.PROFILE_NAMES):
        ref = P64[q]
        # fluctuation moments are compared on the scale of the quantity's
        # range over y (a mean crossing zero has no relative error)
        scale = max(float(np.max(np.abs(ref))), 1e-300)
        out["profile_error"][name] = float(np.max(np.abs(P32[q] - ref))) / scale
    out["max_profile_error"] = max(out["profile_error"].values())
    out["state_error"] = [float(np.max(np.abs(U32[v][interior] - U64[v][interior])) /
                                np.max(np.abs(U64[v][interior]))) for v in range(5)]
    return out


def footprint(a):
    free0, total = nv.device_memory(0)
    ch = _channel(a)
    cfg = nv.channel_config(ch, a.nx, a.ny, a.nz)
    cfg.precision = 1 if a.precision == "mixed" else 0
    out = {"workload": "channel %dx%dx%d, Re_tau=%g, beta=%.1f, skew=%.2f, %s residuals%s"
                       % (a.nx, a.ny, a.nz, a.re_tau, a.beta, a.skew, a.precision,
                          ", learned guess (ring of %d records, %d^3 crops)" % (a.ring, a.crop) if a.net else "")}
    with nv.Context(cfg) as ctx:
        dt = ctx.channel_init(ch)
        net = None
        if a.net:
            net = nv.Net(seed=1)
            net.ring(a.ring, a.nx, a.ny, a.nz)
            ctx.attach_net(net)
            ctx.set_rule(1)
            ctx.set_guess("quadratic")
        print("steps:", file=sys.stderr, flush=True)
        out["run"] = _run(ctx, dt, a, a.steps)
        if net is not None:
            # a few training steps at this resolution allocate the training
            # work set; then one step with the network's guess
            loss = net.train(a.train, batch=a.batch, crop=(a.crop, a.crop, a.crop), lr=1e-3, seed=3)
            out["train_loss_first_last"] = [float(loss[0]), float(loss[-1])]
            ctx.set_guess("network")
            out["run_network"] = _run(ctx, dt, a, 1)
            out["network_guess_taken"] = bool(ctx.guess_info()[0])
        free1, _ = nv.device_memory(0)
        out["device_memory_used_gb"] = (free0 - free1) / 1e9
        out["device_memory_total_gb"] = total / 1e9
        out["fraction_of_hbm"] = (free0 - free1) / total
        out["field_gb"] = 8.0 * (a.nx + 6) * (a.ny + 4) * (a.nz + 4) / 1e9
        if net is not None:
            ctx.attach_net(None)
            net.close()
    return out


def learned(a):
    ch = _channel(a)
    cfg = nv.channel_config(ch, a.nx, a.ny, a.nz)
    cfg.precision = 1 if a.precision == "mixed" else 0
    out = {"workload": "channel %dx%dx%d, Re_tau=%g, beta=%.1f, skew=%.2f, acoustic CFL %.2f, %s residuals; "
                       "%d recorded steps, %d Adam steps (lr 1e-3, batch %d, %d^3 crops), %d held-out steps; "
                       "recorded at max_v er_v <= %.0e E_v (E the running maxima); held-out runs at "
                       "max_v er_v <= tol %s for each tol"
                       % (a.nx, a.ny, a.nz, a.re_tau, a.beta, a.skew, a.cfl, a.precision, a.gen, a.train,
                          a.batch, a.crop, a.held, a.tol,
                          "E_v" if a.held_rule == "running-max" else
                          "er_v(first) of the quadratic guess at the same step (the same threshold for every guess)")}
    with nv.Context(cfg) as ctx:
        dt = ctx.channel_init(ch)
        net = nv.Net(seed=a.seed)
        net.ring(a.gen + 8, a.nx, a.ny, a.nz)
        ctx.attach_net(net)
        ctx.set_rule(1)
        ctx.set_guess("quadratic")
        ctx.sync()
        t0 = time.perf_counter()
        for s in range(a.gen):
            ctx.filter()
            it, _, _ = ctx.channel_step(dt, a.tol, a.max_inner)
            if (s + 1) % 100 == 0:
                print("  recorded %d steps (%d inner iterations on the last)" % (s + 1, it), file=sys.stderr,
                      flush=True)
        ctx.sync()
        out["record_seconds"] = time.perf_counter() - t0
        loss = net.train(a.train, batch=a.batch, crop=(a.crop, min(a.crop, a.ny), a.crop), lr=1e-3, seed=a.seed)
        tm = net.timing()
        out["train"] = {"steps": a.train, "seconds": tm["train_ms"] / 1e3, "loss_first20": float(np.mean(loss[:20])),
                        "loss_last20": float(np.mean(loss[-20:]))}
        print("  trained: loss %.3e -> %.3e" % (out["train"]["loss_first20"], out["train"]["loss_last20"]),
              file=sys.stderr, flush=True)
        state = [ctx.get_many(f, 5) for f in (nv.U_, nv.U, nv.UQ)]
        E = ctx.rule_state()
        nrec = net.records()
        out["held_out"] = {}
        def restart(guess):
            for f, arrs in zip((nv.U_, nv.U, nv.UQ), state):
                ctx.set_many(f, arrs)
            ctx.rule_state(E)
            net.rewind(nrec)
            ctx.set_guess(guess)

        for tol in [float(t) for t in a.held_tols.split(",")]:
            runs = {}
            Eq = None
            if a.held_rule == "quadratic-drop":
                # thresholds: tol times the quadratic guess's first residual of each step
                restart("quadratic")
                ctx.set_rule(0)
                Eq = []
                for s in range(a.held):
                    ctx.filter()
                    it, e1, en = ctx.channel_step(dt, tol, a.max_inner)
                    Eq.append(e1)
                ctx.set_rule(1)
            for guess in a.guesses.split(","):
                restart(guess)
                inner, first, met, taken = [], [], 0, 0
                p0 = net.timing()
                ctx.sync()
                t0 = time.perf_counter()
                for s in range(a.held):
                    ctx.filter()
                    ref = E if Eq is None else Eq[s]
                    if Eq is not None:
                        ctx.rule_state(ref)
                    it, e1, en = ctx.channel_step(dt, tol, a.max_inner)
                    inner.append(it)
                    first.append(e1 / ref)
                    met += bool(np.all(en <= tol * ref))
                    if guess == "network":
                        taken += bool(ctx.guess_info()[0])
                ctx.sync()
                el = time.perf_counter() - t0
                p1 = net.timing()
                r = {"inner_iterations_per_step": float(np.mean(inner)), "ms_per_step": 1e3 * el / a.held,
                     "steps_meeting_tolerance": met,
                     "first_residual_over_E": [float(x) for x in np.mean(first, axis=0)]}
                if guess == "network":
                    r["network_guess_taken"] = taken
                    n = p1["predicts"] - p0["predicts"]
                    r["predict_ms"] = (p1["predict_ms"] - p0["predict_ms"]) / n if n else None
                runs[guess] = r
                print("  tol %.0e %s: %.2f inner/step, %.2f ms/step" % (tol, guess, r["inner_iterations_per_step"],
                                                                          r["ms_per_step"]), file=sys.stderr, flush=True)
            res = {"runs": runs}
            if "quadratic" in runs and "network" in runs:
                q = np.array(runs["quadratic"]["first_residual_over_E"])
                res["network_first_residual_vs_quadratic"] = [
                    float(x) for x in np.array(runs["network"]["first_residual_over_E"]) / q]
                res["network_inner_vs_quadratic"] = (runs["network"]["inner_iterations_per_step"] /
                                                     runs["quadratic"]["inner_iterations_per_step"])
                res["network_wall_vs_quadratic"] = runs["network"]["ms_per_step"] / runs["quadratic"]["ms_per_step"]
            if "extrapolate" in runs and "network" in runs:
                res["network_inner_vs_extrapolate"] = (runs["network"]["inner_iterations_per_step"] /
                                                       runs["extrapolate"]["inner_iterations_per_step"])
                res["network_wall_vs_extrapolate"] = runs["network"]["ms_per_step"] / runs["extrapolate"]["ms_per_step"]
            out["held_out"]["%.0e" % tol] = res
        out["E"] = [float(x) for x in E]
        ctx.attach_net(None)
        net.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["precision", "footprint", "learned"])
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--re-tau", type=float, default=395.0)
    ap.add_argument("--beta", type=float, default=2.4)
    ap.add_argument("--skew", type=float, default=0.05)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-inner", type=int, default=200)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--net", action="store_true")
    ap.add_argument("--ring", type=int, default=8)
    ap.add_argument("--train", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--gen", type=int, default=3000, help="learned: recorded steps")
    ap.add_argument("--held", type=int, default=500, help="learned: held-out steps")
    ap.add_argument("--guesses", default="network,quadratic,extrapolate", help="learned: guesses compared")
    ap.add_argument("--held-tols", default="1e-6,1e-4", help="learned: inner drops of the held-out runs")
    ap.add_argument("--held-rule", choices=["running-max", "quadratic-drop"], default="quadratic-drop",
                    help="learned: the absolute tolerance of the held-out runs: tol E (E the recording run's "
                         "running maxima) or, per step, tol times the quadratic guess's first residual")
    ap.add_argument("--seed", type=int, default=20261203)
    a = ap.parse_args(argv)
    out = {"precision": precision, "footprint": footprint, "learned": learned}[a.what](a)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
