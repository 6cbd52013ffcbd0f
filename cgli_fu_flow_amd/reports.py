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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["precision", "footprint"])
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
    a = ap.parse_args(argv)
    out = precision(a) if a.what == "precision" else footprint(a)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
