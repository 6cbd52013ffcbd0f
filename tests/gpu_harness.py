"""GPU-side helpers for the parity tests: upload oracle inputs to a context.

TEST INFRASTRUCTURE ONLY."""
import numpy as np

from cgli_fu_flow_amd import native as nv


def config_from(orc_cfg):
    c = nv.Config()
    for name, _ in nv.Config._fields_:
        setattr(c, name, getattr(orc_cfg, name))
    return c


def upload_inputs(ctx, f, with_levels=True):
    ctx.set_many(nv.U_, f["U_"])
    if with_levels:
        ctx.set_many(nv.U, f["U"])
        ctx.set_many(nv.UQ, f["Uq"])
    ctx.set(nv.J, f["J"])
    for base, tag in ((nv.MET, ""), (nv.MET_U, "_u"), (nv.MET_V, "_v"), (nv.MET_W, "_w")):
        for q, n in enumerate(nv.METRIC_NAMES):
            ctx.set(base + q, f[n + tag])
    ctx.set(nv.J_U, f["J_u"])
    ctx.set(nv.J_V, f["J_v"])
    ctx.set(nv.J_W, f["J_w"])


def rel(a, b):
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    return num / den if den else num
