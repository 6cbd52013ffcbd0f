"""Inner iterations replayed from HIP graphs against the same calls made eagerly.

A graph records the same launches the eager calls make, so every replay must
give the eager result bit for bit.  The channel tests replay each captured
iteration many times, with the step's eager work (the first iteration, the
time-level shift, the host's stopping test) and synchronisations between the
replays, and compare every physical step; the generic test replays one
captured iteration on reset inputs and checks the capture rules.
"""
import numpy as np
import pytest

from oracle_harness import Case, load_oracle, make_inputs
from gpu_harness import config_from, upload_inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol,max_inner", [(0, 0.0, 4), (1, 1e-6, 40)])
def test_channel_graph_steps_equal_eager_steps(precision, tol, max_inner):
    from cgli_fu_flow_amd import native as nv
    nx, ny, nz = 16, 24, 12
    print("channel graph precision", precision, "tol", tol, flush=True)
    ch = nv.channel(180.0, beta=2.2, skew=0.05, xi_mode=1)
    cfg = nv.channel_config(ch, nx, ny, nz)
    cfg.precision = precision
    with nv.Context(cfg) as a, nv.Context(cfg) as b:
        dt = a.channel_init(ch)
        b.channel_init(ch)
        b.set_graph(True)
        for step in range(4):
            ra = a.channel_step(dt, tol=tol, max_inner=max_inner)
            rb = b.channel_step(dt, tol=tol, max_inner=max_inner)
            assert ra[0] == rb[0], (step, ra[0], rb[0])
            assert np.array_equal(ra[1], rb[1]) and np.array_equal(ra[2], rb[2]), step
            assert a.last_step_fp32() == b.last_step_fp32()
            for q in range(5):
                assert np.array_equal(a.get(nv.U_ + q), b.get(nv.U_ + q)), (step, q)
        if precision == 1:
            assert 0 < b.last_step_fp32() < ra[0]   # both graph keys were used
        # a schedule change makes the graphs stale: they are captured again
        for ctx in (a, b):
            ctx.set_lusgs_mode(1)
        ra = a.channel_step(dt, tol=tol, max_inner=max_inner)
        rb = b.channel_step(dt, tol=tol, max_inner=max_inner)
        assert ra[0] == rb[0] and np.array_equal(ra[2], rb[2])
        for q in range(5):
            assert np.array_equal(a.get(nv.U_ + q), b.get(nv.U_ + q)), ("schedule 1", q)


def test_graph_of_one_iteration_replays_the_eager_result():
    from cgli_fu_flow_amd import native as nv
    orc = load_oracle()
    case = Case(20, 24, 17, 1, 5, 6, 0, (0.002, 0.004, 0.003))
    f = make_inputs(orc, case.slab, 20261203 + 9)
    print("generic graph", (20, 24, 17), flush=True)

    def iteration(ctx):
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        ctx.viscous()
        ctx.dual_time_stepping_async(2e-4, 1e-4, sweeps=2)

    with nv.Context(config_from(case.cfg(orc))) as ctx:
        upload_inputs(ctx, f)
        iteration(ctx)
        er0 = ctx.dual_time_stepping_wait()
        U0 = ctx.get_many(nv.U_, 5)
        with pytest.raises(nv.CgliError) as e:   # nothing queued
            ctx.dual_time_stepping_wait()
        assert e.value.code == -5

        ctx.graph_begin()
        iteration(ctx)
        with pytest.raises(nv.CgliError) as e:   # waits for the device: refused
            ctx.get(nv.U_)
        assert e.value.code == -5
        with pytest.raises(nv.CgliError) as e:
            ctx.set_lusgs_mode(0)
        assert e.value.code == -5
        g = ctx.graph_end()
        try:
            for rep in range(3):
                ctx.set_many(nv.U_, f["U_"])   # the captured iteration's input again
                g.launch()
                er = ctx.dual_time_stepping_wait()
                assert np.array_equal(er, er0), rep
                for q in range(5):
                    assert np.array_equal(ctx.get(nv.U_ + q), U0[q]), (rep, q)
                # NL equals NR on the interior after every replay, also when NL
                # was read (and so written) after the previous one
                nl, nr = ctx.get_many(nv.NL, 5), ctx.get_many(nv.NR, 5)
                inner = (slice(3, 23), slice(2, 26), slice(2, 19))
                for q in range(5):
                    assert np.array_equal(nl[q][inner], nr[q][inner]), (rep, q)
                ctx.set(nv.NL, np.zeros(ctx.shape))   # clobbered: the next replay must restore it
            ctx.set_lusgs_mode(1)
            with pytest.raises(nv.CgliError) as e:   # stale after the schedule change
                g.launch()
            assert e.value.code == -5
        finally:
            g.close()
