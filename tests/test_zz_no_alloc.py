"""cgli_dplusgs_noalloc against cgli_dplusgs, through the C ABI.

The allocation-free entry point runs the same kernels with the norm's
partial sums in the library's fixed per-device scratch, so its results
equal cgli_dplusgs's bit for bit.  Two calls in a row on one device, the
second on a smaller slab (the scratch keeps the first call's partials), each
against cgli_dplusgs on the same inputs; then the size and missing-field
errors.
"""
import numpy as np
import pytest

from oracle_harness import Case, load_oracle, make_inputs
from gpu_harness import config_from, upload_inputs

pytestmark = pytest.mark.gpu
SEED = 20261203


def dplusgs_once(nv, orc, dims, seed, noalloc):
    X, Y, Z = dims
    case = Case(X, Y, Z, 1, 3, 3, 0, (0.001, 0.005, 0.002))
    f = make_inputs(orc, case.slab, seed)
    with nv.Context(config_from(case.cfg(orc))) as ctx:
        upload_inputs(ctx, f)
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        ctx.viscous()
        # the fields the entry point writes exist before it runs
        zero = np.zeros(ctx.shape)
        for fid in range(nv.MR, nv.RES + 5):
            ctx.set(fid, zero)
        if noalloc:
            er = ctx.dplusgs_noalloc(2e-4, 1e-4, sweeps=4)
        else:
            er = ctx.dplusgs(2e-4, 1e-4, sweeps=4)
        return er, ctx.get_many(nv.U_, 5), ctx.get_many(nv.NR, 5), ctx.get_many(nv.MR, 5)


def test_noalloc_equals_dplusgs_on_two_calls():
    from cgli_fu_flow_amd import native as nv
    orc = load_oracle()
    for dims in ((40, 24, 30), (12, 10, 8)):   # the second call on a smaller slab
        print("dplusgs_noalloc", dims, flush=True)
        want = dplusgs_once(nv, orc, dims, SEED, noalloc=False)
        got = dplusgs_once(nv, orc, dims, SEED, noalloc=True)
        assert np.array_equal(got[0], want[0]), (dims, got[0], want[0])
        for a, b in zip(got[1:], want[1:]):
            for q in range(5):
                assert np.array_equal(a[q], b[q]), dims
        assert np.all(np.isfinite(got[0])) and np.all(got[0] > 0)


def test_noalloc_rejects_what_it_cannot_do():
    from cgli_fu_flow_amd import native as nv
    orc = load_oracle()
    # 300^3 cells exceed the fixed scratch: rejected before anything runs
    case = Case(300, 300, 300, 1, 3, 3, 0, (0.001, 0.005, 0.002))
    with nv.Context(config_from(case.cfg(orc))) as ctx:
        with pytest.raises(nv.CgliError) as e:
            ctx.dplusgs_noalloc(2e-4, 1e-4)
        assert e.value.code == -9
    # outputs never allocated: the entry point does not allocate them
    case = Case(12, 10, 8, 1, 3, 3, 0, (0.001, 0.005, 0.002))
    f = make_inputs(orc, case.slab, SEED)
    with nv.Context(config_from(case.cfg(orc))) as ctx:
        upload_inputs(ctx, f)
        ctx.flux_x(1.0, 1e-4)
        ctx.flux_y(1.0, 1e-4)
        ctx.flux_z(1.0, 1e-4)
        ctx.viscous()
        with pytest.raises(nv.CgliError) as e:
            ctx.dplusgs_noalloc(2e-4, 1e-4)
        assert e.value.code == -5
