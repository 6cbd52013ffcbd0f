"""The restart file written and read by the library against the oracle.

cgli_write_restart / cgli_read_restart (Output.cpp:167-284,
Initial_condition.cpp:112-219) move U*J and Uq*J through the device; the file
must equal, byte for byte, what the oracle's restatement (pinned against the
reference's own Output and Initial_condition on the CPU) writes from the same
downloaded fields, and the fields read back must equal the oracle's Q/J
exactly.  A restarted channel then steps on like the uninterrupted one, to
rounding: U*J/J need not return U's last bit, in the reference as here.
"""
import os
import sys

import numpy as np
import pytest

from gpu_harness import rel

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_restart as orr  # noqa: E402

pytestmark = pytest.mark.gpu


def test_channel_restart_round_trip(tmp_path):
    from cgli_fu_flow_amd import native as nv
    nx, ny, nz = 16, 24, 12
    print("restart round trip", (nx, ny, nz), flush=True)
    ch = nv.channel(180.0, beta=2.2, skew=0.05, xi_mode=1)
    cfg = nv.channel_config(ch, nx, ny, nz)
    path = str(tmp_path / "LES00002.bin")
    with nv.Context(cfg) as a, nv.Context(cfg) as b:
        dt = a.channel_init(ch)
        for _ in range(2):
            a.channel_step(dt, tol=0.0, max_inner=3)
        U, Uq, J = a.get_many(nv.U, 5), a.get_many(nv.UQ, 5), a.get(nv.J)
        a.write_restart(path)
        orr.write(str(tmp_path / "oracle.bin"), U, Uq, J, nx, ny, nz)
        assert open(path, "rb").read() == open(str(tmp_path / "oracle.bin"), "rb").read()

        b.channel_init(ch)
        b.read_restart(path)
        Uo, Uqo = orr.read(path, J, nx, ny, nz)
        U_b, Ub, Uqb = b.get_many(nv.U_, 5), b.get_many(nv.U, 5), b.get_many(nv.UQ, 5)
        for q in range(5):
            np.testing.assert_array_equal(orr.interior(Ub[q], nx, ny, nz), Uo[q])
            np.testing.assert_array_equal(orr.interior(U_b[q], nx, ny, nz), Uo[q])
            np.testing.assert_array_equal(orr.interior(Uqb[q], nx, ny, nz), Uqo[q])
            # Q/J returns U to within one rounding of the product and quotient
            assert rel(orr.interior(Ub[q], nx, ny, nz), orr.interior(U[q], nx, ny, nz)) <= 1e-15

        for ctx in (a, b):
            ctx.channel_step(dt, tol=0.0, max_inner=3)
        for q in range(5):
            err = rel(b.get(nv.U + q), a.get(nv.U + q))
            assert err <= 1e-12, (q, err)


def test_slab_writes_its_part_of_every_block(tmp_path):
    """A middle slab (gstart 5 of 20 planes) updates only its cells of the ten
    blocks of an existing file and truncates nothing."""
    from cgli_fu_flow_amd import native as nv
    nx, ny, nz, nxg, gstart = 8, 10, 6, 20, 5
    print("restart slab", (nx, ny, nz), "gstart", gstart, flush=True)
    cfg = nv.reference_config(nx, ny, nz)
    cfg.first, cfg.last, cfg.gstart, cfg.gend0, cfg.nx_global = 0, 0, gstart, gstart + nx - 1, nxg
    rng = np.random.default_rng(11)
    shape = (nx + 6, ny + 4, nz + 4)
    U = [rng.uniform(0.5, 2.0, shape) for _ in range(5)]
    Uq = [rng.uniform(0.5, 2.0, shape) for _ in range(5)]
    J = rng.uniform(0.7, 1.3, shape)
    pre = np.full(10 * nxg * ny * nz + 17, -3.25)   # longer than the blocks
    path, opath = str(tmp_path / "LES00001.bin"), str(tmp_path / "oracle.bin")
    pre.tofile(path)
    pre.tofile(opath)
    with nv.Context(cfg) as ctx:
        ctx.set_many(nv.U, U)
        ctx.set_many(nv.UQ, Uq)
        ctx.set(nv.J, J)
        ctx.write_restart(path)
        orr.write(opath, U, Uq, J, nx, ny, nz, gstart=gstart, nx_global=nxg)
        assert open(path, "rb").read() == open(opath, "rb").read()

        ctx.set_many(nv.U, [np.zeros(shape)] * 5)
        ctx.read_restart(path)
        Uo, _ = orr.read(path, J, nx, ny, nz, gstart=gstart, nx_global=nxg)
        for q in range(5):
            np.testing.assert_array_equal(orr.interior(ctx.get(nv.U + q), nx, ny, nz), Uo[q])

        # a file that ends inside the last block, and a missing file
        short = str(tmp_path / "short.bin")
        pre[: 9 * nxg * ny * nz + gstart * ny * nz + 3].tofile(short)
        for bad in (short, str(tmp_path / "missing.bin")):
            with pytest.raises(nv.CgliError) as e:
                ctx.read_restart(bad)
            assert e.value.code == -8
