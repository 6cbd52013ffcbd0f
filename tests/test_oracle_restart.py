"""The restart file's CPU restatement (oracle/oracle_restart.py) against the
reference's own writer and reader.

Output.cpp:167-284 and Initial_condition.cpp:112-219, compiled from the
reference's sources into oracle/_ref (where the reference tree was present at
build time), write and read LES%05d.bin / initial.bin in the current
directory; each test runs them inside its own temporary directory.  The
layout check needs no reference build: every value written encodes its
block and cell, so a swapped index or block order fails it.
"""
import ctypes as C
import glob
import os
import sys

import numpy as np
import pytest

from oracle_harness import REF_DIR, dp, load_oracle, ptrs, ref_cfg_from

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_restart as orr  # noqa: E402

SEED = 20261203


def division(X, NP, myid):
    """Mpi_division.h: (gstart, gcount) of rank myid."""
    Xdv = X // NP
    Xr = X - Xdv * NP
    if myid < Xr:
        return myid * (Xdv + 1), Xdv + 1
    return myid * Xdv + Xr, Xdv


def seeded_state(orc, shape, seed):
    """Positive U, Uq and J fields (the values need not be physical)."""
    def fill(s, lo, hi):
        a = np.empty(shape)
        orc.orc_fill(dp(a), a.size, s, lo, hi)
        return a
    U = [fill(seed + q, 0.5, 2.0) for q in range(5)]
    Uq = [fill(seed + 10 + q, 0.5, 2.0) for q in range(5)]
    J = fill(seed + 20, 0.7, 1.3)
    return U, Uq, J


def test_layout_encodes_block_and_cell(tmp_path):
    nx, ny, nz = 4, 6, 5
    shape = (nx + 6, ny + 4, nz + 4)
    i, j, k = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    code = lambda b: b * 1e6 + i * 1e4 + j * 1e2 + k
    U = [code(b) for b in range(5)]
    Uq = [code(b) for b in range(5, 10)]
    J = np.ones(shape)
    nxg, gstart = 9, 3   # a middle slab of a 9-plane problem
    path = str(tmp_path / "LES00001.bin")
    orr.write(path, U, Uq, J, nx, ny, nz, gstart=gstart, nx_global=nxg)
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == 9 * nxg * ny * nz + (gstart + nx) * ny * nz   # ends with the last block's part
    full = np.zeros(10 * nxg * ny * nz)
    full[:raw.size] = raw
    blocks = full.reshape(10, nxg, ny, nz)
    for b in range(10):
        want = b * 1e6 + (np.arange(nx)[:, None, None] + 3) * 1e4 + (np.arange(ny)[None, :, None] + 2) * 1e2 + \
            (np.arange(nz)[None, None, :] + 2)
        np.testing.assert_array_equal(blocks[b, gstart:gstart + nx], want)
        assert not blocks[b, :gstart].any() and not blocks[b, gstart + nx:].any()
    Ur, Uqr = orr.read(path, J, nx, ny, nz, gstart=gstart, nx_global=nxg)
    for b in range(5):
        np.testing.assert_array_equal(Ur[b], orr.interior(U[b], nx, ny, nz))
        np.testing.assert_array_equal(Uqr[b], orr.interior(Uq[b], nx, ny, nz))
    with pytest.raises(EOFError):
        orr.read(path, J, nx + 1, ny, nz, gstart=gstart, nx_global=nxg)


def _ref_cases():
    out = []
    for path in sorted(glob.glob(os.path.join(REF_DIR, "libref_*.so"))):
        tag = os.path.basename(path)[len("libref_"):-3]
        nproc = int(tag.split("_np")[1].split("_")[0])
        for myid in sorted({0, nproc // 2}):
            out.append(pytest.param(path, myid, id="%s-rank%d" % (tag, myid)))
    return out or [pytest.param(None, 0, id="no-reference-build")]


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


@pytest.mark.parametrize("path,myid", _ref_cases())
def test_oracle_matches_reference_restart_io(orc, tmp_path, monkeypatch, path, myid):
    if path is None:
        pytest.skip("reference build not present")
    ref = C.CDLL(path)
    ref.ref_init()
    ref.ref_output.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    ref.ref_initial_condition.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 4
    (X, Y, Z, Ym, Zm, NP, _, _), _ = ref_cfg_from(ref)
    gstart, nx = division(X, NP, myid)
    shape = (nx + 6, Ym, Zm)
    U, Uq, J = seeded_state(orc, shape, SEED + 31 * myid)
    monkeypatch.chdir(tmp_path)
    step = 7
    ref.ref_output(nx + 6, myid, step, C.cast(ptrs(U), C.c_void_p), C.cast(ptrs(Uq), C.c_void_p),
                   C.cast(dp(J), C.c_void_p))
    ref_bytes = (tmp_path / ("LES%05d.bin" % step)).read_bytes()
    orr.write(str(tmp_path / "oracle.bin"), U, Uq, J, nx, Y, Z, gstart=gstart, nx_global=X)
    assert (tmp_path / "oracle.bin").read_bytes() == ref_bytes

    # read back: the reference from initial.bin, the oracle from the same bytes
    (tmp_path / "initial.bin").write_bytes(ref_bytes)
    U_r = [np.zeros(shape) for _ in range(5)]
    U_n = [np.zeros(shape) for _ in range(5)]
    Uq_r = [np.zeros(shape) for _ in range(5)]
    ref.ref_initial_condition(nx + 6, myid, C.cast(ptrs(U_r), C.c_void_p), C.cast(ptrs(U_n), C.c_void_p),
                              C.cast(ptrs(Uq_r), C.c_void_p), C.cast(dp(J), C.c_void_p))
    Uo, Uqo = orr.read(str(tmp_path / "initial.bin"), J, nx, Y, Z, gstart=gstart, nx_global=X)
    for q in range(5):
        np.testing.assert_array_equal(orr.interior(U_n[q], nx, Y, Z), Uo[q])
        np.testing.assert_array_equal(orr.interior(U_r[q], nx, Y, Z), Uo[q])
        np.testing.assert_array_equal(orr.interior(Uq_r[q], nx, Y, Z), Uqo[q])
