"""CPU restatement of the reference's restart file -- TEST INFRASTRUCTURE ONLY
(tests, the smoke check and the CPU baseline may use it; the library never
does).

Output.cpp:167-284 writes LES%05d.bin: ten blocks, U1..U5 then U1q..U5q,
each ge*Y_out*Z_out doubles with ge = gend0[nproc-1] + 1 = X_out
(Output.cpp:220-226).  A value is U*J (Output.cpp:190-201); the interior cell
(i, j, k) of a rank sits at (i-3)*Y_out*Z_out + (j-2)*Z_out + (k-2) of that
rank's part, which starts gstart*Y_out*Z_out doubles into the block
(Output.cpp:232-279, the MPI_File_set_view displacements).  Native byte
order.  Initial_condition.cpp:112-219 reads it back: U = Q/J, Uq = Q_old/J and
U_ = U on the interior (Initial_condition.cpp:196-212).

Fields are [X_np][Y_m][Z_m] arrays (X_np = nx+6, Y_m = ny+4, Z_m = nz+4).

Pinned (tests/test_oracle_restart.py) against the reference's own Output and
Initial_condition, compiled from its sources, on one rank and on the middle
rank of three, and by a layout check in which every value encodes its block
and cell.
"""
import os

import numpy as np


def interior(a, nx, ny, nz):
    """The interior cells i = 3..gend, j = 2..ny+1, k = 2..nz+1 (a view)."""
    return a[3:nx + 3, 2:ny + 2, 2:nz + 2]


def block_offset(b, nx_global, ny, nz, gstart):
    """Byte offset of a rank's part of block b (Output.cpp:232-279)."""
    return 8 * (b * nx_global * ny * nz + gstart * ny * nz)


def write(path, U, Uq, J, nx, ny, nz, gstart=0, nx_global=None):
    """Write (or update, without truncating) a restart file like Output.cpp."""
    nxg = nx if nx_global is None else nx_global
    Ji = interior(J, nx, ny, nz)
    with open(path, "r+b" if os.path.exists(path) else "w+b") as f:
        for b in range(10):
            src = U[b] if b < 5 else Uq[b - 5]
            q = np.ascontiguousarray(interior(src, nx, ny, nz) * Ji, dtype=np.float64)
            f.seek(block_offset(b, nxg, ny, nz, gstart))
            f.write(q.tobytes())


def read(path, J, nx, ny, nz, gstart=0, nx_global=None):
    """Interior U and Uq (5 arrays each, [nx][ny][nz]) as Initial_condition.cpp
    sets them from the file; U_ equals U."""
    nxg = nx if nx_global is None else nx_global
    Ji = interior(J, nx, ny, nz)
    n = nx * ny * nz
    out = []
    with open(path, "rb") as f:
        for b in range(10):
            f.seek(block_offset(b, nxg, ny, nz, gstart))
            raw = f.read(8 * n)
            if len(raw) != 8 * n:
                raise EOFError("restart file %s ends inside block %d" % (path, b))
            out.append(np.frombuffer(raw, dtype=np.float64).reshape(nx, ny, nz) / Ji)
    return out[:5], out[5:]
