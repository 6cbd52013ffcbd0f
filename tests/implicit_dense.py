"""Dense block form of the implicit operator of the dual-time inner iteration.

TEST INFRASTRUCTURE ONLY.  Assembles, in numpy, the 5x5-block matrices of
the scalar-diagonal LU-SGS / DP-LUR operator on the interior cells of one
slab, from the split-flux formulas of the reference (DPLUSGS.cpp:585-990 for
the lower and upper neighbour blocks, Dual_time_stepping.cpp:1054-1170 and
1384-1524 for the same blocks in the hyperplane sweeps) and the diagonal of
Dual_time_stepping.cpp:853-895:

  forward   NL_c = D_c (MR_c - sum_lower Lblk(c,n) NL_n)
  backward  NR_c = NL_c - D_c sum_upper Ublk(c,n) NR_n
  Jacobi    x'_c = D_c (MR_c - sum_lower Lblk(c,n) x_n - sum_upper Ublk(c,n) x_n)

so that the LU-SGS pair solves (D^-1 + L) D (D^-1 + U) x = MR and the Jacobi
iteration converges (where it does) to (D^-1 + L + U) x = MR.  Increments of
cells outside the interior are zero (Dual_time_stepping.cpp:198-323), so
those neighbours drop out.
"""
import numpy as np

MET = ["xidx", "xidy", "xidz", "etdx", "etdy", "etdz", "ztdx", "ztdy", "ztdz"]


def primitives(U_, J, K, e):
    rho = U_[0] * J
    u, v, w = U_[1] / U_[0], U_[2] / U_[0], U_[3] / U_[0]
    VV = u * u + v * v + w * w
    P = (U_[4] * J - 0.5 * rho * VV) * (K - 1)
    C = K * P / rho            # c^2 (no square root), as in the reference
    T = P / rho
    beta = np.maximum(VV / C, e)
    return dict(rho=rho, u=u, v=v, w=w, P=P, C=C, T=T, beta=beta)


def split_blocks(q, mx, my, mz, K):
    """Plus-split block P and the backward block M of one face direction at
    one cell's state and metric row (DPLUSGS.cpp:612-634, 836-866)."""
    b, rho, P, T, C = q["beta"], q["rho"], q["P"], q["T"], q["C"]
    Un = mx * q["u"] + my * q["v"] + mz * q["w"]
    S = np.sqrt(Un * Un * (b - 1) ** 2 + 4 * b * C * (mx * mx + my * my + mz * mz))
    d = abs(((b + 1) * abs(Un) + S) / 2)
    p11, p22 = 0.5 * (b * Un + d), 0.5 * (Un + d)
    Pb = np.zeros((5, 5))
    Pb[0, :4] = [p11, 0.5 * P * K * b * mx, 0.5 * P * K * b * my, 0.5 * P * K * b * mz]
    Pb[1, 0], Pb[1, 1] = 0.5 * mx / rho, p22
    Pb[2, 0], Pb[2, 2] = 0.5 * my / rho, p22
    Pb[3, 0], Pb[3, 3] = 0.5 * mz / rho, p22
    Pb[4] = [0.5 * Un * (K - 1) * (b - 1) / rho / K, 0.5 * mx * b * T * (K - 1), 0.5 * my * b * T * (K - 1),
             0.5 * mz * b * T * (K - 1), p22]
    Mb = Pb.copy()
    Mb[0, 0] = 0.5 * (b * Un - d)
    for r in range(1, 5):
        Mb[r, r] = 0.5 * (Un - d)
    return Pb, Mb


def diag_block(q, m, deltaT, dxi, deta, dzeta, K):
    """D_c: the inverse-diagonal action x = D r of Dual_time_stepping.cpp:887-895."""
    b, rho = q["beta"], q["rho"]
    lam = 0.0
    for d, dd in enumerate((dxi, deta, dzeta)):
        mx, my, mz = m[3 * d], m[3 * d + 1], m[3 * d + 2]
        Un = mx * q["u"] + my * q["v"] + mz * q["w"]
        S = np.sqrt(Un * Un * (b - 1) ** 2 + 4 * b * q["C"] * (mx * mx + my * my + mz * mz))
        lam += ((b + 1) * abs(Un) + S) / 2 / dd
    d11 = 2 * deltaT / (3 * b + 2 * lam * deltaT)
    d22 = 2 * deltaT / (3 + 2 * lam * deltaT)
    d51 = -6 * (K - 1) * (b - 1) * deltaT / (K * (3 + 2 * lam * deltaT) * (3 * b + 2 * lam * deltaT) * rho)
    D = np.diag([d11, d22, d22, d22, d22])
    D[4, 0] = d51
    return D


class DenseOperator:
    """D (block diagonal), L and U over the interior cells (i outer, j, k
    inner), for the state U_ and metrics of one slab in the reference layout."""

    def __init__(self, U_, J, met, deltaT, e, dxi, deta, dzeta, K=1.4):
        nxp, Ym, Zm = J.shape
        self.cells = [(i, j, k) for i in range(3, nxp - 3) for j in range(2, Ym - 2) for k in range(2, Zm - 2)]
        index = {c: n for n, c in enumerate(self.cells)}
        N = len(self.cells)
        self.D = np.zeros((5 * N, 5 * N))
        self.L = np.zeros((5 * N, 5 * N))
        self.U = np.zeros((5 * N, 5 * N))
        q = primitives(U_, J, K, e)
        at = lambda c: {k: v[c] for k, v in q.items()}
        deltas = (dxi, deta, dzeta)
        for n, c in enumerate(self.cells):
            mc = [met[t][c] for t in range(9)]
            self.D[5 * n:5 * n + 5, 5 * n:5 * n + 5] = diag_block(at(c), mc, deltaT, dxi, deta, dzeta, K)
            for d in range(3):
                off = [0, 0, 0]
                off[d] = 1
                lo = (c[0] - off[0], c[1] - off[1], c[2] - off[2])
                up = (c[0] + off[0], c[1] + off[1], c[2] + off[2])
                if lo in index:
                    m = [met[3 * d + t][lo] for t in range(3)]
                    Pb, _ = split_blocks(at(lo), *m, K)
                    a = index[lo]
                    self.L[5 * n:5 * n + 5, 5 * a:5 * a + 5] = -Pb / deltas[d]
                if up in index:
                    m = [met[3 * d + t][up] for t in range(3)]
                    _, Mb = split_blocks(at(up), *m, K)
                    a = index[up]
                    self.U[5 * n:5 * n + 5, 5 * a:5 * a + 5] = Mb / deltas[d]
        self.Dinv = np.linalg.inv(self.D)

    def gather(self, fields):
        """interior values of 5 fields as one vector (cell-major, 5 per cell)"""
        return np.array([[f[c] for f in fields] for c in self.cells]).reshape(-1)

    def lusgs_operator(self):
        return (self.Dinv + self.L) @ self.D @ (self.Dinv + self.U)

    def jacobi(self, MR, sweeps):
        x = self.D @ MR
        for _ in range(sweeps):
            x = self.D @ (MR - (self.L + self.U) @ x)
        return x
