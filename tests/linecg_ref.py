"""NumPy restatement of the fast path's pressure solver, for its GPU tests:
CG with the reference's L1 stopping test, cap and breakdown guards
(LinearSystem.cpp:118-210), preconditioned by the exact inverse of each
wall-normal column's tridiagonal part (Thomas algorithm), on the pressure
system in the device's cell-row form (diagonal, the coupling with the cell
above, with the cell to the right, and the wake coupling of row 0).

The preconditioner has no counterpart in the reference; this module pins the
HIP kernels to their definition, not to the reference."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def matrix(dP, pN, pE, pW, ni, nj, partner_cell):
    """The pressure matrix from the device rows (symmetric in the interior;
    the wake couplings are per row)."""
    n = ni * nj
    c = np.arange(n)
    i, j = c % ni, c // ni
    rows, cols, vals = [c], [c], [dP]
    m = j > 0
    rows.append(c[m]); cols.append(c[m] - ni); vals.append(pN[c[m] - ni])
    m = (j == 0) & (partner_cell[i] >= 0)
    rows.append(c[m]); cols.append(partner_cell[i[m]]); vals.append(pW[c[m]])
    m = j < nj - 1
    rows.append(c[m]); cols.append(c[m] + ni); vals.append(pN[c[m]])
    m = i > 0
    rows.append(c[m]); cols.append(c[m] - 1); vals.append(pE[c[m] - 1])
    m = i < ni - 1
    rows.append(c[m]); cols.append(c[m] + 1); vals.append(pE[c[m]])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def line_solve(dP, pN, r, ni, nj):
    """z = M^-1 r, M = blockdiag over columns of tridiag(pN below, dP, pN above)."""
    d = dP.reshape(nj, ni)
    u = pN.reshape(nj, ni)
    rr = r.reshape(nj, ni)
    y = np.zeros_like(d)
    inv = np.zeros_like(d)
    cp = np.zeros(ni)
    yp = np.zeros(ni)
    for jj in range(nj):
        lo = u[jj - 1] if jj > 0 else np.zeros(ni)
        den = d[jj] - lo * cp
        inv[jj] = np.where(np.abs(den) > 0, 1.0 / np.where(den == 0, 1.0, den), 1.0)
        y[jj] = (rr[jj] - lo * yp) * inv[jj]
        cp = u[jj] * inv[jj]
        yp = y[jj]
    z = np.zeros_like(d)
    zn = np.zeros(ni)
    for jj in range(nj - 1, -1, -1):
        z[jj] = y[jj] - (u[jj] * inv[jj]) * zn
        zn = z[jj]
    return z.reshape(-1)


def line_cg(A, dP, pN, b, ni, nj, tol, cap):
    """Returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    n0 = np.abs(r).sum()
    if not n0 > 0:
        return x, 0
    z = line_solve(dP, pN, r, ni, nj)
    p = z.copy()
    rz = r @ z
    for it in range(cap):
        ap = A @ p
        pap = p @ ap
        if not abs(pap) > 0:
            return x, it
        a = rz / pap
        x += a * p
        r -= a * ap
        if np.abs(r).sum() <= tol * n0:
            return x, it + 1
        z = line_solve(dP, pN, r, ni, nj)
        rzn = r @ z
        if not abs(rz) > 0:
            return x, it + 1
        p = z + (rzn / rz) * p
        rz = rzn
    return x, cap


def refined_line_cg(A, dP, pN, b, ni, nj, tol, cap, passes):
    """FP64-residual iterative refinement: pass k solves A d = b - A x_k."""
    x = np.zeros_like(b)
    rhs = b
    for k in range(passes):
        d, _ = line_cg(A, dP, pN, rhs, ni, nj, tol, cap)
        x = x + d if k else d
        rhs = b - A @ x
    return x
