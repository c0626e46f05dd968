"""Host-side products with layout-L matrices for the large-N checks (test
infrastructure).  A flat layout-L array a of an N x N matrix holds tile (m,n)
at b*b*(m + MT*n), column-major inside, so a.reshape(MT, MT, b, b)[n, m] is
tile (m,n) transposed.  Every routine walks one tile column at a time, so an
8 GiB matrix is never copied whole.

  matvec(a, N, b, X, part)  : Y = op(A) X for X of shape (N, k) with
      part "full"        A
           "upper"       U: tiles m < n and the upper triangle of tiles (m,m)
           "upper_t"     U^T
           "unit_lower"  L: tiles m > n, the strict lower triangle of tiles
                         (m,m) and a unit diagonal
  inf_norm(a, N, b)         : max row sum of |A|
"""
import numpy as np


def _tiles(a, N, b):
    MT = N // b
    return a.reshape(MT, MT, b, b), MT


def matvec(a, N, b, X, part="full"):
    X = np.asarray(X, dtype=np.float64)
    vec = X.ndim == 1
    X = X.reshape(N, -1)
    k = X.shape[1]
    t, MT = _tiles(a, N, b)
    Xb = X.reshape(MT, b, k)
    Y = np.zeros((MT, b, k))
    for n in range(MT):
        col = t[n]                                     # col[m] = tile (m, n)^T
        if part == "full":
            # Y_m += T_mn X_n  for every m
            Y += np.matmul(col.transpose(0, 2, 1), Xb[n])
        elif part == "upper":
            if n > 0:
                Y[:n] += np.matmul(col[:n].transpose(0, 2, 1), Xb[n])
            Y[n] += np.triu(col[n].T) @ Xb[n]
        elif part == "upper_t":
            # Y_n = sum_{m <= n} T_mn^T X_m
            if n > 0:
                Y[n] += np.einsum("mcr,mrk->ck", col[:n], Xb[:n], optimize=True)
            Y[n] += np.triu(col[n].T).T @ Xb[n]
        elif part == "unit_lower":
            if n + 1 < MT:
                Y[n + 1:] += np.matmul(col[n + 1:].transpose(0, 2, 1), Xb[n])
            Y[n] += (np.tril(col[n].T, -1) + np.eye(b)) @ Xb[n]
        else:
            raise ValueError(part)
    Y = Y.reshape(N, k)
    return Y[:, 0] if vec else Y


def inf_norm(a, N, b):
    t, MT = _tiles(a, N, b)
    rows = np.zeros((MT, b))
    for n in range(MT):
        rows += np.abs(t[n]).sum(axis=1)               # sum over the tile's columns c
    return rows.max()
