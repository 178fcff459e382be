"""FP64 reference for alpha values and their gradients (test infrastructure).

Independent of the closed forms the kernels use: the orthocentre c of a
simplex is the point of its affine hull with equal power to every vertex,
found by solving the radical-plane system with numpy.linalg.solve; its
barycentric coordinates lam come from the same solve.  The gradient of
rho = |c - p_q|^2 - w_q with the combinatorics fixed is
d rho/d p_q = 2 lam_q (p_q - c), d rho/d r_q = -2 lam_q r_q.
"""
import numpy as np


def ortho(P, w):
    """P: (m, k+1, 3), w: (m, k+1) -> c (m, 3), lam (m, k+1), rho (m)"""
    p0 = P[:, 0, :]
    U = P[:, 1:, :] - p0[:, None, :]                       # (m, k, 3)
    b = (U * U).sum(-1) + w[:, :1] - w[:, 1:]               # (m, k)
    G = 2.0 * np.einsum("mia,mja->mij", U, U)               # (m, k, k)
    t = np.linalg.solve(G, b[..., None])[..., 0]            # (m, k)
    x = np.einsum("mi,mia->ma", t, U)
    c = p0 + x
    lam = np.concatenate([1.0 - t.sum(-1, keepdims=True), t], axis=1)
    rho = (x * x).sum(-1) - w[:, 0]
    return c, lam, rho


def balls_f64(xyzr, solvent):
    """ball centres and radii as the pipeline holds them: FP32 inputs keep
    r = fl32(r_raw + s); FP64 inputs r = r_raw + s"""
    xyzr = np.asarray(xyzr)
    if xyzr.dtype == np.float32:
        r = (xyzr[:, 3] + np.float32(solvent)).astype(np.float64)
    else:
        r = xyzr[:, 3].astype(np.float64) + solvent
    return xyzr[:, :3].astype(np.float64), r


def rho(pos, r, rows):
    P = pos[rows]
    w = r[rows] ** 2
    return ortho(P, w)[2]


def gradient(pos, r, lists, g):
    """dL/dpos (n, 3), dL/dr (n) for L with dL/drho = g[dim] per row"""
    n = len(pos)
    gp = np.zeros((n, 3))
    gr = np.zeros(n)
    for dim, rows in lists.items():
        if len(rows) == 0:
            continue
        P = pos[rows]
        w = r[rows] ** 2
        c, lam, _ = ortho(P, w)
        k = 2.0 * lam * g[dim][:, None]                      # (m, k+1)
        np.add.at(gp, rows.ravel(), (k[..., None] * (P - c[:, None, :])).reshape(-1, 3))
        np.add.at(gr, rows.ravel(), (-k * r[rows]).ravel())
    return gp, gr


def local_loss(pos, r, lists, ball, weights=None):
    """sum of rho over the rows of `lists` that contain `ball` (the rows that
    a perturbation of that ball changes)"""
    s = 0.0
    for dim, rows in lists.items():
        sel = (rows == ball).any(axis=1)
        if sel.any():
            v = rho(pos, r, rows[sel])
            s += v.sum() if weights is None else (weights[dim][sel] * v).sum()
    return s
