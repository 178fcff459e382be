"""torch.autograd binding of the alpha values and their adjoint.

`alpha_values(xyzr, solvent, alpha)` runs the pipeline on the balls
(n x 4 tensor: x, y, z, r_raw; float32 through pac_compute, float64
through pac_compute_f64), returns the FP64 alpha values rho of the output
triangles and tets (and the lists themselves), and differentiates rho with
respect to the balls through pac_alpha_adjoint with the combinatorics held
fixed.  r = r_raw + solvent, so dL/dr_raw = dL/dr.
"""
import numpy as np
import torch

from . import compute, compute_f64


class AlphaValues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyzr, solvent, alpha, kw):
        x = xyzr.detach().cpu().numpy()
        if xyzr.dtype == torch.float64:
            res = compute_f64(x, solvent, alpha, **kw)
        else:
            res = compute(x.astype(np.float32), solvent, alpha, **kw)
        lists = res.lists()
        rho_tri = torch.from_numpy(res.alpha_values(2)).to(xyzr.device)
        rho_tet = torch.from_numpy(res.alpha_values(3)).to(xyzr.device)
        ctx.res = res
        ctx.in_dtype = xyzr.dtype
        tris = torch.from_numpy(lists[2]).to(xyzr.device)
        tets = torch.from_numpy(lists[3]).to(xyzr.device)
        ctx.mark_non_differentiable(tris, tets)
        return rho_tri, rho_tet, tris, tets

    @staticmethod
    def backward(ctx, g_tri, g_tet, _g_tris, _g_tets):
        gt = None if g_tri is None else g_tri.detach().double().cpu().numpy()
        gk = None if g_tet is None else g_tet.detach().double().cpu().numpy()
        dxyz, dr = ctx.res.adjoint(gt, gk)
        g = np.concatenate([dxyz, dr[:, None]], axis=1)
        dev = g_tri.device if g_tri is not None else g_tet.device
        return torch.from_numpy(g).to(dev, ctx.in_dtype), None, None, None


def alpha_values(xyzr, solvent=1.4, alpha=0.0, **kw):
    """(rho_tri, rho_tet, tris, tets); rho_* are differentiable in xyzr."""
    return AlphaValues.apply(xyzr, solvent, alpha, kw)
