"""torch.autograd binding of the alpha values and their adjoint.

`alpha_values(xyzr, solvent, alpha)` runs the pipeline on the balls
(n x 4 tensor: x, y, z, r_raw; float32 through pac_compute, float64
through pac_compute_f64), returns the FP64 alpha values rho of the output
triangles and tets (and the lists themselves), and differentiates rho with
respect to the balls through pac_alpha_adjoint with the combinatorics held
fixed.  r = r_raw + solvent, so dL/dr_raw = dL/dr.

A GPU tensor stays on the device: the library reads it in place
(device_inputs, on torch's current stream), writes the lists and rho into
device tensors, and the backward pass hands device gradient buffers to
pac_alpha_adjoint.  A CPU tensor goes through host memory.
"""
import numpy as np
import torch

from . import compute, compute_device, compute_device_f64, compute_f64


def _device_forward(xyzr, solvent, alpha, kw):
    x = xyzr.detach().contiguous()
    dev = x.device
    stream = torch.cuda.current_stream(dev)
    opts = dict(kw, device=dev.index if dev.index is not None else 0,
                stream=stream.cuda_stream)
    n = x.shape[0]
    if x.dtype == torch.float64:
        res = compute_device_f64(x.data_ptr(), n, solvent, alpha, **opts)
    else:
        x = x.float()
        res = compute_device(x.data_ptr(), n, solvent, alpha, **opts)
    c = res.counts()
    tris = torch.empty((int(c[2]), 3), dtype=torch.int32, device=dev)
    tets = torch.empty((int(c[3]), 4), dtype=torch.int32, device=dev)
    rho_tri = torch.empty(int(c[2]), dtype=torch.float64, device=dev)
    rho_tet = torch.empty(int(c[3]), dtype=torch.float64, device=dev)
    # the library copies on its own stream and waits: the buffers torch
    # just allocated must exist before it writes them
    stream.synchronize()
    res.copy_list_to(2, tris.data_ptr())
    res.copy_list_to(3, tets.data_ptr())
    res.copy_alpha_to(2, rho_tri.data_ptr())
    res.copy_alpha_to(3, rho_tet.data_ptr())
    return res, rho_tri, rho_tet, tris, tets


class AlphaValues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyzr, solvent, alpha, kw):
        if xyzr.is_cuda:
            res, rho_tri, rho_tet, tris, tets = _device_forward(xyzr, solvent,
                                                                alpha, kw)
        else:
            x = xyzr.detach().numpy()
            if xyzr.dtype == torch.float64:
                res = compute_f64(x, solvent, alpha, **kw)
            else:
                res = compute(x.astype(np.float32), solvent, alpha, **kw)
            lists = res.lists()
            rho_tri = torch.from_numpy(res.alpha_values(2))
            rho_tet = torch.from_numpy(res.alpha_values(3))
            tris = torch.from_numpy(lists[2])
            tets = torch.from_numpy(lists[3])
        ctx.res = res
        ctx.in_dtype = xyzr.dtype
        ctx.n = xyzr.shape[0]
        ctx.mark_non_differentiable(tris, tets)
        return rho_tri, rho_tet, tris, tets

    @staticmethod
    def backward(ctx, g_tri, g_tet, _g_tris, _g_tets):
        ref = g_tri if g_tri is not None else g_tet
        if ref.is_cuda:
            dev = ref.device
            gt = None if g_tri is None else g_tri.detach().double().contiguous()
            gk = None if g_tet is None else g_tet.detach().double().contiguous()
            dxyz = torch.empty((ctx.n, 3), dtype=torch.float64, device=dev)
            dr = torch.empty(ctx.n, dtype=torch.float64, device=dev)
            # the library runs on its own stream: torch's work first
            torch.cuda.current_stream(dev).synchronize()
            ctx.res.adjoint_to(None if gt is None else gt.data_ptr(),
                               None if gk is None else gk.data_ptr(),
                               dxyz.data_ptr(), dr.data_ptr())
            g = torch.cat([dxyz, dr[:, None]], dim=1)
            return g.to(ctx.in_dtype), None, None, None
        gt = None if g_tri is None else g_tri.detach().double().numpy()
        gk = None if g_tet is None else g_tet.detach().double().numpy()
        dxyz, dr = ctx.res.adjoint(gt, gk)
        g = np.concatenate([dxyz, dr[:, None]], axis=1)
        return torch.from_numpy(g).to(ctx.in_dtype), None, None, None


def alpha_values(xyzr, solvent=1.4, alpha=0.0, **kw):
    """(rho_tri, rho_tet, tris, tets); rho_* are differentiable in xyzr."""
    return AlphaValues.apply(xyzr, solvent, alpha, kw)
