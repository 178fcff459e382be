"""GPU: adjoint of the alpha values (pac_alpha_adjoint) and its torch
binding, with the combinatorics held fixed.

Checked three ways:
  - against an FP64 reference gradient built independently (orthocentre
    by numpy.linalg.solve, tests/adjoint_ref.py), to 1e-9 relative;
  - against central finite differences of the FP64 reference loss (step
    1e-6 A) on 256 random coordinates, to 1e-6 relative;
  - through torch.autograd: loss.backward() equals the adjoint.
Loss: L = sum of rho over the triangles and tets of the complex (rho, the
squared orthoradius, is the alpha value).
"""
import sys

import numpy as np
import pytest

import pac
import adjoint_ref as ref

pytestmark = pytest.mark.gpu

FD_STEP = 1e-6
FD_REL = 1e-6
REF_REL = 1e-9


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(xyzr, solvent, compute):
    res = compute(xyzr, solvent, 0.0)
    L = res.lists()
    lists = {2: L[2].astype(np.int64), 3: L[3].astype(np.int64)}
    pos, r = ref.balls_f64(xyzr, solvent)
    return res, lists, pos, r


def _check_gradients(res, lists, pos, r, seed, what):
    T, K = len(lists[2]), len(lists[3])
    assert T > 0 and K > 0
    # alpha values agree with the independent FP64 reference
    for dim in (2, 3):
        g_rho = res.alpha_values(dim)
        r_rho = ref.rho(pos, r, lists[dim])
        scale = np.abs(r_rho) + r[lists[dim]].max(axis=1) ** 2
        assert (np.abs(g_rho - r_rho) <= 1e-9 * scale).all(), what
    ones = {2: np.ones(T), 3: np.ones(K)}
    dxyz, dr = res.adjoint(ones[2], ones[3])
    rp, rr = ref.gradient(pos, r, lists, ones)
    scale = max(np.abs(rp).max(), np.abs(rr).max())
    assert np.abs(dxyz - rp).max() <= REF_REL * scale, what
    assert np.abs(dr - rr).max() <= REF_REL * scale, what
    # central differences on 256 random coordinates (x, y, z or r)
    rng = np.random.default_rng(seed)
    in_complex = np.unique(np.concatenate([lists[2].ravel(), lists[3].ravel()]))
    balls = rng.choice(in_complex, 256, replace=False)
    axes = rng.integers(0, 4, 256)
    worst = 0.0
    for b, a in zip(balls, axes):
        def loss(delta):
            p2, r2 = pos.copy(), r.copy()
            if a < 3:
                p2[b, a] += delta
            else:
                r2[b] += delta
            return ref.local_loss(p2, r2, lists, b)
        fd = (loss(FD_STEP) - loss(-FD_STEP)) / (2 * FD_STEP)
        g = dxyz[b, a] if a < 3 else dr[b]
        err = abs(fd - g) / max(abs(g), 1.0)
        worst = max(worst, err)
    _say("%s: max FD relative error %.3g over 256 coordinates" % (what, worst))
    assert worst <= FD_REL, what


def test_adjoint_fp32_globule():
    _say("case: adjoint 3000-ball globule (FP32 inputs)")
    b = pac.synth_globule(3000, None, 1.3, seed=4)
    res, lists, pos, r = _setup(b, 1.4, pac.compute)
    _check_gradients(res, lists, pos, r, 4, "fp32 globule")


def test_adjoint_partial_and_zero_inputs():
    _say("case: adjoint with tets only / zero gradients")
    b = pac.synth_globule(1000, 16.84, 1.3, seed=1)
    res, lists, pos, r = _setup(b, 1.4, pac.compute)
    K = len(lists[3])
    g = np.linspace(-1.0, 2.0, K)
    dxyz, dr = res.adjoint(None, g)
    rp, rr = ref.gradient(pos, r, {3: lists[3]}, {3: g})
    scale = max(np.abs(rp).max(), 1.0)
    assert np.abs(dxyz - rp).max() <= REF_REL * scale
    assert np.abs(dr - rr).max() <= REF_REL * scale
    z, zr = res.adjoint(None, None)
    assert not z.any() and not zr.any()


def test_torch_autograd():
    import torch
    from pac.torch_ops import alpha_values
    _say("case: torch autograd through the adjoint")
    b = pac.synth_globule(1500, None, 1.3, seed=7)
    x = torch.tensor(b, requires_grad=True)
    rho_tri, rho_tet, tris, tets = alpha_values(x, 1.4, 0.0)
    loss = rho_tri.sum() + (rho_tet ** 2).sum()
    loss.backward()
    res = pac.compute(b, 1.4, 0.0)
    dxyz, dr = res.adjoint(np.ones(len(tris)), 2.0 * res.alpha_values(3))
    gx = x.grad.double().numpy()
    scale = max(np.abs(dxyz).max(), 1.0)
    assert np.abs(gx[:, :3] - dxyz).max() <= 1e-5 * scale  # grad stored in FP32
    assert np.abs(gx[:, 3] - dr).max() <= 1e-5 * scale


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_torch_autograd_device_tensors(dtype):
    """a GPU tensor stays on the device: the lists, rho and the gradient
    come back as device tensors equal to the host path's (same kernels,
    same order: bit-exact)"""
    import torch
    from pac.torch_ops import alpha_values
    _say("case: torch autograd on device tensors (%s)" % dtype)
    if dtype == "float64":
        b = pac.synth_globule_f64(2500, None, 1.3, seed=8, sigma=0.01)
        res = pac.compute_f64(b, 1.4, 0.0)
    else:
        b = pac.synth_globule(2500, None, 1.3, seed=8)
        res = pac.compute(b, 1.4, 0.0)
    x = torch.tensor(b, device="cuda:0", requires_grad=True)
    rho_tri, rho_tet, tris, tets = alpha_values(x, 1.4, 0.0)
    for t in (rho_tri, rho_tet, tris, tets):
        assert t.is_cuda
    L = res.lists()
    assert np.array_equal(tris.cpu().numpy(), L[2])
    assert np.array_equal(tets.cpu().numpy(), L[3])
    assert np.array_equal(rho_tri.detach().cpu().numpy(), res.alpha_values(2))
    assert np.array_equal(rho_tet.detach().cpu().numpy(), res.alpha_values(3))
    loss = rho_tri.sum() + (rho_tet ** 2).sum()
    loss.backward()
    assert x.grad.is_cuda
    dxyz, dr = res.adjoint(np.ones(len(L[2])), 2.0 * rho_tet.detach().cpu().numpy())
    gx = x.grad.double().cpu().numpy()
    if dtype == "float64":
        assert np.array_equal(gx[:, :3], dxyz) and np.array_equal(gx[:, 3], dr)
    else:
        scale = max(np.abs(dxyz).max(), 1.0)
        assert np.abs(gx[:, :3] - dxyz).max() <= 1e-5 * scale  # grad stored in FP32
        assert np.abs(gx[:, 3] - dr).max() <= 1e-5 * scale


def test_adjoint_c4_fp64():
    """the differentiation workload: 20,000 balls, continuous FP64 centres
    with N(0, 0.01 A) jitter, solvent 1.4, alpha 0; FD on 256 coordinates.
    An FP32-forward variant (the same balls rounded to FP32) reports its
    gradient error against the FP64 one."""
    _say("case: adjoint C4 20k FP64")
    b = pac.synth_globule_f64(20000, 45.7, 1.3, seed=4, sigma=0.01)
    res, lists, pos, r = _setup(b, 1.4, pac.compute_f64)
    _check_gradients(res, lists, pos, r, 44, "C4 fp64")
    g64 = res.adjoint(np.ones(len(lists[2])), np.ones(len(lists[3])))
    b32 = b.astype(np.float32)
    r32 = pac.compute(b32, 1.4, 0.0)
    L32 = r32.lists()
    c32 = r32.counts()
    g32 = r32.adjoint(np.ones(int(c32[2])), np.ones(int(c32[3])))
    # balls whose incident triangles / tets are the same in both complexes
    touched = set()
    for d in (2, 3):
        a = {tuple(sorted(t)) for t in lists[d].tolist()}
        c = {tuple(sorted(t)) for t in L32[d].tolist()}
        for t in a ^ c:
            touched.update(t)
    keep = np.ones(len(b), dtype=bool)
    keep[list(touched)] = False
    G64 = np.concatenate([g64[0], g64[1][:, None]], 1)[keep]
    G32 = np.concatenate([g32[0], g32[1][:, None]], 1)[keep]
    rel = np.linalg.norm(G32 - G64) / np.linalg.norm(G64)
    _say("C4 FP32-forward gradient relative error vs FP64: %.3e over %d balls "
         "with identical incident simplices (%d balls touched by the %s vs %s "
         "complex difference)" % (rel, keep.sum(), len(touched),
                                   list(res.counts()), list(c32)))
    assert rel < 1e-4
