"""Gradient-based homography alignment on the differentiable F32 warp.

The HIP entry points (mvx_WarpPerspective_F32_Bilinear, mvx_warp_perspective_f32_adjoint /
_l2, mvx_Remap_F32_Bilinear, mvx_remap_f32_adjoint, mvx_PyramidHalf_F32,
mvx_pyramid_half_f32_adjoint) are wrapped here for batches of F32 images held in torch tensors of
shape (count, height, width); torch only provides the memory and the stream.

Alignment workload: `count` pairs of F32 images in [0, 1].  Each source is the single-channel
scene of ``synth.scene_u8`` (seed 5000 + pair) divided by 255 and smoothed by a Gaussian of sigma
1.5; each target is the source warped by a seeded homography (rotation U(-10, 10) degrees,
translation U(-40, 40) px, perspective U(-2e-4, 2e-4) per px, about the image centre; seed 5100 +
pair) plus N(0, 0.01^2) noise.  The loss is the L2 photometric error summed over a 3-level
pyramid (mean over pixels per level); pyramid levels of source and target are computed once on
the GPU and the warp at level l uses S_l H S_l^-1 with S_l = diag(2^-l, 2^-l, 1).  The 8 free
parameters p describe H in centre-normalised coordinates (x' = (x - c) / s, s = width / 2):
H = N^-1 (I + P) N, P holding p in all entries but the last, so every parameter moves the image
border by about s pixels per unit.  Adam updates p from dL/dp, obtained from the fused
loss-and-gradient kernel's dL/dH by the chain rule.
"""
import ctypes
import math

import numpy as np
import torch

from . import kernels, synth

c = ctypes
LEVELS = 3


def _ptr(t):
    return c.c_void_p(t.data_ptr()) if t is not None else c.c_void_p(None)


def _stream():
    return c.c_void_p(torch.cuda.current_stream().cuda_stream)


def _geom(t):
    """(count, h, w, row stride bytes, batch stride bytes) of a contiguous (count, h, w[, 2]) F32 tensor"""
    assert t.dtype == torch.float32 and t.is_contiguous()
    n, h, w = t.shape[:3]
    per = 2 if t.dim() == 4 else 1
    return n, h, w, 4 * w * per, 4 * w * h * per


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} returned {rc}")


# ------------------------------------------------------------------ kernel wrappers -------------
def warp(src, H, out_hw=None):
    """bilinear F32 warp of every image by its H (count x 9 float64 on the device)"""
    n, sh, sw, ss, sb = _geom(src)
    dh, dw = out_hw or (sh, sw)
    out = torch.empty((n, dh, dw), dtype=torch.float32, device=src.device)
    K = kernels()
    _check(K.mvx_WarpPerspective_F32_Bilinear(_stream(), n, dw, dh, _ptr(out), 4 * dw, c.c_size_t(4 * dw * dh), sw, sh,
                                              _ptr(src), ss, c.c_size_t(sb), _ptr(H)), "mvx_WarpPerspective_F32_Bilinear")
    return out


def _workspace(n, dw, dh):
    K = kernels()
    K.mvx_warp_perspective_f32_workspace_size.restype = c.c_size_t
    size = K.mvx_warp_perspective_f32_workspace_size(n, dw, dh)
    return torch.empty(max(size, 4) // 4, dtype=torch.float32, device="cuda"), size


def warp_adjoint(grad_out, src, H, want_src=True):
    """(dL/dsrc or None, dL/dH as count x 9 float64) for dL/d(out) = grad_out"""
    n, sh, sw, ss, sb = _geom(src)
    _, dh, dw, gs, gb = _geom(grad_out)
    gsrc = torch.zeros_like(src) if want_src else None
    gH = torch.empty((n, 9), dtype=torch.float64, device=src.device)
    ws, size = _workspace(n, dw, dh)
    _check(kernels().mvx_warp_perspective_f32_adjoint(_stream(), n, dw, dh, _ptr(grad_out), gs, c.c_size_t(gb), sw, sh, _ptr(src),
                                                      ss, c.c_size_t(sb), _ptr(H), _ptr(gsrc), ss, c.c_size_t(sb), _ptr(gH),
                                                      _ptr(ws), c.c_size_t(size)), "mvx_warp_perspective_f32_adjoint")
    return gsrc, gH


class L2Grad:
    """fused warp + L2 loss + dL/dH for one pyramid level of a batch (workspace kept)"""

    def __init__(self, src, tgt):
        self.src, self.tgt = src, tgt
        self.n, self.sh, self.sw, self.ss, self.sb = _geom(src)
        _, self.dh, self.dw, self.ts, self.tb = _geom(tgt)
        self.ws, self.size = _workspace(self.n, self.dw, self.dh)

    def __call__(self, H, weight, loss, gH, accumulate, gsrc=None):
        _check(kernels().mvx_warp_perspective_f32_l2(
            _stream(), self.n, self.dw, self.dh, _ptr(self.tgt), self.ts, c.c_size_t(self.tb), self.sw, self.sh, _ptr(self.src),
            self.ss, c.c_size_t(self.sb), _ptr(H), c.c_float(weight), _ptr(loss), _ptr(gH), _ptr(gsrc), self.ss,
            c.c_size_t(self.sb), int(accumulate), _ptr(self.ws), c.c_size_t(self.size)), "mvx_warp_perspective_f32_l2")


def remap(src, mapxy):
    n, sh, sw, ss, sb = _geom(src)
    _, dh, dw, ms, mb = _geom(mapxy)
    out = torch.empty((n, dh, dw), dtype=torch.float32, device=src.device)
    _check(kernels().mvx_Remap_F32_Bilinear(_stream(), n, dw, dh, _ptr(out), 4 * dw, c.c_size_t(4 * dw * dh), sw, sh, _ptr(src), ss,
                                            c.c_size_t(sb), _ptr(mapxy), ms, c.c_size_t(mb)), "mvx_Remap_F32_Bilinear")
    return out


def remap_adjoint(grad_out, src, mapxy, want_src=True, want_map=True):
    n, sh, sw, ss, sb = _geom(src)
    _, dh, dw, gs, gb = _geom(grad_out)
    _, _, _, ms, mb = _geom(mapxy)
    gsrc = torch.zeros_like(src) if want_src else None
    gmap = torch.empty_like(mapxy) if want_map else None
    _check(kernels().mvx_remap_f32_adjoint(_stream(), n, dw, dh, _ptr(grad_out), gs, c.c_size_t(gb), sw, sh, _ptr(src), ss,
                                           c.c_size_t(sb), _ptr(mapxy), ms, c.c_size_t(mb), _ptr(gsrc), ss, c.c_size_t(sb),
                                           _ptr(gmap), ms, c.c_size_t(mb)), "mvx_remap_f32_adjoint")
    return gsrc, gmap


def pyr_half(x):
    n, sh, sw, ss, sb = _geom(x)
    dh, dw = (sh + 1) // 2, (sw + 1) // 2
    out = torch.empty((n, dh, dw), dtype=torch.float32, device=x.device)
    _check(kernels().mvx_PyramidHalf_F32(_stream(), n, dw, dh, _ptr(out), 4 * dw, c.c_size_t(4 * dw * dh), sw, sh, _ptr(x), ss,
                                         c.c_size_t(sb)), "mvx_PyramidHalf_F32")
    return out


def pyr_half_adjoint(grad_out, src_hw, into=None):
    n, dh, dw, gs, gb = _geom(grad_out)
    sh, sw = src_hw
    gsrc = into if into is not None else torch.zeros((n, sh, sw), dtype=torch.float32, device=grad_out.device)
    _check(kernels().mvx_pyramid_half_f32_adjoint(_stream(), n, sw, sh, _ptr(gsrc), 4 * sw, c.c_size_t(4 * sw * sh), dw, dh,
                                                  _ptr(grad_out), gs, c.c_size_t(gb)), "mvx_pyramid_half_f32_adjoint")
    return gsrc


def pyramid(x, levels=LEVELS):
    out = [x]
    for _ in range(levels - 1):
        out.append(pyr_half(out[-1]))
    return out


# ------------------------------------------------------------------ parameterisation -----------
def norm_matrix(w, h):
    s = w / 2.0
    return np.array([[1 / s, 0, -(w / 2.0) / s], [0, 1 / s, -(h / 2.0) / s], [0, 0, 1]], dtype=np.float64)


def H_from_p(p, w, h):
    """p: (count, 8) float64 tensor -> (count, 3, 3) pixel homographies (destination -> source)"""
    N = torch.as_tensor(norm_matrix(w, h), device=p.device)
    Ni = torch.as_tensor(np.linalg.inv(norm_matrix(w, h)), device=p.device)
    P = torch.zeros((p.shape[0], 9), dtype=torch.float64, device=p.device)
    P[:, :8] = p
    Hh = torch.eye(3, dtype=torch.float64, device=p.device) + P.view(-1, 3, 3)
    return Ni @ Hh @ N


def grad_p_from_grad_H(gH, w, h):
    """dL/dp from dL/dH_pixel (count x 9): dL/dHhat = N^-T G N^T"""
    N = torch.as_tensor(norm_matrix(w, h), device=gH.device)
    Ni = torch.as_tensor(np.linalg.inv(norm_matrix(w, h)), device=gH.device)
    G = Ni.T @ gH.view(-1, 3, 3) @ N.T
    return G.reshape(-1, 9)[:, :8]


def level_H(Hpix, level):
    """S H S^-1 with S = diag(2^-l, 2^-l, 1), as count x 9 float64"""
    s = 0.5 ** level
    S = torch.diag(torch.tensor([s, s, 1.0], dtype=torch.float64, device=Hpix.device))
    Si = torch.diag(torch.tensor([1 / s, 1 / s, 1.0], dtype=torch.float64, device=Hpix.device))
    return (S @ Hpix @ Si).reshape(-1, 9).contiguous()


def level_grad_to_base(gH, level):
    """dL/dH_base = S^T G S^-T for the level's dL/dH_l (count x 9 float64)"""
    s = 0.5 ** level
    sc = torch.tensor([s, s, 1.0], dtype=torch.float64, device=gH.device)
    G = gH.view(-1, 3, 3)
    return (sc[:, None] * G / sc[None, :]).reshape(-1, 9)


def corner_error(H_est, H_true, w, h):
    """mean distance (px) between the images of the 4 image corners under the two homographies"""
    C = torch.tensor([[0, 0, 1], [w - 1, 0, 1], [0, h - 1, 1], [w - 1, h - 1, 1]], dtype=torch.float64, device=H_est.device).T
    a = H_est @ C
    b = H_true @ C
    a = a[:, :2] / a[:, 2:3]
    b = b[:, :2] / b[:, 2:3]
    return torch.linalg.norm(a - b, dim=1).mean(dim=1)


# ------------------------------------------------------------------ workload --------------------
def true_params(count, w, h, seed0=5100):
    """per pair: centre-normalised parameters of rotation U(-10,10) deg, translation U(-40,40) px,
    perspective U(-2e-4,2e-4) per px"""
    s = w / 2.0
    out = np.zeros((count, 8), np.float64)
    for i in range(count):
        u = synth.uniform01(seed0 + i, 6, 5)
        th = math.radians(-10 + 20 * u[0])
        tx, ty = -40 + 80 * u[1], -40 + 80 * u[2]
        px, py = (-2e-4 + 4e-4 * u[3]) * s, (-2e-4 + 4e-4 * u[4]) * s
        out[i] = [math.cos(th) - 1, -math.sin(th), tx / s, math.sin(th), math.cos(th) - 1, ty / s, px, py]
    return out


def gaussian_blur(x, sigma=1.5):
    """data preparation only: separable Gaussian (radius ceil(3 sigma)), edges replicated"""
    r = int(math.ceil(3 * sigma))
    k = torch.exp(-torch.arange(-r, r + 1, dtype=torch.float64) ** 2 / (2 * sigma * sigma))
    k = (k / k.sum()).to(torch.float32).to(x.device)
    y = torch.nn.functional.pad(x[:, None], (r, r, 0, 0), mode="replicate")
    y = torch.nn.functional.conv2d(y, k.view(1, 1, 1, -1))
    y = torch.nn.functional.pad(y, (0, 0, r, r), mode="replicate")
    return torch.nn.functional.conv2d(y, k.view(1, 1, -1, 1))[:, 0].contiguous()


def make_pairs(count, w=1024, h=1024, seed0=5000, device="cuda"):
    """(src, tgt, H_true) of the alignment workload"""
    src = torch.stack([synth.scene_u8_device(w, h, seed0 + i, device).to(torch.float32) / 255.0 for i in range(count)])
    src = gaussian_blur(src, 1.5)
    p = torch.as_tensor(true_params(count, w, h, seed0 + 100), device=device)
    Ht = H_from_p(p, w, h)
    tgt = warp(src, Ht.reshape(-1, 9).contiguous())
    noise = torch.stack([_normal(seed0 + 100 + i, w * h, device).view(h, w) for i in range(count)])
    return src, (tgt + 0.01 * noise).contiguous(), Ht


def _normal(seed, n, device):
    u1 = 1.0 - synth._uniform01_t(seed, 7, n, device)
    u2 = synth._uniform01_t(seed, 8, n, device)
    return (torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2)).to(torch.float32)


class Aligner:
    """Adam on p over the L2 loss summed over the pyramid levels (mean per level).

    ``coarse_init`` picks each pair's starting point from a grid of rotations and translations scored
    on the coarsest level (the loss surface of a 40 px shift has local minima that plain descent from
    the identity can settle in); ``run`` replays one captured HIP graph per Adam step (the fused
    loss-and-gradient launches, the chain rule to dL/dp and the optimiser update), so a step costs its
    kernels rather than their host-side launches."""

    def __init__(self, src, tgt, levels=LEVELS):
        self.n, self.h, self.w = src.shape
        dev = src.device
        self.ps = pyramid(src, levels)
        self.pt = pyramid(tgt, levels)
        self.levels = levels
        self.l2 = [L2Grad(a, b) for a, b in zip(self.ps, self.pt)]
        self.loss = [torch.zeros(self.n, dtype=torch.float64, device=dev) for _ in range(levels)]
        self.gH = [torch.zeros((self.n, 9), dtype=torch.float64, device=dev) for _ in range(levels)]
        # constant matrices on the device once, so a step makes no host-to-device copies
        self.N = torch.as_tensor(norm_matrix(self.w, self.h), device=dev)
        self.Ni = torch.as_tensor(np.linalg.inv(norm_matrix(self.w, self.h)), device=dev)
        self.S = [torch.diag(torch.tensor([0.5 ** l, 0.5 ** l, 1.0], dtype=torch.float64, device=dev)) for l in range(levels)]
        self.Si = [torch.diag(torch.tensor([2.0 ** l, 2.0 ** l, 1.0], dtype=torch.float64, device=dev)) for l in range(levels)]
        self.sc = [torch.tensor([0.5 ** l, 0.5 ** l, 1.0], dtype=torch.float64, device=dev) for l in range(levels)]

    def _H(self, p):
        P = torch.zeros((p.shape[0], 9), dtype=torch.float64, device=p.device)
        P[:, :8] = p
        return self.Ni @ (torch.eye(3, dtype=torch.float64, device=p.device) + P.view(-1, 3, 3)) @ self.N

    def loss_and_grad(self, p, levels=None):
        """(loss per pair, dL/dp) at p over the given pyramid levels (default all) -- one fused
        launch + one reduction per level"""
        Hpix = self._H(p)
        g = torch.zeros((self.n, 9), dtype=torch.float64, device=p.device)
        levels = range(self.levels) if levels is None else levels
        for l in levels:
            lv = self.ps[l]
            Hl = (self.S[l] @ Hpix @ self.Si[l]).reshape(-1, 9).contiguous()
            self.l2[l](Hl, 1.0 / (lv.shape[1] * lv.shape[2]), self.loss[l], self.gH[l], accumulate=False)
            g += (self.sc[l][:, None] * self.gH[l].view(-1, 3, 3) / self.sc[l][None, :]).reshape(-1, 9)
        G = self.Ni.T @ g.view(-1, 3, 3) @ self.N.T
        return sum(self.loss[l] for l in levels), G.reshape(-1, 9)[:, :8]

    def coarse_init(self, rotations_deg=(-8.0, -4.0, 0.0, 4.0, 8.0), shift_px=48.0, step_px=8.0):
        """per pair, the rotation / translation grid point (about the image centre, full-resolution
        pixels) with the smallest loss on the coarsest pyramid level"""
        dev = self.ps[0].device
        top = self.levels - 1
        s = self.w / 2.0
        shifts = torch.arange(-shift_px, shift_px + 0.5 * step_px, step_px, dtype=torch.float64)
        best = torch.full((self.n,), float("inf"), dtype=torch.float64, device=dev)
        p0 = torch.zeros((self.n, 8), dtype=torch.float64, device=dev)
        cand = torch.zeros((self.n, 8), dtype=torch.float64, device=dev)
        for r in rotations_deg:
            th = math.radians(r)
            for tx in shifts.tolist():
                for ty in shifts.tolist():
                    cand[:] = torch.tensor([math.cos(th) - 1, -math.sin(th), tx / s, math.sin(th), math.cos(th) - 1, ty / s, 0.0, 0.0],
                                           dtype=torch.float64)
                    loss, _ = self.loss_and_grad(cand, levels=[top])
                    better = loss < best
                    best = torch.where(better, loss, best)
                    p0 = torch.where(better[:, None], cand, p0)
        return p0

    def run(self, steps=200, lr=4e-3, p0=None, graph=True):
        dev = self.ps[0].device
        p = torch.zeros((self.n, 8), dtype=torch.float64, device=dev) if p0 is None else p0.clone()
        p.requires_grad_(True)
        opt = torch.optim.Adam([p], lr=lr, capturable=True)   # the same update arithmetic eager or replayed
        if not graph:
            for _ in range(steps):
                _, g = self.loss_and_grad(p.detach())
                p.grad = g
                opt.step()
            return p.detach()
        step = self.capture_step(p, opt)
        for _ in range(steps):
            step.replay()
        return p.detach()

    def capture_step(self, p, opt, warmup=3):
        """a CUDA graph of one Adam step on p (the optimiser must be capturable); the warm-up steps
        run eagerly on a side stream first, as graph capture requires, and are undone"""
        p0 = p.detach().clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                _, g = self.loss_and_grad(p.detach())
                p.grad = g
                opt.step()
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            _, g = self.loss_and_grad(p.detach())
            p.grad = g
            opt.step()
        # back to the starting point and a fresh optimiser state
        with torch.no_grad():
            p.copy_(p0)
            for st in opt.state.values():
                for k, v in st.items():
                    if torch.is_tensor(v):
                        v.zero_()
        return graph
