"""Seeded synthetic inputs for the vision-graph workloads (the reference ships no image data).

Random numbers come from a counter-based generator: value i of stream (seed, stream) is the upper
32 bits of splitmix64's finaliser applied to i plus a per-stream key, so any slice can be produced
independently and the sequence never depends on a library's RNG implementation.  Tests commit
SHA-256 hashes of generated inputs, which pins the generator.

Workloads
* ``uniform_u8(w, h, seed)``: i.i.d. uniform U8 (the 1920x1080 Gaussian->Sobel->Magnitude/Phase graph).
* ``scene_rgb(w, h, seed)``: per channel, 128 + the sum of 8 random 2-D sinusoids (amplitude
  U(10,40), period log-uniform in [16,512] px, orientation U(0,pi), phase U(0,2pi)), then 40
  axis-aligned rectangles of constant random intensity (corner uniform over the image, size
  U(16,512), painted in order), then N(0, 8^2) noise by Box-Muller, rounded half-to-even and
  clipped to [0,255] (the 3840x2160 vision graph input).
* ``scene_u8(w, h, seed)``: the same process for one channel.
* ``warp_matrices(w, h, seed)``: a near-identity affine (2x3) and perspective (3x3) VX matrix about
  the image centre -- rotation U(-5,5) degrees, scale U(0.9,1.1), translation U(-50,50) px,
  perspective terms U(-1e-4,1e-4) -- as the float32 arrays the VX matrix objects store.
* ``lk_pair(w, h, seed)``: the tracking workload -- a scene and a copy translated by a seeded
  (tx, ty) with |t| in [2, 8] px (bilinear, edges clamped), and the translation.
* ``lk_keypoints(w, h, n, seed, margin=16)``: n vx_keypoint_t records at seeded integer positions
  inside the margin (tracking_status 1, other fields 0).
* ``barrel_map(w, h, k1)``: the float remap table of a barrel distortion r' = r(1 + k1 r^2) about
  the centre (r normalised by the half-diagonal): (h, w, 2) float32 source (x, y) per output pixel.
"""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(x):
    x = x ^ (x >> np.uint64(30))
    x = x * _M1
    x = x ^ (x >> np.uint64(27))
    x = x * _M2
    return x ^ (x >> np.uint64(31))


def u32(seed, stream, n, offset=0):
    """n uint32 values of stream (seed, stream), starting at counter `offset`."""
    with np.errstate(over="ignore"):
        key = _mix64(np.array([seed], dtype=np.uint64) * _GOLD + np.uint64(stream + 1))[0]
        idx = np.arange(offset, offset + n, dtype=np.uint64)
        return (_mix64(idx * _GOLD + key) >> np.uint64(32)).astype(np.uint32)


def uniform01(seed, stream, n, offset=0):
    """n doubles in [0, 1) with 32-bit resolution."""
    return u32(seed, stream, n, offset).astype(np.float64) * (1.0 / 4294967296.0)


def uniform_u8(w, h, seed=0):
    return (u32(seed, 1, w * h) >> np.uint32(24)).astype(np.uint8).reshape(h, w)


def _scene_channel(w, h, seed, ch):
    p = uniform01(seed, 100 + ch, 8 * 4)
    xs = np.arange(w, dtype=np.float64)
    ys = np.arange(h, dtype=np.float64)
    img = np.full((h, w), 128.0)
    for k in range(8):
        amp = 10.0 + 30.0 * p[4 * k]
        period = 16.0 * np.exp(np.log(512.0 / 16.0) * p[4 * k + 1])
        theta = np.pi * p[4 * k + 2]
        phase = 2.0 * np.pi * p[4 * k + 3]
        a = 2.0 * np.pi * np.cos(theta) / period
        b = 2.0 * np.pi * np.sin(theta) / period
        # sin(a x + b y + phase) = sin(a x) cos(b y + phase) + cos(a x) sin(b y + phase)
        sx, cx = np.sin(a * xs), np.cos(a * xs)
        sy, cy = np.sin(b * ys + phase), np.cos(b * ys + phase)
        img += amp * (np.outer(cy, sx) + np.outer(sy, cx))
    return img


def _rects(w, h, seed, nch):
    r = uniform01(seed, 200, 40 * (4 + nch))
    out = []
    for k in range(40):
        q = r[k * (4 + nch):(k + 1) * (4 + nch)]
        x0 = int(q[0] * w)
        y0 = int(q[1] * h)
        rw = 16 + int(q[2] * (512 - 16 + 1))
        rh = 16 + int(q[3] * (512 - 16 + 1))
        vals = [int(v * 256) for v in q[4:]]
        out.append((x0, y0, rw, rh, vals))
    return out


def _noise(w, h, seed, ch, sigma=8.0):
    n = w * h
    u1 = uniform01(seed, 300 + ch, n)
    u2 = uniform01(seed, 400 + ch, n)
    u1 = 1.0 - u1   # (0, 1]
    return (sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).reshape(h, w)


def _scene(w, h, seed, nch):
    rects = _rects(w, h, seed, nch)
    planes = []
    for ch in range(nch):
        img = _scene_channel(w, h, seed, ch)
        for (x0, y0, rw, rh, vals) in rects:
            img[y0:y0 + rh, x0:x0 + rw] = vals[ch]
        img += _noise(w, h, seed, ch)
        planes.append(np.clip(np.rint(img), 0, 255).astype(np.uint8))
    return planes


def scene_rgb(w, h, seed=0):
    """(h, w, 3) uint8 RGB scene."""
    return np.ascontiguousarray(np.stack(_scene(w, h, seed, 3), axis=-1))


def scene_u8(w, h, seed=0):
    """(h, w) uint8 single-channel scene."""
    return _scene(w, h, seed, 1)[0]


def warp_matrices(w, h, seed=0):
    """(affine 6 floats, perspective 9 floats) in the VX storage order: x_src = m0*x + m2*y + m4,
    y_src = m1*x + m3*y + m5; perspective (a d g / b e h / c f i) with z = g*x + h*y + i."""
    p = uniform01(seed, 500, 8)
    ang = np.deg2rad(-5.0 + 10.0 * p[0])
    sc = 0.9 + 0.2 * p[1]
    tx, ty = -50.0 + 100.0 * p[2], -50.0 + 100.0 * p[3]
    g, hh = -1e-4 + 2e-4 * p[4], -1e-4 + 2e-4 * p[5]
    c, s = np.cos(ang) * sc, np.sin(ang) * sc
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    # rotate/scale about the centre, then translate
    a00, a01, a10, a11 = c, -s, s, c
    bx = cx - (a00 * cx + a01 * cy) + tx
    by = cy - (a10 * cx + a11 * cy) + ty
    affine = np.array([a00, a10, a01, a11, bx, by], np.float32)
    persp = np.array([a00, a10, g, a01, a11, hh, bx, by, 1.0 - g * cx - hh * cy], np.float32)
    return affine, persp


def barrel_map(w, h, k1):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    rn = np.hypot(cx, cy)
    dx, dy = (xx - cx) / rn, (yy - cy) / rn
    f = 1.0 + k1 * (dx * dx + dy * dy)
    return np.ascontiguousarray(np.stack([cx + dx * f * rn, cy + dy * f * rn], -1).astype(np.float32))


KEYPOINT = np.dtype([("x", np.int32), ("y", np.int32), ("strength", np.float32), ("scale", np.float32),
                     ("orientation", np.float32), ("tracking_status", np.int32), ("error", np.float32)])


def shift_image(img, tx, ty):
    """out(x, y) = img(x - tx, y - ty), bilinear, coordinates clamped to the image, rounded"""
    h, w = img.shape
    xs = np.clip(np.arange(w, dtype=np.float64) - tx, 0, w - 1)
    ys = np.clip(np.arange(h, dtype=np.float64) - ty, 0, h - 1)
    x0 = np.minimum(np.floor(xs).astype(np.int64), w - 2)
    y0 = np.minimum(np.floor(ys).astype(np.int64), h - 2)
    fx, fy = xs - x0, ys - y0
    f = img.astype(np.float64)
    top = f[y0][:, x0] * (1 - fx) + f[y0][:, x0 + 1] * fx
    bot = f[y0 + 1][:, x0] * (1 - fx) + f[y0 + 1][:, x0 + 1] * fx
    out = top * (1 - fy)[:, None] + bot * fy[:, None]
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def lk_pair(w, h, seed=1):
    p = uniform01(seed, 600, 2)
    mag = 2.0 + 6.0 * p[0]
    ang = 2.0 * np.pi * p[1]
    tx, ty = mag * np.cos(ang), mag * np.sin(ang)
    old = scene_u8(w, h, seed)
    return old, shift_image(old, tx, ty), (tx, ty)


def lk_keypoints(w, h, n, seed=1, margin=16):
    r = u32(seed, 700, 2 * n)
    kp = np.zeros(n, KEYPOINT)
    kp["x"] = margin + (r[0::2] % np.uint32(w - 2 * margin)).astype(np.int32)
    kp["y"] = margin + (r[1::2] % np.uint32(h - 2 * margin)).astype(np.int32)
    kp["tracking_status"] = 1
    return kp


def batch_sizes(n=256, seed=0):
    """The uneven-size batch: n (width, height) pairs, width uniform over {320, 336, ..., 4096} and
    height uniform over {240, 256, ..., 2160} (stream 4 of `seed`)."""
    u = u32(seed, 4, 2 * n).astype(np.uint64)
    ws = 320 + 16 * ((u[0::2] * np.uint64(237)) >> np.uint64(32)).astype(np.int64)
    hs = 240 + 16 * ((u[1::2] * np.uint64(121)) >> np.uint64(32)).astype(np.int64)
    return [(int(a), int(b)) for a, b in zip(ws, hs)]


def _signed64(v):
    v = int(v) & 0xFFFFFFFFFFFFFFFF
    return v - (1 << 64) if v >= (1 << 63) else v


def _mix64_t(x):
    import torch
    x = x ^ ((x >> 30) & ((1 << 34) - 1))
    x = x * _signed64(_M1)
    x = x ^ ((x >> 27) & ((1 << 37) - 1))
    x = x * _signed64(_M2)
    return x ^ ((x >> 31) & ((1 << 33) - 1))


def _uniform01_t(seed, stream, n, device):
    """uniform01 computed with torch on `device` (int64 arithmetic wraps like uint64)"""
    import torch
    with np.errstate(over="ignore"):
        key = _mix64(np.array([seed], dtype=np.uint64) * _GOLD + np.uint64(stream + 1))[0]
    idx = torch.arange(n, dtype=torch.int64, device=device)
    v = (_mix64_t(idx * _signed64(_GOLD) + _signed64(key)) >> 32) & 0xFFFFFFFF
    return v.to(torch.float64) * (1.0 / 4294967296.0)


def scene_u8_device(w, h, seed=0, device="cuda"):
    """``scene_u8`` computed with torch on `device` (float64), for large batches: the same random
    streams and formulas; transcendental functions may differ from numpy's in the last bit, which
    moves a pixel by one level only where the noisy value sits on a rounding boundary."""
    import torch
    p = uniform01(seed, 100, 8 * 4)
    xs = torch.arange(w, dtype=torch.float64, device=device)
    ys = torch.arange(h, dtype=torch.float64, device=device)
    img = torch.full((h, w), 128.0, dtype=torch.float64, device=device)
    for k in range(8):
        amp = 10.0 + 30.0 * p[4 * k]
        period = 16.0 * np.exp(np.log(512.0 / 16.0) * p[4 * k + 1])
        theta = np.pi * p[4 * k + 2]
        phase = 2.0 * np.pi * p[4 * k + 3]
        a = 2.0 * np.pi * np.cos(theta) / period
        b = 2.0 * np.pi * np.sin(theta) / period
        sx, cx = torch.sin(a * xs), torch.cos(a * xs)
        sy, cy = torch.sin(b * ys + phase), torch.cos(b * ys + phase)
        img += amp * (torch.outer(cy, sx) + torch.outer(sy, cx))
    for (x0, y0, rw, rh, vals) in _rects(w, h, seed, 1):
        img[y0:y0 + rh, x0:x0 + rw] = vals[0]
    u1 = 1.0 - _uniform01_t(seed, 300, w * h, device)
    u2 = _uniform01_t(seed, 400, w * h, device)
    img += (8.0 * torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * np.pi * u2)).reshape(h, w)
    return torch.clamp(torch.round(img), 0, 255).to(torch.uint8)


def stitch_rig(W=7680, H=3840, sw=1920, sh=1080, cams=6, overlap=64, device="cuda"):
    """The stitch variant of the uneven-batch configuration: `cams` pinhole cameras of sw x sh around
    the vertical axis (camera c at yaw 360/cams * c degrees plus a seeded +-2 degree jitter, 90 degree
    horizontal field of view, radial lens term k1 in +-0.1; stream 4100 + c of seed 0), each with an F32
    map of the W x H equirectangular frame into its image (entries outside the image are -1), the U8
    blend weights of a linear ramp over `overlap` pixels around the mid-angle seam between neighbours
    (255 inside a camera's own sector, 0 where it does not see), and Loom's camera-selection maps for
    groups of 8 pixels: camId0 (U8: a single camera id, 31 for none, or a blend code 128 + count - 2),
    camId1 / camId2 (U16: three 5-bit ids each, 31 where unused).  Maps, weights and selection are torch
    tensors on `device`; the camera images (scene generator seeds 4000 + c) are numpy arrays."""
    import torch
    dev = torch.device(device)
    f = (sw / 2.0) / np.tan(np.deg2rad(45.0))
    cx, cy = (sw - 1) / 2.0, (sh - 1) / 2.0
    lon = ((torch.arange(W, dtype=torch.float64, device=dev) + 0.5) / W * 2.0 * np.pi - np.pi)
    lat = (np.pi / 2.0 - (torch.arange(H, dtype=torch.float64, device=dev) + 0.5) / H * np.pi)
    px_per_deg = W / 360.0
    maps, weights, yaws = [], [], []
    for c in range(cams):
        p = uniform01(0, 4100 + c, 2)
        yaw = 360.0 / cams * c + (-2.0 + 4.0 * p[0])
        k1 = -0.1 + 0.2 * p[1]
        yaws.append(yaw)
        d = torch.remainder(lon[None, :] - np.deg2rad(yaw) + np.pi, 2.0 * np.pi) - np.pi   # yaw offset, (1, W)
        cl = torch.cos(lat)[:, None]
        xcam, ycam, zcam = cl * torch.sin(d), torch.sin(lat)[:, None].expand(H, 1), cl * torch.cos(d)
        ok = zcam > 1e-6
        zs = torch.where(ok, zcam, torch.ones_like(zcam))
        u, v = f * xcam / zs, -f * ycam / zs
        r2 = (u * u + v * v) / (f * f)
        s = 1.0 + k1 * r2
        mx, my = cx + u * s, cy + v * s
        inside = ok & (mx >= 0) & (my >= 0) & (mx <= sw - 1) & (my <= sh - 1)
        m = torch.stack([torch.where(inside, mx, torch.full_like(mx, -1.0)), torch.where(inside, my, torch.full_like(my, -1.0))], -1)
        maps.append(m.to(torch.float32).contiguous())
        # ramp: 1 up to (seam - overlap/2), 0 beyond (seam + overlap/2), seam at half the camera spacing
        half = 180.0 / cams
        dist_px = torch.abs(torch.rad2deg(d)) * px_per_deg                      # (1, W)
        ramp = torch.clamp(0.5 + (half * px_per_deg - dist_px) / overlap, 0.0, 1.0)
        wt = torch.round(ramp * 255.0).expand(H, W) * inside.to(torch.float64)
        weights.append(wt.to(torch.uint8))
    wts = torch.stack(weights)                                                   # (cams, H, W)
    groups = W // 8
    used = (wts.reshape(cams, H, groups, 8) > 0).any(-1)                           # (cams, H, groups)
    full = (wts.reshape(cams, H, groups, 8) == 255).all(-1)
    n = used.sum(0)
    ids = torch.full((6, H, groups), 31, dtype=torch.int32, device=dev)
    order = torch.cumsum(used.to(torch.int32), 0) - 1                            # rank of each used camera
    for c in range(cams):
        for k in range(6):
            ids[k] = torch.where(used[c] & (order[c] == k), torch.full_like(ids[k], c), ids[k])
    single = (n == 1)
    single_full = torch.zeros_like(single)
    single_id = torch.full((H, groups), 31, dtype=torch.int32, device=dev)
    for c in range(cams):
        hit = single & used[c]
        single_full |= hit & full[c]
        single_id = torch.where(hit, torch.full_like(single_id, c), single_id)
    code = torch.where(n <= 2, torch.full_like(n, 128), 126 + n).to(torch.int32)
    cam0 = torch.where(n == 0, torch.full_like(code, 31), torch.where(single_full, single_id, code))
    cam1 = (ids[0] | (ids[1] << 5) | (ids[2] << 10)).to(torch.int32)
    cam2 = (ids[3] | (ids[4] << 5) | (ids[5] << 10)).to(torch.int32)
    images = [scene_rgb(sw, sh, 4000 + c) for c in range(cams)]
    return dict(images=images, maps=maps, weights=wts.reshape(cams * H, W).contiguous(), camId0=cam0.to(torch.uint8).contiguous(),
                camId1=cam1.to(torch.int16).contiguous(), camId2=cam2.to(torch.int16).contiguous(), yaws=yaws, W=W, H=H, sw=sw, sh=sh)
