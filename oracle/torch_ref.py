"""Dense FP64 reference of the forward frame, differentiable with torch autograd.

TEST INFRASTRUCTURE ONLY (the gradient checker).  It restates project.slang:33-90 (with
splat/common.slang, splat/volume.slang) and blend.slang:54-101 in float64 torch operations on the
training parameterisation (means, log-scales, raw quaternions, opacity logits, SH), evaluated on
a FIXED discrete structure taken from the FP32 oracle: for every pixel the ordered list of splats
the FP32 blend actually used (oracle.blend_decisions).  Culling, tile rectangles, the 1/255 skip
and early termination are piecewise constant in the parameters, so with the structure fixed the
function is smooth and both autograd and central finite differences apply.  The 0.99 alpha clamp
and max(colour, 0) are kept as clamps (zero gradient where active), as in the forward.
"""
from __future__ import annotations

import math

import numpy as np
import torch

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)
C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435)


def _sh_color(sh, d, deg):
    """splat/common.slang:35-80 in float64, band-3 '+ feature 12' included.  sh: (N, 48), d: (N, 3)."""
    x, y, z = d[:, 0:1], d[:, 1:2], d[:, 2:3]

    def f(idx):  # (N, 3): coefficient idx of channels r, g, b
        return torch.stack([sh[:, 3 + c * 15 + idx - 1] for c in range(3)], 1)

    res = C0 * sh[:, 0:3]
    if deg > 0:
        res = res - C1 * y * f(1) + C1 * z * f(2) - C1 * x * f(3)
        if deg > 1:
            xx, yy, zz, xy, yz, zx = x * x, y * y, z * z, x * y, y * z, z * x
            res = (res + C2[0] * xy * f(4) + C2[1] * yz * f(5) + C2[2] * (2.0 * zz - xx - yy) * f(6)
                   + C2[3] * zx * f(7) + C2[4] * (xx - yy) * f(8))
            if deg > 2:
                res = (res + C3[0] * y * (3.0 * xx - yy) * f(9) + C3[1] * xy * z * f(10)
                       + C3[2] * y * (4.0 * zz - xx - yy) * f(11)
                       + C3[3] * z * (2.0 * zz - 3.0 * xx - 3.0 * yy) + f(12)
                       + C3[4] * x * (4.0 * zz - xx - yy) * f(13) + C3[5] * z * (xx - yy) * f(14)
                       + C3[6] * x * (xx - 3.0 * yy) * f(15))
    return torch.clamp(res + 0.5, min=0.0)


def splats(means, log_scales, quats, opacity_logits, sh, view, proj_view, focal_ndc, width, height, deg,
           modifier=1.0):
    """Per-Gaussian screen quantities: (px, py, conic a, b, c, colour (N,3), opacity)."""
    V = torch.as_tensor(np.asarray(view, np.float64).reshape(4, 4), dtype=torch.float64)
    PV = torch.as_tensor(np.asarray(proj_view, np.float64).reshape(4, 4), dtype=torch.float64)
    ones = torch.ones_like(means[:, :1])
    p4 = torch.cat([means, ones], 1)
    v = p4 @ V.T
    clip = p4 @ PV.T
    vx, vy, vz = v[:, 0], v[:, 1], v[:, 2]
    ndc_x, ndc_y = clip[:, 0] / clip[:, 3], clip[:, 1] / clip[:, 3]
    px = ((ndc_x + 1.0) * width - 1.0) * 0.5
    py = ((ndc_y + 1.0) * height - 1.0) * 0.5
    s = torch.exp(log_scales) * modifier
    qx, qy, qz, qw = quats[:, 0], quats[:, 1], quats[:, 2], quats[:, 3]
    R = torch.stack([
        torch.stack([1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)], -1),
        torch.stack([2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)], -1),
        torch.stack([2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)], -1)], -2)
    M = R * s[:, None, :]
    S3 = M @ M.transpose(1, 2)
    fx = 0.5 * width * float(focal_ndc[0])
    fy = 0.5 * height * float(focal_ndc[1])
    zero = torch.zeros_like(vz)
    J = torch.stack([torch.stack([fx / vz, zero, -fx * vx / (vz * vz)], -1),
                     torch.stack([zero, fy / vz, -fy * vy / (vz * vz)], -1)], -2)
    T = J @ V[:3, :3]
    S2 = T @ S3 @ T.transpose(1, 2)
    ca, cb, cc = S2[:, 0, 0] + 0.3, S2[:, 1, 0], S2[:, 1, 1] + 0.3
    det = ca * cc - cb * cb
    cam = -(V[:3, :3].T @ V[:3, 3])
    d = means - cam
    d = d / torch.sqrt((d * d).sum(1, keepdim=True))
    color = _sh_color(sh, d, deg)
    return px, py, cc / det, -cb / det, ca / det, color, torch.sigmoid(opacity_logits)


def render(params, cam, deg, vals, counts, positions):
    """FP64 image (H, W, 3) blending, for every pixel, exactly the splats listed for it."""
    means, log_scales, quats, logits, sh = params
    W, H = cam.width, cam.height
    px, py, ca, cb, cc, color, op = splats(means, log_scales, quats, logits, sh, cam.view, cam.proj_view,
                                            cam.focal_ndc, W, H, deg)
    counts = torch.as_tensor(counts.astype(np.int64))
    pix = torch.repeat_interleave(torch.arange(W * H), counts)
    g = torch.as_tensor(np.asarray(vals, np.int64)[np.asarray(positions, np.int64)])
    fxp = (pix % W).to(torch.float64)
    fyp = (pix // W).to(torch.float64)
    dx = px[g] - fxp
    dy = py[g] - fyp
    power = -0.5 * (ca[g] * dx * dx + cc[g] * dy * dy) - cb[g] * dx * dy
    alpha = torch.clamp(op[g] * torch.exp(power), max=0.99)
    l1m = torch.log1p(-alpha)
    cs = torch.cumsum(l1m, 0)
    starts = torch.cumsum(counts, 0) - counts
    before = torch.where(starts > 0, cs[(starts - 1).clamp(min=0)], torch.zeros_like(starts, dtype=torch.float64))
    T = torch.exp(cs - l1m - before[pix])
    img = torch.zeros(W * H, 3, dtype=torch.float64).index_add(0, pix, color[g] * (alpha * T)[:, None])
    return img.reshape(H, W, 3)


def make_params(records, dtype=torch.float64):
    """Training parameters whose forward reproduces the records: log-scales = log(scale), opacity
    logits = logit(opacity); means, quaternions, SH as stored.  (scale modifier must be 1.)"""
    r = np.asarray(records, np.float64)
    op = np.clip(r[:, 3], 1e-12, 1 - 1e-12)
    return [torch.tensor(r[:, 0:3], dtype=dtype), torch.tensor(np.log(r[:, 8:11]), dtype=dtype),
            torch.tensor(r[:, 4:8], dtype=dtype), torch.tensor(np.log(op) - np.log1p(-op), dtype=dtype),
            torch.tensor(r[:, 12:60], dtype=dtype)]


def records_from_params(params) -> np.ndarray:
    """FP32 records the forward kernel consumes for the given parameters."""
    means, ls, q, lg, sh = [p.detach().to(torch.float32) for p in params]
    n = means.shape[0]
    rec = torch.cat([means, torch.sigmoid(lg)[:, None], q, torch.exp(ls), torch.ones(n, 1), sh], 1)
    return rec.numpy().astype(np.float32)


def grads(params, cam, deg, vals, counts, positions, upstream):
    ps = [p.detach().clone().requires_grad_(True) for p in params]
    img = render(ps, cam, deg, vals, counts, positions)
    loss = (img * torch.as_tensor(upstream, dtype=torch.float64)).sum()
    loss.backward()
    return [p.grad.detach().numpy() for p in ps], float(loss.detach())


def finite_difference(params, cam, deg, vals, counts, positions, upstream, which, idx, h=1e-6):
    """Central difference of sum(image * upstream) with respect to params[which][idx]."""
    up = torch.as_tensor(upstream, dtype=torch.float64)
    vals_ = []
    for sgn in (1.0, -1.0):
        ps = [p.detach().clone() for p in params]
        flat = ps[which].reshape(-1)
        step = h * max(1.0, abs(float(flat[idx])))
        flat[idx] += sgn * step
        vals_.append((float((render(ps, cam, deg, vals, counts, positions) * up).sum()), step))
    return (vals_[0][0] - vals_[1][0]) / (2.0 * vals_[0][1])


def relative_error(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    denom = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / denom) if denom > 0 else float(np.linalg.norm(a))


__all__ = ["splats", "render", "make_params", "records_from_params", "grads", "finite_difference",
           "relative_error", "math"]
