"""The FP64 dense gradient reference (oracle/torch_ref.py) pinned on the CPU.

  - its forward, evaluated on the FP32 frame's own blend decisions, reproduces the FP32 oracle
    image to FP32 rounding;
  - its autograd gradients agree with central finite differences in FP64 for every parameter
    group (means, log-scales, quaternions, opacity logits, SH), including Gaussians whose alpha
    hits the 0.99 clamp and SH degree 0..3;
  - a textbook case: one isotropic Gaussian, d(image)/d(opacity logit) and d/d(DC SH) in closed form.
"""
import numpy as np
import pytest
import torch

from conftest import small_scene



@pytest.fixture(autouse=True)
def float64_default():
    """FP64 default for tensors made inside these tests only (restored afterwards)."""
    prev = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(prev)

GROUPS = ("means", "log_scales", "quats", "opacity_logits", "sh")


def setup_frame(orc, n=2000, w=64, h=64, seed=3, deg=3, log_scale_mean=-3.0, eye=(0, 0, -3), up=(0, -1, 0)):
    from oracle import torch_ref
    g = small_scene(n, seed=seed, log_scale_mean=log_scale_mean)
    params = torch_ref.make_params(g)
    rec = torch_ref.records_from_params(params)
    cam = orc.make_camera(list(eye), [0, 0, 0], list(up), w, h)
    r = orc.render(rec, cam, deg)
    counts, pos = orc.blend_decisions(r["splats"], r["vals"], r["ranges"], w, h)
    return params, rec, cam, r, counts, pos


@pytest.mark.parametrize("eye,up", [((0, 0, -3), (0, -1, 0)), ((2.1, -1.4, -1.6), (0.3, 0.2, 1.0))])
def test_dense_forward_matches_fp32_oracle(orc, eye, up):
    """Also from an oblique camera (rotated view, off-axis means), where the oracle's projection is
    pinned independently by tests/test_oracle_geometry.py."""
    from oracle import torch_ref
    params, rec, cam, r, counts, pos = setup_frame(orc, eye=eye, up=up, w=80, h=56)
    img = torch_ref.render(params, cam, 3, r["vals"], counts, pos).detach().numpy()
    assert counts.sum() > 5000
    err = np.abs(img - r["image"]).max()
    assert err < 2e-5, err


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_autograd_matches_finite_differences(orc, deg):
    from oracle import torch_ref
    eye, up = ((0, 0, -3), (0, -1, 0)) if deg % 2 == 0 else ((1.8, 1.2, -2.0), (-0.2, 1.0, 0.4))
    params, rec, cam, r, counts, pos = setup_frame(orc, n=600, w=48, h=48, seed=10 + deg, deg=deg, eye=eye, up=up)
    rng = np.random.default_rng(deg)
    up = rng.normal(size=(48, 48, 3))
    gr, _ = torch_ref.grads(params, cam, deg, r["vals"], counts, pos, up)
    used = np.unique(np.asarray(r["vals"])[pos])
    for which, name in enumerate(GROUPS):
        width = params[which].reshape(params[which].shape[0], -1).shape[1]
        picks = rng.choice(used, size=6, replace=False)
        for gi in picks:
            for comp in rng.choice(width, size=min(width, 3), replace=False):
                idx = int(gi) * width + int(comp)
                fd = torch_ref.finite_difference(params, cam, deg, r["vals"], counts, pos, up, which, idx)
                an = gr[which].reshape(-1)[idx]
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-7, (name, gi, comp, fd, an)


def test_clamped_alpha_has_no_opacity_gradient(orc):
    from oracle import torch_ref
    from torpedo_amd.engine import gaussian_point
    g = gaussian_point([0, 0, 0], 0.999, (0, 0, 0, 1), (1.0, 1.0, 1.0), np.full(48, 0.1, np.float32))[None]
    params = torch_ref.make_params(g)
    rec = torch_ref.records_from_params(params)
    cam = orc.make_camera([0, 0, -3], [0, 0, 0], [0, -1, 0], 32, 32)
    r = orc.render(rec, cam, 3)
    counts, pos = orc.blend_decisions(r["splats"], r["vals"], r["ranges"], 32, 32)
    up = np.zeros((32, 32, 3))
    up[15, 15] = 1.0  # centre pixel: o * exp(power) > 0.99 -> clamped
    gr, _ = torch_ref.grads(params, cam, 3, r["vals"], counts, pos, up)
    assert gr[3][0] == 0.0
    assert np.abs(gr[4]).sum() > 0  # colour still receives gradient


def test_single_gaussian_closed_form(orc):
    """C = c * alpha at every pixel: dC/dlogit = c * alpha * (1 - o), dC/d(DC) = C0 * alpha."""
    from oracle import torch_ref
    from torpedo_amd.engine import gaussian_point
    g = gaussian_point([0.01, -0.02, 0.0], 0.6, (0, 0, 0, 1), (0.05, 0.05, 0.05), np.zeros(48, np.float32))[None]
    params = torch_ref.make_params(g)
    rec = torch_ref.records_from_params(params)
    cam = orc.make_camera([0, 0, -3], [0, 0, 0], [0, -1, 0], 32, 32)
    r = orc.render(rec, cam, 0)
    counts, pos = orc.blend_decisions(r["splats"], r["vals"], r["ranges"], 32, 32)
    px, py, ca, cb, cc, color, op = [t.detach().numpy() for t in torch_ref.splats(*params, cam.view, cam.proj_view,
                                                                                  cam.focal_ndc, 32, 32, 0)]
    up = np.zeros((32, 32, 3))
    up[..., 0] = 1.0
    gr, _ = torch_ref.grads(params, cam, 0, r["vals"], counts, pos, up)
    expect_logit = 0.0
    expect_dc = 0.0
    for p in range(32 * 32):
        if counts[p] == 0:
            continue
        x, y = p % 32, p // 32
        dx, dy = px[0] - x, py[0] - y
        a = op[0] * np.exp(-0.5 * (ca[0] * dx * dx + cc[0] * dy * dy) - cb[0] * dx * dy)
        expect_logit += color[0, 0] * a * (1 - op[0])
        expect_dc += 0.28209479177387814 * a
    assert abs(gr[3][0] - expect_logit) < 1e-10 * max(1, abs(expect_logit))
    assert abs(gr[4][0, 0] - expect_dc) < 1e-10 * max(1, abs(expect_dc))
