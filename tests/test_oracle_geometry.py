"""The oracle's preprocess pinned in general geometry, against mathematics it does not share.

Every other projection pin looks down the optical axis of a camera whose view is [I | t]; there a
rotation of the covariance, the off-axis terms of the Jacobian and the view rotation W all drop
out.  These tests take rotated anisotropic Gaussians seen from random lookAt cameras (non-square
images, so fx != fy; off-axis means; non-identity W; model matrices) and check the oracle
(oracle/oracle_core.inc, which restates project.slang:33-90, splat/volume.slang:3-63 and
splat/common.slang:35-119) against:

  R(q)          scipy.spatial.transform.Rotation.from_quat (scalar-last x, y, z, w), so the 3D
                covariance is R diag(s)^2 R^T built without the oracle's quaternion formula;
  J W           central finite differences, in float64, of the actual pixel mapping
                p -> ndc2pix((PV p).xy / (PV p).w) at the Gaussian's (model-transformed) mean: the
                EWA affine approximation is exactly that derivative, so the expected 2D
                covariance (J W) S3 (J W)^T + 0.3 I owes nothing to computeJacobian's formula;
  SH basis      scipy.special.sph_harm_y (complex, Condon-Shortley phase) turned into the real
                basis; 3DGS's basis is that real basis times (-1)^m, coefficient k = l^2 + l + m
                (common.slang:35-80), band 3's m = 0 term added unscaled with feature 12 added
                raw (the typo at common.slang:69);
  camera pos    numpy's inverse of the view matrix (getCameraWorldPosition, common.slang:88-92).

Each test also shows that it has power: the same expectation built with a plausible mistake
(W transposed, the quaternion conjugated, x and y swapped, the model matrix put into W, an SH
term's sign flipped) lands far outside the tolerance.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation
from scipy.special import sph_harm_y

C0 = 0.28209479177387814
C1 = 0.4886025119029199


def record(pos, quat, scale, opacity=0.7, sh=None):
    from torpedo_amd.engine import gaussian_point
    return gaussian_point(pos, opacity, quat, list(scale) + [1.0], sh)


def random_camera(orc, rng, w, h, dist=(3.0, 8.0), center_spread=0.3):
    while True:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        up = rng.normal(size=3)
        up /= np.linalg.norm(up)
        if abs(np.dot(d, up)) < 0.9:
            break
    center = rng.uniform(-center_spread, center_spread, 3)
    eye = center - d * rng.uniform(*dist)
    return orc.make_camera(eye.astype(np.float32), center.astype(np.float32), up.astype(np.float32), w, h), center


def pix_map(cam, q):
    """Pixel position of world point q (float64), through the camera's own projView."""
    PV = cam.proj_view.astype(np.float64)
    c = PV @ np.array([q[0], q[1], q[2], 1.0])
    ndc = c[:2] / c[3]
    return ((ndc + 1.0) * np.array([cam.width, cam.height]) - 1.0) * 0.5


def fd_jacobian(cam, q, h=1e-4):
    """d pixel / d world position at q by central differences: the 2x3 matrix J W."""
    J = np.zeros((2, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        J[:, k] = (pix_map(cam, q + e) - pix_map(cam, q - e)) / (2 * h)
    return J


def expected_cov2d(JW, R, s):
    S3 = R @ np.diag(np.asarray(s, np.float64) ** 2) @ R.T
    return JW @ S3 @ JW.T + 0.3 * np.eye(2)


def conic_residual(conic, cov):
    """max |conic * cov - I| over the 2x2 entries: 0 when the conic is exactly cov's inverse."""
    a, b, c = conic
    M = np.array([[a, b], [b, c]], np.float64)
    return float(np.abs(M @ cov - np.eye(2)).max())


def make_scene(rng, n, center, spread=0.6, log_scale=(-3.2, 0.5)):
    quats = rng.normal(size=(n, 4))
    quats /= np.linalg.norm(quats, axis=1, keepdims=True)
    scales = np.exp(rng.normal(log_scale[0], log_scale[1], size=(n, 3)))
    means = center + rng.uniform(-spread, spread, size=(n, 3))
    recs = np.stack([record(means[i], quats[i].astype(np.float32), scales[i].astype(np.float32))
                     for i in range(n)])
    return recs


@pytest.mark.parametrize("case", range(6))
def test_projection_against_scipy_rotation_and_fd_jacobian(orc, case):
    rng = np.random.default_rng(100 + case)
    w, h = [(640, 360), (333, 517), (1920, 1080), (256, 256), (800, 600), (1600, 900)][case]
    cam, center = random_camera(orc, rng, w, h)
    recs = make_scene(rng, 400, center)
    sp = orc.preprocess(recs, cam, 0)
    vis = np.nonzero(sp["tiles"] > 0)[0]
    assert vis.size > 100, vis.size
    V = cam.view.astype(np.float64)
    W3 = V[:3, :3]
    worst = {"ok": 0.0, "W transposed": np.inf, "quaternion conjugated": np.inf, "x/y swapped": np.inf}
    for i in vis:
        m = recs[i, 0:3].astype(np.float64)
        q = recs[i, 4:8].astype(np.float64)
        s = recs[i, 8:11].astype(np.float64)
        R = Rotation.from_quat(q).as_matrix()  # scalar-last (x, y, z, w)
        JW = fd_jacobian(cam, m)
        cov = expected_cov2d(JW, R, s)
        conic = sp["copac"][i, :3].astype(np.float64)
        worst["ok"] = max(worst["ok"], conic_residual(conic, cov))
        # the same expectation under plausible mistakes
        J = JW @ W3.T  # J alone (W is orthonormal)
        worst["W transposed"] = min(worst["W transposed"], conic_residual(conic, expected_cov2d(J @ W3.T, R, s)))
        worst["quaternion conjugated"] = min(worst["quaternion conjugated"],
                                             conic_residual(conic, expected_cov2d(JW, R.T, s)))
        worst["x/y swapped"] = min(worst["x/y swapped"], conic_residual(conic, expected_cov2d(JW[::-1], R, s)))
        # image position and depth against float64 of the same mapping
        px = pix_map(cam, m)
        assert np.allclose(sp["texel"][i, :2], px, rtol=0, atol=2e-4 * max(1.0, np.abs(px).max())), (i, px)
        z = (V @ np.append(m, 1.0))[2]
        assert abs(sp["texel"][i, 2] - z) <= 4e-7 * z
        # radius = ceil(3 sqrt(lambda_max)) with lambda from the exact covariance (skip ceil edges)
        det = cov[0, 0] * cov[1, 1] - cov[0, 1] ** 2
        mid = 0.5 * (cov[0, 0] + cov[1, 1])
        r = 3 * math.sqrt(mid + math.sqrt(max(0.1, mid * mid - det)))
        if abs(r - round(r)) > 1e-3:
            assert sp["texel"][i, 3] == math.ceil(r), (i, r, sp["texel"][i, 3])
    print(f"case {case}: {vis.size} visible, conic residual {worst['ok']:.2e}; "
          + ", ".join(f"{k} {v:.2e}" for k, v in worst.items() if k != "ok"))
    assert worst["ok"] < 2e-4, worst
    # the planted mistakes are caught on the smallest-residual Gaussian already
    for k in ("W transposed", "quaternion conjugated", "x/y swapped"):
        assert worst[k] > 100 * max(worst["ok"], 1e-6), (k, worst)


def test_model_matrix_moves_the_mean_but_not_the_covariance(orc):
    """project.slang:48-57, volume.slang:44-51: the model matrix places the mean; W is the view's
    3x3 alone, so the covariance is not rotated or scaled by the model."""
    rng = np.random.default_rng(7)
    cam, center = random_camera(orc, rng, 640, 480, center_spread=0.0)
    recs = make_scene(rng, 200, np.zeros(3), spread=0.4)
    Rm = Rotation.from_euler("xyz", [0.4, -0.7, 1.1]).as_matrix()
    M = np.eye(4)
    M[:3, :3] = 1.3 * Rm
    M[:3, 3] = [0.1, -0.2, 0.05]
    models = M.astype(np.float32).reshape(1, 16)
    ent = np.zeros(len(recs), np.uint32)
    sp = orc.preprocess(recs, cam, 0, ent_idx=ent, models=models)
    vis = np.nonzero(sp["tiles"] > 0)[0]
    assert vis.size > 50
    Mf = models.reshape(4, 4).astype(np.float64)
    err, err_model_in_w = 0.0, np.inf
    for i in vis:
        m = recs[i, 0:3].astype(np.float64)
        R = Rotation.from_quat(recs[i, 4:8].astype(np.float64)).as_matrix()
        s = recs[i, 8:11].astype(np.float64)
        qw = (Mf @ np.append(m, 1.0))[:3]
        JW = fd_jacobian(cam, qw)  # derivative at the transformed mean, w.r.t. world position
        conic = sp["copac"][i, :3].astype(np.float64)
        err = max(err, conic_residual(conic, expected_cov2d(JW, R, s)))
        err_model_in_w = min(err_model_in_w, conic_residual(conic, expected_cov2d(JW @ Mf[:3, :3], R, s)))
        assert np.allclose(sp["texel"][i, :2], pix_map(cam, qw), atol=2e-3)
    assert err < 2e-4, err
    assert err_model_in_w > 1e-2, err_model_in_w


def test_rotated_scale_swaps_conic_axes(orc):
    """Closed form on the C1 axis (W = I, x = y = 0): scales (s1, s2, s3) rotated 90 degrees about z
    give cov2D = (f/z)^2 diag(s2^2, s1^2) + 0.3 I, i.e. the unrotated conic's a and c swapped."""
    cam = orc.make_camera([0, 0, -3], [0, 0, 0], [0, -1, 0], 1280, 720)
    s1, s2, s3 = 0.08, 0.02, 0.05
    c45 = math.sqrt(0.5)
    g = np.stack([record([0, 0, 0], (0, 0, 0, 1), (s1, s2, s3)), record([0, 0, 0], (0, 0, c45, c45), (s1, s2, s3))])
    sp = orc.preprocess(g, cam, 0)
    f = 0.5 * 720 * math.sqrt(3.0)
    ax, ay = 1 / ((f * s1 / 3) ** 2 + 0.3), 1 / ((f * s2 / 3) ** 2 + 0.3)
    a0, b0, c0 = sp["copac"][0, :3]
    a1, b1, c1 = sp["copac"][1, :3]
    assert abs(a0 - ax) < 1e-6 * ax and abs(c0 - ay) < 1e-6 * ay and b0 == 0
    assert abs(a1 - ay) < 1e-5 * ay and abs(c1 - ax) < 1e-5 * ax and abs(b1) < 1e-6 * ax


def test_off_axis_isotropic_closed_form(orc):
    """An isotropic Gaussian at (x0, 0, 0) seen from (0, 0, -z) (W = I): EWA's J has the off-axis
    term -f x0 / z^2, so cov2D = s^2 f^2 diag(1/z^2 + x0^2/z^4, 1/z^2) + 0.3 I exactly."""
    z = 3.0
    cam = orc.make_camera([0, 0, -z], [0, 0, 0], [0, -1, 0], 1280, 720)
    s = 0.04
    x0s = [-1.2, -0.5, 0.25, 0.9]
    g = np.stack([record([x0, 0, 0], (0, 0, 0, 1), (s, s, s)) for x0 in x0s])
    sp = orc.preprocess(g, cam, 0)
    fy = 0.5 * 720 * float(np.float32(math.sqrt(3.0)))
    fx = 0.5 * 1280 * float(cam.focal_ndc[0])
    for i, x0 in enumerate(x0s):
        cxx = s * s * fx * fx * (1 / z ** 2 + x0 ** 2 / z ** 4) + 0.3
        cyy = s * s * fy * fy / z ** 2 + 0.3
        a, b, c = sp["copac"][i, :3]
        assert abs(a - 1 / cxx) < 2e-6 / cxx and abs(c - 1 / cyy) < 2e-6 / cyy and abs(b) < 1e-7, (x0, a, c)
        # dropping the off-axis term would be a relative change of x0^2 / z^2 in cov_xx
        assert abs(1 / (s * s * fx * fx / z ** 2 + 0.3) - a) > 1e-3 * a


# ------------------------------------------------------------------ spherical harmonics

def sh_basis(l, m, d):
    """3DGS's basis function (l, m) at unit direction d: (-1)^m times the real SH built from scipy."""
    x, y, z = d
    theta = math.acos(max(-1.0, min(1.0, z)))
    phi = math.atan2(y, x)
    Y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        real = math.sqrt(2.0) * (-1) ** m * Y.real
    elif m < 0:
        real = math.sqrt(2.0) * (-1) ** m * Y.imag
    else:
        real = Y.real
    return (-1) ** m * real


def lm_of(k):
    l = int(math.isqrt(k))
    return l, k - l * l - l


def camera_facing(orc, rng, d, target):
    """A camera at target - 4 d looking at target (so the SH direction is d), random up."""
    while True:
        up = rng.normal(size=3)
        if abs(np.dot(up / np.linalg.norm(up), d)) < 0.8:
            break
    eye = (target - 4.0 * d).astype(np.float32)
    return orc.make_camera(eye, target.astype(np.float32), up.astype(np.float32), 64, 64)


def sh_direction(cam, target):
    """normalize(mean - camera position), the camera position from numpy's inverse of the view."""
    pos = np.linalg.inv(cam.view.astype(np.float64))[:3, 3]
    d = target - pos
    return d / np.linalg.norm(d)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_sh_basis_against_scipy(orc, deg):
    rng = np.random.default_rng(20 + deg)
    v = 0.25
    dc = 1.0 / C0  # C0 * dc = 1 keeps every channel far from the max(., 0) clamp
    ks = list(range(1, 16))
    worst, worst_flip = 0.0, np.inf
    for trial in range(16):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        target = rng.uniform(-0.5, 0.5, 3)
        cam = camera_facing(orc, rng, d, target)
        dd = sh_direction(cam, target)
        recs = []
        for k in ks:
            sh = np.zeros(48, np.float32)
            sh[0:3] = dc
            sh[3 + (k % 3) * 15 + (k - 1)] = v
            recs.append(record(target, (0, 0, 0, 1), (0.02, 0.02, 0.02), sh=sh))
        col = orc.preprocess(np.stack(recs), cam, deg)["color"].astype(np.float64)
        typo = sh_basis(3, 0, dd) if deg == 3 else 0.0
        for j, k in enumerate(ks):
            l, m = lm_of(k)
            expect = np.full(3, 1.0 + 0.5 + typo)
            if l <= deg:
                expect[k % 3] += v if k == 12 and deg == 3 else v * sh_basis(l, m, dd)
            worst = max(worst, float(np.abs(col[j] - expect).max()))
            if l <= deg and k != 12:
                flipped = expect.copy()
                flipped[k % 3] -= 2 * v * sh_basis(l, m, dd)
                if abs(sh_basis(l, m, dd)) > 0.05:
                    worst_flip = min(worst_flip, float(np.abs(col[j] - flipped).max()))
    print(f"degree {deg}: max |color - scipy basis| {worst:.2e}, sign-flip distance {worst_flip:.2e}")
    assert worst < 1e-5, worst
    assert worst_flip > 1e-2


def test_sh_band1_x_y_not_swapped(orc):
    """Band 1 read back as a direction: with features 1, 2, 3 = (-s, s, -s) / C1 in channels r, g,
    b the colour is 0.5 + C0 dc + s (y, z, x) -- a swap of x and y shows in channels r and b."""
    rng = np.random.default_rng(5)
    s = 0.2
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        target = rng.uniform(-0.5, 0.5, 3)
        cam = camera_facing(orc, rng, d, target)
        sh = np.zeros(48, np.float32)
        sh[0:3] = 1.0 / C0
        sh[3 + 0 * 15 + 0] = -s / C1   # feature 1 (multiplies -C1 y), red
        sh[3 + 1 * 15 + 1] = s / C1    # feature 2 (multiplies +C1 z), green
        sh[3 + 2 * 15 + 2] = -s / C1   # feature 3 (multiplies -C1 x), blue
        col = orc.preprocess(record(target, (0, 0, 0, 1), (0.02,) * 3, sh=sh)[None], cam, 1)["color"][0]
        dd = sh_direction(cam, target)
        got = (col.astype(np.float64) - 1.5) / s
        assert np.allclose(got, [dd[1], dd[2], dd[0]], atol=2e-5), (got, dd)


def test_camera_position_is_the_inverse_view_translation(orc):
    """getCameraWorldPosition (common.slang:88-92) = -R^T t = inv(V)[:3, 3], read through band 1:
    a wrong camera position (e.g. -R t) moves the direction and so the colour."""
    rng = np.random.default_rng(9)
    s = 0.2
    mutant = []
    for _ in range(20):
        cam, center = random_camera(orc, rng, 320, 240)
        target = center + rng.uniform(-0.2, 0.2, 3)
        sh = np.zeros(48, np.float32)
        sh[0:3] = 1.0 / C0
        sh[3 + 0 * 15 + 0] = -s / C1
        sh[3 + 1 * 15 + 1] = s / C1
        sh[3 + 2 * 15 + 2] = -s / C1
        sp = orc.preprocess(record(target, (0, 0, 0, 1), (0.02,) * 3, sh=sh)[None], cam, 1)
        if sp["tiles"][0] == 0:
            continue
        got = (sp["color"][0].astype(np.float64) - 1.5) / s
        V = cam.view.astype(np.float64)
        pos = np.linalg.inv(V)[:3, 3]
        d = (target - pos) / np.linalg.norm(target - pos)
        assert np.allclose(got, [d[1], d[2], d[0]], atol=2e-5)
        wrong = -(V[:3, :3] @ V[:3, 3])
        dw = (target - wrong) / np.linalg.norm(target - wrong)
        mutant.append(float(np.abs(got - [dw[1], dw[2], dw[0]]).max()))
    assert len(mutant) >= 10
    print(f"-R t instead of -R^T t: colour distance min {min(mutant):.2e}, median {np.median(mutant):.2e}")
    assert min(mutant) > 1e-3 and np.median(mutant) > 2e-2
