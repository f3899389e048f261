"""torch.autograd.Function around the HIP forward and backward kernels.

    image = rasterize(means, log_scales, quats, opacity_logits, sh, engine, camera, sh_degree)

Parameters are CUDA (HIP) float tensors in the training parameterisation: means (N, 3),
log-scales (N, 3), quaternions (N, 4, xyzw, fed unnormalised as volume.slang:28-35 does),
opacity logits (N,), SH (N, 48).  They are packed into the 60-float GaussianPoint records
(GaussianGeometry.h:10-30: opacity = sigmoid(logit), scale = exp(log-scale), modifier 1) on the
device, rendered by tpd_render, and differentiated by tpd_backward, which recomputes the per-pixel
transmittance instead of storing it.  Returns the FP32 image (H, W, 3).

torch must be imported before the library is first loaded (torch carries its own HIP runtime and
the library binds to it); importing this module does that.
"""
from __future__ import annotations

import torch

from . import _capi
from .engine import GaussianEngine, PerspectiveCamera


def _check_runtime():
    if _capi._LIB is not None and not _capi.LOADED_AFTER_TORCH:
        raise RuntimeError("libtpd_hip.so was loaded before torch; import torch (or torpedo_amd.autograd) first")


def pack_records(means, log_scales, quats, opacity_logits, sh) -> torch.Tensor:
    n = means.shape[0]
    return torch.cat([means.reshape(n, 3), torch.sigmoid(opacity_logits.reshape(n, 1)), quats.reshape(n, 4),
                      torch.exp(log_scales.reshape(n, 3)), torch.ones(n, 1, device=means.device, dtype=means.dtype),
                      sh.reshape(n, 48)], 1).to(torch.float32).contiguous()


class RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, engine: GaussianEngine,
                camera: PerspectiveCamera, sh_degree: int):
        _check_runtime()
        stream = torch.cuda.current_stream(means.device).cuda_stream
        rec = pack_records(means, log_scales, quats, opacity_logits, sh)
        n = rec.shape[0]
        engine.bindDeviceScene(rec.data_ptr(), n, sh_degree)
        view, pv, fo = camera.uniforms()
        engine.setCamera(view, pv, fo, camera.width, camera.height)
        engine.render(stream)
        img = torch.empty(camera.height, camera.width, 3, device=means.device, dtype=torch.float32)
        engine.copyImageTo(img.data_ptr(), stream)
        ctx.engine = engine
        ctx.records = rec  # the engine renders from this memory; keep it alive for the backward
        ctx.frame = engine.frameCount
        ctx.shapes = (means.shape, log_scales.shape, quats.shape, opacity_logits.shape, sh.shape)
        ctx.dtypes = (means.dtype, log_scales.dtype, quats.dtype, opacity_logits.dtype, sh.dtype)
        return img

    @staticmethod
    def backward(ctx, grad_image):
        engine = ctx.engine
        if engine.frameCount != ctx.frame:
            raise RuntimeError("rasterize backward: the engine rendered another frame after this one")
        dev = grad_image.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        g = grad_image.to(torch.float32).contiguous()
        n = ctx.records.shape[0]
        gm = torch.empty(n, 3, device=dev, dtype=torch.float32)
        gls = torch.empty(n, 3, device=dev, dtype=torch.float32)
        gq = torch.empty(n, 4, device=dev, dtype=torch.float32)
        glo = torch.empty(n, device=dev, dtype=torch.float32)
        gsh = torch.empty(n, 48, device=dev, dtype=torch.float32)
        engine.backward(g.data_ptr(), gm.data_ptr(), gls.data_ptr(), gq.data_ptr(), glo.data_ptr(), gsh.data_ptr(),
                        stream=stream)
        outs = [t.reshape(s).to(d) for t, s, d in zip((gm, gls, gq, glo, gsh), ctx.shapes, ctx.dtypes)]
        return (*outs, None, None, None)


def rasterize(means, log_scales, quats, opacity_logits, sh, engine: GaussianEngine, camera: PerspectiveCamera,
              sh_degree: int = 3) -> torch.Tensor:
    return RasterizeGaussians.apply(means, log_scales, quats, opacity_logits, sh, engine, camera, sh_degree)
