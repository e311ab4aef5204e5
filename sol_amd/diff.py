"""Differentiable rendering over solhip_backward (adjoint path replay).

Parameters, per material and per light of an uploaded scene:
  albedo    (M, 3)  diffuse albedo kd (diffuse materials, and the diffuse lobe of the mixed
                    diffuse-Phong material)
  log_ns    (M,)    natural log of the Phong exponent ns used to evaluate the BSDF (Phong materials,
                    and the Phong lobe of the mixed material); sampling keeps the scene's
                    ns_sample, so sampling decisions do not move with it
  emission  (L, 3)  intensity of every light
The other entries (albedo of a Phong or specular material, log_ns of a diffuse one) are not read
by the renderer; their gradients are zero.

`render(albedo, log_ns, emission)` is a torch.autograd.Function: the forward pass writes the
tables and renders the image on the GPU (solhip_render into a device tensor); the backward pass
calls solhip_backward with the incoming image gradient. Sampling decisions, MIS weights and
Russian roulette are constants of the gradient (the derivative of the estimator with its random
numbers frozen), so a finite-difference check with common random numbers must disable Russian
roulette, whose survival probability moves with the throughput.

torch serves only as device memory here; all rendering arithmetic runs in the HIP kernels.
"""
import copy
import math

import numpy as np
import torch

from . import _ffi as F
from .gpu import PathTracerConfig


def base_parameters(materials, lights):
    """Parameter arrays (albedo, log_ns, emission) read from material and light tables."""
    albedo = np.array([[m.kd[0], m.kd[1], m.kd[2]] for m in materials], np.float32).reshape(-1, 3)
    log_ns = np.array([math.log(m.ns) if m.ns > 0 else 0.0 for m in materials], np.float32)
    emission = np.array([[l.intensity[0], l.intensity[1], l.intensity[2]] for l in lights], np.float32).reshape(-1, 3)
    return albedo, log_ns, emission


def apply_parameters(materials, lights, albedo, log_ns, emission):
    """Copies of the tables with kd = albedo (diffuse and mixed), ns = exp(log_ns) (Phong and mixed;
    ns_sample unchanged) and intensity = emission: exactly the entries solhip_backward writes
    gradients for."""
    albedo = np.asarray(albedo, np.float32)
    log_ns = np.asarray(log_ns, np.float32)
    emission = np.asarray(emission, np.float32)
    mats = []
    for i, m in enumerate(materials):
        m2 = copy.copy(m)
        if m.type in (F.BSDF_DIFFUSE, F.BSDF_INTERP):
            for c in range(3):
                m2.kd[c] = float(albedo[i, c])
        if m.type in (F.BSDF_PHONG, F.BSDF_INTERP):
            m2.ns = float(np.float32(math.exp(float(log_ns[i]))))
        mats.append(m2)
    ls = []
    for i, l in enumerate(lights):
        l2 = copy.copy(l)
        for c in range(3):
            l2.intensity[c] = float(emission[i, c])
        ls.append(l2)
    return mats, ls


class DiffRenderer:
    """An uploaded scene, an image size and a sample range, rendered as a differentiable function of
    (albedo, log_ns, emission)."""

    def __init__(self, gscene, width, height, sample_index, sample_count, config=None, renderer=F.RENDERER_WAVEFRONT):
        self.gscene = gscene
        self.width, self.height = width, height
        self.sample_index, self.sample_count = sample_index, sample_count
        self.config = config or PathTracerConfig()
        self.renderer = renderer
        self.materials = gscene.scene.materials()
        self.lights = gscene.scene.lights()

    def parameters(self):
        return base_parameters(self.materials, self.lights)

    def set_parameters(self, albedo, log_ns, emission):
        mats, lights = apply_parameters(self.materials, self.lights, albedo, log_ns, emission)
        self.gscene.set_materials(mats)
        self.gscene.set_lights(lights)

    def render_with(self, albedo, log_ns, emission):
        """Image (3, H, W) float32 device tensor for the given parameters (no autograd)."""
        self.set_parameters(albedo, log_ns, emission)
        img = torch.zeros((3, self.height, self.width), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.gscene.render_device(img.data_ptr(), self.width, self.height, self.sample_index, self.sample_count,
                                  self.config, renderer=self.renderer)
        return img

    def gradients(self, d_image, albedo, log_ns, emission):
        """(d_albedo (M, 3), d_log_ns (M,), d_emission (L, 3)) device tensors for dLoss/dImage."""
        self.set_parameters(albedo, log_ns, emission)
        d_image = d_image.detach().to(device="cuda", dtype=torch.float32).contiguous()
        dm = torch.zeros((len(self.materials), 4), dtype=torch.float32, device="cuda")
        dl = torch.zeros((len(self.lights), 3), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.gscene.backward_device(d_image.data_ptr(), self.width, self.height, self.sample_index,
                                    self.sample_count, dm.data_ptr(), dl.data_ptr(), self.config)
        # kd gradients exist for diffuse and mixed materials, log ns for Phong and mixed ones: the
        # entries apply_parameters moves (the kernel writes no others)
        return dm[:, :3].contiguous(), dm[:, 3].contiguous(), dl

    def __call__(self, albedo, log_ns, emission):
        return _Render.apply(self, albedo, log_ns, emission)


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, albedo, log_ns, emission):
        a = albedo.detach().cpu().numpy()
        n = log_ns.detach().cpu().numpy()
        e = emission.detach().cpu().numpy()
        ctx.renderer = renderer
        ctx.params = (a, n, e)
        ctx.devices = (albedo.device, log_ns.device, emission.device)
        return renderer.render_with(a, n, e)

    @staticmethod
    def backward(ctx, grad_image):
        da, dn, de = ctx.renderer.gradients(grad_image, *ctx.params)
        return None, da.to(ctx.devices[0]), dn.to(ctx.devices[1]), de.to(ctx.devices[2])
