"""Differentiable rendering of material parameters by path replay.

forward(materials) renders the resident scene with the given material table;
backward(grad_image) returns dLoss/dparameters by re-tracing every path of the
same frame (mrad_backward).  Gradients come per primitive (the reference's
global order) and are summed per material for the material-table parameters.

PathReplayFunction is the torch.autograd.Function form of the same pair:
    image = PathReplayFunction.apply(albedo, emission, renderer)
with albedo and emission (materials x 3) CPU tensors.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .scene import GpuScene
from .tracer import GpuCamera, GpuRayTracer


def primitive_material_indices(scene: GpuScene) -> np.ndarray:
    """Material index of every primitive in the reference's global order
    (spheres, boxes, affine spheres, affine boxes, planes)."""
    parts = [scene.sphere_data()[:, 4], scene.box_data()[:, 6], scene.affine_sphere_data()[:, 21],
             scene.affine_box_data()[:, 21], scene.plane_data()[:, 23]]
    return np.concatenate(parts).astype(np.int64)


def l2_loss(image: np.ndarray, target: np.ndarray) -> tuple[float, np.ndarray]:
    """Mean squared pixel error and its gradient with respect to the image."""
    diff = image.astype(np.float64) - target.astype(np.float64)
    n = diff.size
    return float(np.sum(diff * diff) / n), (2.0 * diff / n).astype(np.float32)


class PathReplayRenderer:
    """Forward/backward pair over one tracer and one frame (camera, seed)."""

    def __init__(self, scene: GpuScene, camera: GpuCamera, width: int, height: int, spp: int, max_depth: int,
                 frame_index: int = 0, frame_seed: int = 0x5EED, rr_start: int = -1, bvh_stack_size: int = 32):
        self.scene, self.camera = scene, camera
        self.width, self.height, self.spp, self.max_depth = width, height, spp, max_depth
        self.frame_index, self.frame_seed = frame_index, frame_seed
        self.tracer = GpuRayTracer(width, height, spp, True, bvh_stack_size=bvh_stack_size)
        if rr_start >= 0:
            self.tracer.set_russian_roulette(rr_start)
        self.tracer.ensure_scene_resident(scene)
        self.base_materials = scene.material_data().copy()
        self.prim_material = primitive_material_indices(scene)
        self.primitive_count = len(self.prim_material)
        self._cam = camera.packed() if isinstance(camera, GpuCamera) else np.asarray(camera, np.float32)

    def close(self):
        self.tracer.close()

    def materials_with(self, albedo=None, emission=None) -> np.ndarray:
        m = self.base_materials.copy()
        if albedo is not None:
            m[:, 0:3] = np.asarray(albedo, np.float32).reshape(-1, 3)
        if emission is not None:
            m[:, 9:12] = np.asarray(emission, np.float32).reshape(-1, 3)
        return m

    def forward(self, materials: np.ndarray | None = None) -> np.ndarray:
        if materials is not None:
            m = np.ascontiguousarray(materials, np.float32).reshape(-1, 13)
            N.check(N.lib.mrad_update_materials(self.tracer._handle(), N.ptr(m, N.fp), len(m)))
        img = np.zeros(self.width * self.height * 3, np.float32)
        self.tracer.render_sample(img, self.scene, self._cam, self.max_depth, self.frame_index, self.frame_seed)
        return img

    def backward(self, grad_image: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Per-primitive (grad_albedo, grad_emission), each primitive_count x 3."""
        g = np.ascontiguousarray(grad_image, np.float32).reshape(-1)
        if g.size != self.width * self.height * 3:
            raise N.InvalidArgument(N.MRAD_ERR_INVALID_ARGUMENT, "gradient image has the wrong size")
        ga = np.zeros(self.primitive_count * 3, np.float32)
        ge = np.zeros(self.primitive_count * 3, np.float32)
        N.check(N.lib.mrad_backward(self.tracer._handle(), N.ptr(g, N.fp), N.ptr(self._cam, N.fp), self.max_depth,
                                    self.frame_index, self.frame_seed & ((1 << 64) - 1), N.ptr(ga, N.fp),
                                    N.ptr(ge, N.fp), self.primitive_count))
        return ga.reshape(-1, 3), ge.reshape(-1, 3)

    def per_material(self, grad_prim: np.ndarray) -> np.ndarray:
        out = np.zeros((len(self.base_materials), 3), np.float64)
        valid = self.prim_material >= 0
        np.add.at(out, self.prim_material[valid], grad_prim[valid].astype(np.float64))
        return out

    def last_kernel_nanos(self) -> int:
        return self.tracer.last_frame_stats()["kernel_nanos"]


try:  # torch is plumbing only: the autograd form is offered when it is importable
    import torch

    class PathReplayFunction(torch.autograd.Function):
        """image = PathReplayFunction.apply(albedo (M,3), emission (M,3), renderer)."""

        @staticmethod
        def forward(ctx, albedo, emission, renderer: PathReplayRenderer):
            m = renderer.materials_with(albedo.detach().cpu().numpy(), emission.detach().cpu().numpy())
            img = renderer.forward(m)
            ctx.renderer = renderer
            return torch.from_numpy(img.reshape(renderer.height, renderer.width, 3).copy())

        @staticmethod
        def backward(ctx, grad_image):
            r = ctx.renderer
            ga, ge = r.backward(grad_image.detach().cpu().numpy().reshape(-1))
            return (torch.from_numpy(r.per_material(ga).astype(np.float32)),
                    torch.from_numpy(r.per_material(ge).astype(np.float32)), None)
except ImportError:  # pragma: no cover
    PathReplayFunction = None
