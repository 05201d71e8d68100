"""Set-up helpers: put a :class:`chysx.scenes.Scene` on the GPU behind a
:class:`chysx.ClothSimulator` (or many scenes behind a :class:`chysx.ClothBatch`).

Device buffers are torch tensors on ``cuda:0`` (PyTorch is only the
allocator here); the simulator receives their raw addresses exactly as the
reference's callers pass Warp/Newton pointers (README.md:113-130).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import ClothBatch, ClothMaterial, ClothSimulator
from .scenes import Scene


@dataclass
class GpuCloth:
    sim: object
    pos: object       # torch (N,3) float32
    vel: object
    inv_mass: object  # torch (N,) float32
    force: Optional[object]
    scene: Scene

    def positions(self) -> np.ndarray:
        return self.pos.detach().cpu().numpy()

    def velocities(self) -> np.ndarray:
        return self.vel.detach().cpu().numpy()

    def set_state(self, pos: np.ndarray, vel: np.ndarray):
        import torch
        self.pos.copy_(torch.as_tensor(np.asarray(pos, np.float32)))
        self.vel.copy_(torch.as_tensor(np.asarray(vel, np.float32)))


def _material(scene: Scene) -> ClothMaterial:
    m = ClothMaterial()
    m.gx, m.gy, m.gz = (float(g) for g in scene.gravity)
    m.damping = float(scene.damping)
    return m


def make_sim(scene: Scene, tol: float = 1e-6, max_iter: int = 10000, springs_active: bool = True,
             force_active: bool = True, device: str = "cuda:0") -> GpuCloth:
    """Simulator for `scene` with the step's extensions switched on (springs
    in the step, external force, PCG to `tol`).  tol=0 and max_iter=50 with
    springs_active=False, force_active=False give the reference's behaviour."""
    import torch
    dev = torch.device(device)
    pos = torch.as_tensor(scene.pos, dtype=torch.float32).to(dev).contiguous()
    vel = torch.as_tensor(scene.vel, dtype=torch.float32).to(dev).contiguous()
    inv_mass = torch.as_tensor(scene.inv_mass, dtype=torch.float32).to(dev).contiguous()
    force = None
    if scene.force is not None:
        force = torch.as_tensor(scene.force, dtype=torch.float32).to(dev).contiguous()
    sim = ClothSimulator()
    sim.set_material(_material(scene))
    sim.set_external_buffers(pos.data_ptr(), vel.data_ptr(), scene.n, inv_mass.data_ptr(),
                             force.data_ptr() if force is not None else 0)
    sim.set_mesh(scene.tris)
    k_s = scene.spring_k
    k_t = scene.stretch_k
    sim.build_springs_from_current_positions(float(k_s) if np.ndim(k_s) == 0 else 1.0)
    if np.ndim(k_s) != 0:
        sim.set_spring_stiffnesses(np.asarray(k_s, np.float32))
    sim.build_fem_stretch_from_current_positions(float(k_t) if np.ndim(k_t) == 0 else 1.0)
    if np.ndim(k_t) != 0:
        sim.set_fem_stretch_stiffnesses(np.asarray(k_t, np.float32))
    if scene.shear_k is not None:
        sim.build_fem_shear_from_current_positions(float(scene.shear_k))
    if scene.bend_k is not None:
        sim.build_bending_from_current_positions(float(scene.bend_k))
    if scene.pin_idx.size:
        sim.set_pins(scene.pin_idx, scene.pin_targets, scene.pin_k)
    sim.set_springs_active(springs_active)
    sim.set_external_force_active(force_active and force is not None)
    sim.set_pcg_tolerance(tol)
    sim.set_pcg_iterations(max_iter)
    return GpuCloth(sim, pos, vel, inv_mass, force, scene)


@dataclass
class GpuBatch:
    batch: object
    pos: object
    vel: object
    inv_mass: object
    offsets: List[int]
    scenes: List[Scene]

    def positions(self, k: int) -> np.ndarray:
        a = self.offsets[k]
        return self.pos[a:a + self.scenes[k].n].detach().cpu().numpy()


def make_batch(scenes: List[Scene], tol: float = 1e-6, max_iter: int = 10000, device: str = "cuda:0",
               force_active: bool = False) -> GpuBatch:
    import torch
    dev = torch.device(device)
    offs = np.cumsum([0] + [s.n for s in scenes]).tolist()
    pos = torch.as_tensor(np.concatenate([s.pos for s in scenes]), dtype=torch.float32).to(dev).contiguous()
    vel = torch.as_tensor(np.concatenate([s.vel for s in scenes]), dtype=torch.float32).to(dev).contiguous()
    im = torch.as_tensor(np.concatenate([s.inv_mass for s in scenes]), dtype=torch.float32).to(dev).contiguous()
    force = None
    if force_active:
        force = torch.as_tensor(np.concatenate([s.force if s.force is not None else np.zeros_like(s.pos)
                                                for s in scenes]), dtype=torch.float32).to(dev).contiguous()
    b = ClothBatch()
    for s in scenes:
        b.add_cloth(s.tris, s.n, s.spring_k if np.ndim(s.spring_k) else float(s.spring_k),
                    s.stretch_k if np.ndim(s.stretch_k) else float(s.stretch_k), s.pin_idx, s.pin_targets, s.pin_k)
    b.set_external_buffers(pos.data_ptr(), vel.data_ptr(), int(offs[-1]), im.data_ptr(),
                           force.data_ptr() if force is not None else 0)
    for k, s in enumerate(scenes):
        b.set_material(k, _material(s))
    b.set_options(tol, max_iter, True, force_active)
    gb = GpuBatch(b, pos, vel, im, offs, scenes)
    gb.force = force
    return gb
