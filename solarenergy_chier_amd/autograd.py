"""Optional PyTorch binding of the differentiable trace: the bilinear flux map as a function of the
per-heliostat aim-point offsets and the sun direction, differentiable with torch.autograd.

    from solarenergy_chier_amd.autograd import FluxMap
    fm = FluxMap(device, scene, grid=0)
    flux = fm(offsets, sun_dir)          # (res_y, res_x) float32 on cuda
    loss = ((flux - 7e5) ** 2).sum()
    loss.backward()                      # offsets.grad (n_heliostats, 2), sun_dir.grad (3,)

Forward is ``qmc_adjoint_forward`` (visibility recorded), backward is ``qmc_adjoint_backward`` fed
torch's dL/dflux: any loss written in torch composes with the hand-written adjoint. The scene's aim
offsets and sun direction are host-side scene parameters, so each forward re-processes the scene
and refreshes the device state (``qmc_update``); both tensors are read on the host. The library runs
on its own HIP stream: the binding synchronises torch's stream before handing it a device buffer and
the library's stream before torch reads one. The product path is the C ABI; torch is only the
caller here.

Import this module (or torch, with its GPU initialised) before the first ``Device`` of the process:
this PyTorch ships its own HIP runtime, the library then shares it; the other way round the
library has already started the system ROCm runtime, and a second runtime in the process cannot
open the GPU. The import checks this and raises.
"""
from __future__ import annotations

import numpy as np
import torch

from .scene import Device, Scene

if not torch.cuda.is_available():
    raise RuntimeError(
        "solarenergy_chier_amd.autograd: torch sees no GPU. Import it (or initialise torch's GPU) before "
        "the first Device of the process, so that the library and torch share one HIP runtime.")
torch.cuda.init()


class _FluxFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, offsets: torch.Tensor, sun_dir: torch.Tensor, fm: "FluxMap") -> torch.Tensor:
        params = (np.array(offsets.detach().cpu().numpy(), np.float32), np.array(sun_dir.detach().cpu().numpy(), np.float32))
        flux = fm._forward(*params)
        ctx.fm = fm
        ctx.params = params
        ctx.stamp = fm._stamp
        ctx.shapes = (offsets.shape, offsets.dtype, offsets.device, sun_dir.dtype, sun_dir.device)
        return flux

    @staticmethod
    def backward(ctx, grad_flux: torch.Tensor):
        fm: FluxMap = ctx.fm
        oshape, odtype, odev, sdtype, sdev = ctx.shapes
        if fm._stamp != ctx.stamp:
            # another forward ran since this one (e.g. loss = f(x1) + f(x2)): the device holds its
            # parameters and visibility mask. Restore this forward's: the same parameters give the same
            # rays, so the recomputed mask is the one this forward recorded.
            fm._forward(*ctx.params)
        g = grad_flux.detach().to(device=fm.torch_device, dtype=torch.float32).contiguous()
        torch.cuda.synchronize(fm.torch_device)
        grad_aim, grad_sun = fm.device.adjoint_backward(fm.scene, fm.grid, g.data_ptr())
        ga = torch.from_numpy(np.ascontiguousarray(grad_aim[fm.helio_slice])).reshape(oshape).to(odev, odtype)
        gs = torch.from_numpy(np.ascontiguousarray(grad_sun)).to(sdev, sdtype)
        return ga, gs, None


class FluxMap:
    """Differentiable flux map of grid `grid` of an uploaded scene (flat heliostats, rectangle
    receiver). Offsets are (n, 2) for the grid's n heliostats (across, up in metres); the sun direction
    is the (3,) light travel direction used as given."""

    def __init__(self, device: Device, scene: Scene, grid: int = 0):
        self.device = device
        self.scene = scene
        self.grid = grid
        gd = scene.grid_desc(grid)
        self.helio_slice = slice(gd.start_heliostat, gd.start_heliostat + gd.n_heliostats)
        self.receiver = gd.receiver
        self.shape = device.flux_shape(self.receiver)
        self.torch_device = torch.device("cuda", device.device_id)
        self._all = np.zeros((scene.counts.n_heliostats, 2), np.float32)
        for h in range(scene.counts.n_heliostats):
            self._all[h] = scene.heliostat_desc(h).aim_offset[:]
        self._stamp = 0  # counts forwards: which one the device state (parameters, mask) belongs to

    def _forward(self, offsets: np.ndarray, sun_dir: np.ndarray) -> torch.Tensor:
        """Set the parameters, run the differentiable forward (recording the visibility mask)."""
        self._set_parameters(offsets, sun_dir)
        flux = torch.empty(self.shape, dtype=torch.float32, device=self.torch_device)
        torch.cuda.synchronize(self.torch_device)
        self.device.adjoint_forward(self.grid, flux.data_ptr())
        self.device.sync()
        self._stamp += 1
        return flux

    def _set_parameters(self, offsets: np.ndarray, sun_dir: np.ndarray) -> None:
        self._all[self.helio_slice] = np.asarray(offsets, np.float32).reshape(-1, 2)
        self.scene.set_aim_offsets(self._all)
        cfg = self.scene.config
        cfg.sun_dir[:] = [float(x) for x in np.asarray(sun_dir, np.float32).reshape(3)]
        self.scene.set_config(cfg)
        self.scene.process()
        self.device.update(self.scene)

    def __call__(self, offsets: torch.Tensor, sun_dir: torch.Tensor) -> torch.Tensor:
        return _FluxFn.apply(offsets, sun_dir, self)
