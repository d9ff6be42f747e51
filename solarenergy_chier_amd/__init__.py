"""MI355X-native quasi-Monte-Carlo heliostat-field ray tracer.

The compute path is ``lib/libqmcrt.so`` (hand-written HIP kernels for gfx950 behind the C ABI of
``include/qmcrt.h``); this package binds it with ctypes and mirrors the reference call surface
(``SolarScene``/``SceneLoader``/``SceneConfiguration``/``SceneProcessor``/
``QuasiMonteCarloRayTracer``/``ImageSmoother``/``RayTracingPipeline``) in :mod:`.api`.
"""
from . import _native
from .scene import Device, GridSpec, HeliostatSpec, ReceiverSpec, Scene, default_config, make_config

__all__ = ["Device", "Scene", "ReceiverSpec", "GridSpec", "HeliostatSpec", "default_config", "make_config",
           "_native"]
