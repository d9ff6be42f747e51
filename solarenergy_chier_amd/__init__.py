"""MI355X-native quasi-Monte-Carlo heliostat-field ray tracer.

The compute path is ``lib/libqmcrt.so`` (hand-written HIP kernels for gfx950 behind the C ABI of
``include/qmcrt.h``); this package binds it with ctypes (:mod:`._native`) and wraps the two C-ABI
objects for numpy callers (:mod:`.scene`: ``Scene`` = qmc_scene, ``Device`` = qmc_ctx). The reference's
C++ call surface (``SolarScene``, ``SceneLoader``, ``SceneProcessor``, ``QuasiMonteCarloRayTracer``,
``ImageSmoother``, ``RayTracingPipeline``) is the host C++ facade ``include/qmcrt_facade.hpp``.
"""
from . import _native
from .scene import Device, GridSpec, HeliostatSpec, ReceiverSpec, Scene, default_config, make_config

__all__ = ["Device", "Scene", "ReceiverSpec", "GridSpec", "HeliostatSpec", "default_config", "make_config",
           "_native"]
