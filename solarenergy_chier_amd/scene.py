"""Numpy-facing wrappers of the two C-ABI objects: :class:`Scene` (host ``qmc_scene``) and
:class:`Device` (``qmc_ctx``: one HIP device, one stream, the device copy of a scene)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _native as N


def _f3(v) -> tuple:
    return tuple(float(x) for x in v)


@dataclass
class ReceiverSpec:
    """Recv block of a .scn file (Scene/SceneLoader.cpp:104-127) plus the res/tilt extensions."""
    type: int = 0
    pos: tuple = (0.0, 0.0, 0.0)
    size: tuple = (1.0, 1.0, 1.0)
    normal: tuple = (0.0, 0.0, 1.0)
    face: int = 0
    res_x: int = 0
    res_y: int = 0
    tilt_deg: float = 0.0

    def to_c(self) -> N.ReceiverDesc:
        d = N.ReceiverDesc()
        d.type = self.type
        d.pos[:] = _f3(self.pos)
        d.size[:] = _f3(self.size)
        d.normal[:] = _f3(self.normal)
        d.face = self.face
        d.res_x, d.res_y = self.res_x, self.res_y
        d.tilt_deg = self.tilt_deg
        return d


@dataclass
class GridSpec:
    """Grid block (Scene/SceneLoader.cpp:130-159)."""
    pos: tuple
    size: tuple
    interval: tuple
    start_heliostat: int
    n_heliostats: int
    heliostat_type: int = 0
    receiver: int = 0
    type: int = 0

    def to_c(self) -> N.GridDesc:
        d = N.GridDesc()
        d.type = self.type
        d.pos[:] = _f3(self.pos)
        d.size[:] = _f3(self.size)
        d.interval[:] = _f3(self.interval)
        d.start_heliostat = self.start_heliostat
        d.n_heliostats = self.n_heliostats
        d.heliostat_type = self.heliostat_type
        d.receiver = self.receiver
        return d


@dataclass
class HeliostatSpec:
    """helio line + sticky modifiers (Scene/SceneLoader.cpp:161-198,280-333)."""
    pos: tuple
    size: tuple
    grid: int = 0
    rows: int = 1
    cols: int = 1
    gap: tuple = (0.0, 0.0)
    surface_property: tuple = (-1.0, -1.0, -1.0, -1.0, -1.0, -1.0)
    sub_count_override: int = 0
    aim_offset: tuple = (0.0, 0.0)
    local_centers: np.ndarray | None = None
    local_normals: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    def to_c(self) -> N.HeliostatDesc:
        d = N.HeliostatDesc()
        d.pos[:] = _f3(self.pos)
        d.size[:] = _f3(self.size)
        d.rows, d.cols = self.rows, self.cols
        d.gap[:] = _f3(self.gap)
        d.surface_property[:] = _f3(self.surface_property)
        d.grid = self.grid
        d.sub_count_override = self.sub_count_override
        d.aim_offset[:] = _f3(self.aim_offset)
        self._keep.clear()
        if self.local_centers is not None:
            a = np.ascontiguousarray(self.local_centers, dtype=np.float32).reshape(-1)
            self._keep.append(a)
            d.local_centers = N.fptr(a)
        if self.local_normals is not None:
            a = np.ascontiguousarray(self.local_normals, dtype=np.float32).reshape(-1)
            self._keep.append(a)
            d.local_normals = N.fptr(a)
        return d


def _as_c(spec):
    """A spec dataclass or an already-built C descriptor."""
    return spec if isinstance(spec, C.Structure) else spec.to_c()


def check_cylinder_columns(width: int, stride: int = 1) -> tuple[int, int, float]:
    """Host check of the trace's cylinder column lookup for a `width`-column map over the float
    cosines in [-1, 1] (every one with stride 1; else every stride-th one and those around each
    breakpoint) and both signs: (pairs checked, mismatches against the double arc cosine's column, the
    acos estimate's largest error in radians). Needs no GPU."""
    n, bad, err = C.c_uint64(0), C.c_uint64(0), C.c_double(0.0)
    N.call("qmc_check_cylinder_columns", int(width), int(stride), C.byref(n), C.byref(bad), C.byref(err))
    return int(n.value), int(bad.value), float(err.value)


def default_config() -> N.Config:
    c = N.Config()
    N.call("qmc_config_default", C.byref(c))
    return c


def make_config(**kw) -> N.Config:
    c = default_config()
    for k, v in kw.items():
        if k == "sun_dir":
            c.sun_dir[:] = _f3(v)
        else:
            setattr(c, k, v)
    return c


class Scene:
    """Host-side scene (qmc_scene): configuration, receivers, grids, heliostats, processed geometry."""

    def __init__(self):
        h = C.c_void_p()
        N.call("qmc_scene_create", C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            N.lib().qmc_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # construction -----------------------------------------------------------------------------
    @classmethod
    def from_files(cls, scn_path: str, json_path: str | None = None, process: bool = True) -> "Scene":
        s = cls()
        if json_path:
            s.load_config(json_path)
        N.call("qmc_scene_load_file", s.handle, str(scn_path).encode())
        if process:
            s.process()
        return s

    @classmethod
    def from_specs(cls, config: N.Config, receivers: Sequence[ReceiverSpec], grids: Sequence[GridSpec],
                   heliostats: Sequence[HeliostatSpec], process: bool = True) -> "Scene":
        s = cls()
        s.set_config(config)
        for r in receivers:
            N.call("qmc_scene_add_receiver", s.handle, C.byref(_as_c(r)))
        for g in grids:
            N.call("qmc_scene_add_grid", s.handle, C.byref(_as_c(g)))
        for h in heliostats:
            N.call("qmc_scene_add_heliostat", s.handle, C.byref(_as_c(h)))
        if process:
            s.process()
        return s

    def load_config(self, path: str) -> int:
        n = C.c_int32(0)
        N.call("qmc_scene_load_config", self.handle, str(path).encode(), C.byref(n))
        return n.value

    def set_config(self, cfg: N.Config) -> None:
        N.call("qmc_scene_set_config", self.handle, C.byref(cfg))

    @property
    def config(self) -> N.Config:
        c = N.Config()
        N.call("qmc_scene_get_config", self.handle, C.byref(c))
        return c

    def process(self) -> None:
        N.call("qmc_scene_process", self.handle)

    def sunshape_table_buie(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """Physical Buie sampler table of this configuration: bin edges [mrad], cumulative
        probability, and the sky density at each bin's two ends."""
        n = C.c_int32(0)
        N.call("qmc_sunshape_table_buie", self.handle, None, None, None, None, C.byref(n))
        th, cdf = np.zeros(n.value + 1), np.zeros(n.value + 1)
        lo, hi = np.zeros(n.value), np.zeros(n.value)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        N.call("qmc_sunshape_table_buie", self.handle, dp(th), dp(cdf), dp(lo), dp(hi), C.byref(n))
        return th, cdf, lo, hi

    # inspection -------------------------------------------------------------------------------
    @property
    def counts(self) -> N.SceneCounts:
        c = N.SceneCounts()
        N.call("qmc_scene_get_counts", self.handle, C.byref(c))
        return c

    def receiver_desc(self, r: int) -> N.ReceiverDesc:
        d = N.ReceiverDesc()
        N.call("qmc_scene_get_receiver_desc", self.handle, r, C.byref(d))
        return d

    def grid_desc(self, g: int) -> N.GridDesc:
        d = N.GridDesc()
        N.call("qmc_scene_get_grid_desc", self.handle, g, C.byref(d))
        return d

    def heliostat_desc(self, h: int) -> N.HeliostatDesc:
        d = N.HeliostatDesc()
        N.call("qmc_scene_get_heliostat_desc", self.handle, h, C.byref(d))
        return d

    def receiver_info(self, r: int) -> N.ReceiverInfo:
        o = N.ReceiverInfo()
        N.call("qmc_scene_get_receiver_info", self.handle, r, C.byref(o))
        return o

    def heliostat_info(self, h: int) -> N.HeliostatInfo:
        o = N.HeliostatInfo()
        N.call("qmc_scene_get_heliostat_info", self.handle, h, C.byref(o))
        return o

    def grid_info(self, g: int) -> N.GridInfo:
        o = N.GridInfo()
        N.call("qmc_scene_get_grid_info", self.handle, g, C.byref(o))
        return o

    def grid_cells(self, g: int) -> tuple[np.ndarray, np.ndarray]:
        gi = self.grid_info(g)
        cs = np.zeros(gi.nx * gi.ny * gi.nz + 1, np.int32)
        m = np.zeros(max(gi.n_match, 1), np.int32)
        N.call("qmc_scene_get_grid_cells", self.handle, g, N.iptr(cs), N.iptr(m))
        return cs, m[: gi.n_match]

    def sub_vertices(self, g: int) -> np.ndarray:
        n = C.c_int32(0)
        N.call("qmc_scene_get_subheliostat_vertices", self.handle, g, None, 0, C.byref(n))
        out = np.zeros((max(n.value, 1), 3), np.float32)
        N.call("qmc_scene_get_subheliostat_vertices", self.handle, g, N.fptr(out), n.value, C.byref(n))
        return out[: n.value]

    def factor(self, h: int) -> float:
        f = C.c_float(0)
        N.call("qmc_scene_get_factor", self.handle, h, C.byref(f))
        return f.value

    def set_aim_offsets(self, offsets: np.ndarray) -> None:
        """Aim-point offsets (heliostats x 2: across, up) in the receiver surface; re-process after."""
        a = np.ascontiguousarray(offsets, np.float32).reshape(-1)
        N.call("qmc_scene_set_aim_offsets", self.handle, N.fptr(a), len(a))

    def check_heliostat(self, h: int) -> None:
        N.call("qmc_scene_check_heliostat", self.handle, h)

    def load_tasks(self, path: str) -> tuple[list[int], bool]:
        nh = self.counts.n_heliostats
        ids = np.zeros(max(nh, 1), np.int32)
        n = C.c_int32(0)
        clean = C.c_int32(0)
        N.call("qmc_task_load", self.handle, str(path).encode(), N.iptr(ids), len(ids), C.byref(n), C.byref(clean))
        return [int(x) for x in ids[: n.value]], bool(clean.value)

    def specs(self):
        """Descriptor arrays (C structs) of the loaded scene: (config, receivers, grids, heliostats)."""
        c = self.counts
        R = (N.ReceiverDesc * max(c.n_receivers, 1))()
        G = (N.GridDesc * max(c.n_grids, 1))()
        H = (N.HeliostatDesc * max(c.n_heliostats, 1))()
        for i in range(c.n_receivers):
            R[i] = self.receiver_desc(i)
        for i in range(c.n_grids):
            G[i] = self.grid_desc(i)
        for i in range(c.n_heliostats):
            H[i] = self.heliostat_desc(i)
        return self.config, R, c.n_receivers, G, c.n_grids, H, c.n_heliostats


class Device:
    """One HIP device context (qmc_ctx) holding the device copy of one processed scene."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        N.call("qmc_create", device, C.byref(h))
        self.handle = h
        self.device_id = device
        self.scene: Scene | None = None

    def close(self):
        if self.handle:
            N.lib().qmc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: Scene) -> None:
        N.call("qmc_upload", self.handle, scene.handle)
        self.scene = scene

    @property
    def stream(self) -> int:
        p = C.c_void_p()
        N.call("qmc_get_stream", self.handle, C.byref(p))
        return int(p.value or 0)

    def sync(self) -> None:
        N.call("qmc_sync", self.handle)

    # pools and offsets --------------------------------------------------------------------------
    def pool_generate(self, seed_sunshape: int = 2, seed_perturbation: int = 4,
                      sunshape_model: int = N.SUNSHAPE_REFERENCE) -> None:
        """Seeded sun-shape and slope-error pools; sunshape_model SUNSHAPE_BUIE draws the physical
        Buie profile instead of the reference's sampler (which leaves 4.65-9.3 mrad empty)."""
        p = N.PoolParams(seed_sunshape, seed_perturbation, sunshape_model, 0)
        N.call("qmc_pool_generate", self.handle, C.byref(p))

    def pool_set(self, dirs: np.ndarray, perts: np.ndarray, L: int) -> None:
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        p = np.ascontiguousarray(perts, np.float32).reshape(-1, 3)
        N.call("qmc_pool_set", self.handle, N.fptr(d), N.fptr(p), d.shape[0], L)

    def pool_from_uniforms(self, ut, up, gx, gz) -> None:
        arrs = [np.ascontiguousarray(a, np.float32) for a in (ut, up, gx, gz)]
        N.call("qmc_pool_from_uniforms", self.handle, *[N.fptr(a) for a in arrs])

    def pool_get(self, P: int) -> tuple[np.ndarray, np.ndarray]:
        d = np.zeros((P, 3), np.float32)
        p = np.zeros((P, 3), np.float32)
        N.call("qmc_pool_get", self.handle, N.fptr(d), N.fptr(p), P)
        return d, p

    def offsets_generate(self, seed: int = 3) -> None:
        N.call("qmc_offsets_generate", self.handle, C.c_uint64(seed))

    def offsets_resample(self, enable: bool, seed: int = 0) -> None:
        """Per-call resampling of the group offsets (the reference draws fresh offsets on every
        rayTracing call): when on, trace number k since this call first regenerates the offsets of the
        heliostats it traces with seed + k (one k per pass: a heliostat listed twice is traced twice)."""
        N.call("qmc_offsets_resample", self.handle, 1 if enable else 0, C.c_uint64(seed))

    def offsets_set(self, h: int, offsets: np.ndarray) -> None:
        a = np.ascontiguousarray(offsets, np.int32)
        N.call("qmc_offsets_set", self.handle, h, N.iptr(a), len(a))

    def offsets_get(self, h: int, F: int) -> np.ndarray:
        a = np.zeros(F + 1, np.int32)
        N.call("qmc_offsets_get", self.handle, h, N.iptr(a), F + 1)
        return a

    def facets(self, h: int, F: int) -> tuple[np.ndarray, np.ndarray]:
        o = np.zeros((F, 3), np.float32)
        n = np.zeros((F, 3), np.float32)
        N.call("qmc_facets_get", self.handle, h, N.fptr(o), N.fptr(n), F)
        return o, n

    # tracing ---------------------------------------------------------------------------------------
    def trace(self, ids: Iterable[int]) -> None:
        a = np.ascontiguousarray(list(ids), np.int32)
        N.call("qmc_trace", self.handle, N.iptr(a), len(a))

    def set_defer_capacity(self, rays: int) -> None:
        """Cap the queue of rays a map-path launch defers to its exact re-trace (0: sized from the
        launch); a launch that defers more re-traces all its rays instead (same results, slower)."""
        N.call("qmc_set_defer_capacity", self.handle, C.c_int64(rays))

    def defer_count(self) -> int:
        """Rays the last map-path launch deferred (waits for the context's stream)."""
        n = C.c_int64(0)
        N.call("qmc_get_defer_count", self.handle, C.byref(n))
        return int(n.value)

    def probe_atomic_rates(self) -> dict:
        """Measured rates (G ops/s) of scattered float / 64-bit global and float LDS atomic adds."""
        r = N.AtomicRates()
        N.call("qmc_probe_atomic_rates", self.handle, C.byref(r))
        return {"global_f32": r.global_f32, "global_f64": r.global_f64, "lds_f32": r.lds_f32}

    def check_fast_sqrt(self) -> tuple[int, list[int]]:
        """Exhaustive check of the short square-root sequences over [2^-8, 2^8]: (arguments checked,
        [short sqrt mismatches, short 1/sqrt mismatches, general-vs-double mismatches])."""
        n = C.c_uint64(0)
        bad = (C.c_uint64 * 3)()
        N.call("qmc_check_fast_sqrt", self.handle, C.byref(n), bad)
        return int(n.value), [int(b) for b in bad]

    def cull_flags(self, ids: Iterable[int]) -> np.ndarray:
        """Per heliostat: bit 0 = shadow traversal needed, bit 1 = blocking traversal needed."""
        a = np.ascontiguousarray(list(ids), np.int32)
        out = np.zeros(len(a), np.uint8)
        N.call("qmc_cull_flags", self.handle, N.iptr(a), len(a), out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    def trace_records(self, h: int, n_rays: int) -> np.ndarray:
        rec = np.zeros(n_rays, dtype=N.RECORD_DTYPE)
        N.call("qmc_trace_records", self.handle, h, rec.ctypes.data_as(C.POINTER(N.RayRecord)), n_rays)
        return rec

    def trace_records_many(self, ids: Iterable[int], kernel: int = N.KERNEL_AUTO) -> np.ndarray:
        """One trace of distinct heliostats with the chosen kernel (KERNEL_AUTO: the one trace() would
        run; or KERNEL_MAP_DECISIONS into it for records of the map path's own decisions); returns every
        ray's record, heliostat after heliostat in list order, each F*L records indexed by ray id."""
        a = np.ascontiguousarray(list(ids), np.int32)
        L = self.scene.config.num_sunshape_lights_per_group
        n = sum(self.scene.heliostat_info(int(h)).n_micro for h in a) * L
        rec = np.zeros(n, dtype=N.RECORD_DTYPE)
        N.call("qmc_trace_records_many", self.handle, N.iptr(a), len(a), kernel,
               rec.ctypes.data_as(C.POINTER(N.RayRecord)), n)
        return rec

    def trace_stats(self, ids: Iterable[int]) -> dict:
        a = np.ascontiguousarray(list(ids), np.int32)
        st = N.TraceStats()
        N.call("qmc_trace_stats_run", self.handle, N.iptr(a), len(a), C.byref(st))
        return st.as_dict()

    # flux ---------------------------------------------------------------------------------------------
    def flux_shape(self, r: int) -> tuple[int, int]:
        p = N._FP()
        w = C.c_int32()
        h = C.c_int32()
        N.call("qmc_flux_device_ptr", self.handle, r, C.byref(p), C.byref(w), C.byref(h))
        return h.value, w.value

    def flux(self, r: int = 0) -> np.ndarray:
        h, w = self.flux_shape(r)
        out = np.zeros((h, w), np.float32)
        N.call("qmc_flux_read", self.handle, r, N.fptr(out), out.size)
        return out

    def flux_write(self, r: int, img: np.ndarray) -> None:
        a = np.ascontiguousarray(img, np.float32)
        N.call("qmc_flux_write", self.handle, r, N.fptr(a), a.size)

    def flux_device_ptr(self, r: int = 0) -> int:
        p = N._FP()
        N.call("qmc_flux_device_ptr", self.handle, r, C.byref(p), None, None)
        return C.cast(p, C.c_void_p).value

    def flux_clear(self, r: int = 0) -> None:
        N.call("qmc_flux_clear", self.handle, r)

    def smooth(self, r: int, radius: int, ratio: float) -> None:
        N.call("qmc_smooth_receiver", self.handle, r, radius, C.c_float(ratio))

    def smooth_device(self, ptr: int, w: int, h: int, radius: int, ratio: float) -> None:
        N.call("qmc_smooth", self.handle, C.cast(C.c_void_p(ptr), N._FP), w, h, radius, C.c_float(ratio))

    # re-configuration and batched traces --------------------------------------------------------------
    def update(self, scene: Scene) -> None:
        """Refresh sun/aim-dependent device state of a re-processed scene with the same layout."""
        N.call("qmc_update", self.handle, scene.handle)

    def batch_set_aims(self, scene: Scene, offsets: np.ndarray) -> None:
        """offsets: (n_sets, heliostats, 2) aim-point offsets (across, up)."""
        a = np.ascontiguousarray(offsets, np.float32)
        N.call("qmc_batch_set_aims", self.handle, scene.handle, N.fptr(a.reshape(-1)), a.shape[0])

    @staticmethod
    def instances(sun_dirs: np.ndarray, dni: np.ndarray, aim_sets: np.ndarray):
        n = len(dni)
        arr = (N.Instance * n)()
        sd = np.asarray(sun_dirs, np.float32).reshape(n, 3)
        for i in range(n):
            arr[i].sun_dir[:] = [float(x) for x in sd[i]]
            arr[i].dni = float(dni[i])
            arr[i].aim_set = int(aim_sets[i])
        return arr

    def trace_batched(self, grid: int, inst, schedule: dict | None = None, device_maps: int | None = None) -> None:
        sched = None
        if schedule is not None:
            sched = C.byref(N.Schedule(schedule["block_threads"], schedule["workgroups_per_cu"],
                                       schedule["lds_flux_tile"], schedule.get("facet_tiles", 0)))
        ptr = C.cast(C.c_void_p(device_maps), N._FP) if device_maps else None
        N.call("qmc_trace_batched", self.handle, grid, inst, len(inst), sched, ptr)

    def batch_read(self, first: int, count: int, shape: tuple[int, int]) -> np.ndarray:
        out = np.zeros((count,) + tuple(shape), np.float32)
        N.call("qmc_batch_read", self.handle, first, count, N.fptr(out))
        return out

    def batch_set_work_budget(self, nbytes: int) -> None:
        """Per-instance device data a batched trace keeps at once (0: the 2 GiB default); a larger
        sweep is traced in chunks of instances."""
        N.call("qmc_batch_set_work_budget", self.handle, C.c_uint64(nbytes))

    def batch_autotune(self, grid: int, inst) -> dict:
        s = N.Schedule()
        N.call("qmc_batch_autotune", self.handle, grid, inst, len(inst), C.byref(s))
        return s.as_dict()

    # differentiable trace ------------------------------------------------------------------------------
    def _dptr(self, ptr: int):
        return C.cast(C.c_void_p(ptr), N._FP)

    def adjoint_forward(self, grid: int, d_flux: int) -> None:
        """Bilinear-deposit forward into the device buffer d_flux; records the visibility mask."""
        N.call("qmc_adjoint_forward", self.handle, grid, self._dptr(d_flux))

    def adjoint_forward_masked(self, grid: int, receiver: int = 0) -> np.ndarray:
        """Frozen-visibility forward at the current parameters, accumulated in double; returns the
        map of `receiver` (the grid's receiver)."""
        h, w = self.flux_shape(receiver)
        out = np.zeros((h, w), np.float64)
        N.call("qmc_adjoint_forward_masked", self.handle, grid, out.ctypes.data_as(C.POINTER(C.c_double)), out.size)
        return out

    def flux_loss(self, d_flux: int, n: int, target: float, peak: float, lam: float, d_grad: int | None = None):
        loss = C.c_double(0.0)
        N.call("qmc_flux_loss", self.handle, self._dptr(d_flux), n, C.c_float(target), C.c_float(peak),
               C.c_float(lam), self._dptr(d_grad) if d_grad else None, C.byref(loss))
        return loss.value

    def adjoint_backward(self, scene: Scene, grid: int, d_grad: int | None = None):
        nh = scene.counts.n_heliostats
        ga = np.zeros(2 * nh, np.float32)
        gs = np.zeros(3, np.float32)
        N.call("qmc_adjoint_backward", self.handle, scene.handle, grid, self._dptr(d_grad) if d_grad else None,
               N.fptr(ga), N.fptr(gs))
        return ga.reshape(nh, 2), gs

    # timing -------------------------------------------------------------------------------------------
    def event_record(self, slot: int) -> None:
        N.call("qmc_event_record", self.handle, slot)

    def kernel_events(self, before: int, after: int) -> None:
        """Record event slots `before` / `after` around every plain trace kernel from now on (-1, -1: off)."""
        N.call("qmc_kernel_events", self.handle, before, after)

    def event_elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float(0)
        N.call("qmc_event_elapsed_ms", self.handle, a, b, C.byref(ms))
        return ms.value
