"""Synthetic heliostat fields used by the tests, the smoke check and the benchmark.

Coordinates follow the tracer: y is up, x east, z north; ``sun_dir`` is the direction sunlight
travels. A sun at azimuth ``az`` (clockwise from north) and elevation ``el`` gives
``sun_dir = -(cos el sin az, sin el, cos el cos az)``.

* :func:`small_field` — 5x5 flat 2x2 m heliostats at 10 m spacing aimed at a 4x4 m receiver at
  40 m, sun az 180 / el 45, 1600 micro-facets x 64 rays per heliostat (2.56M rays), 128x128 map.
* :func:`ps10_field` — 624 flat 12.84x9.45 m heliostats in a seeded radial-staggered north field
  (radii 100-750 m), 12x12 m receiver at 100 m tilted 12.5 deg, sun el 50, 48,384 micro-facets x
  16 rays per heliostat (483M rays), 256x256 map, 5 m grid cells.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .scene import GridSpec, HeliostatSpec, ReceiverSpec, make_config  # noqa: F401  (re-exported)


def sun_direction(azimuth_deg: float, elevation_deg: float) -> tuple[float, float, float]:
    az, el = math.radians(azimuth_deg), math.radians(elevation_deg)
    v = np.array([-math.cos(el) * math.sin(az), -math.sin(el), -math.cos(el) * math.cos(az)], np.float64)
    v /= np.linalg.norm(v)
    return tuple(float(np.float32(x)) for x in v)


@dataclass
class Workload:
    name: str
    config: object
    receivers: list
    grids: list
    heliostats: list
    smooth_radius: int = 1
    smooth_ratio: float = 0.15

    @property
    def n_heliostats(self) -> int:
        return len(self.heliostats)


def small_field(L: int = 64, groups: int = 1024) -> Workload:
    cfg = make_config(sun_dir=sun_direction(180.0, 45.0), dni=1000.0, csr=0.1, num_sunshape_groups=groups,
                      num_sunshape_lights_per_group=L, inverse_transform_sampling_groups=64,
                      receiver_pixel_length=0.03125, helio_disturb_std=0.002, helio_pixel_length=0.05,
                      helio_reflected_rate=0.88)
    recv = ReceiverSpec(type=0, pos=(0.0, 40.0, 0.0), size=(4.0, 4.0, 0.0), normal=(0.0, 0.0, 1.0), face=0)
    hs = []
    for x in (-20.0, -10.0, 0.0, 10.0, 20.0):
        for z in (30.0, 40.0, 50.0, 60.0, 70.0):
            hs.append(HeliostatSpec(pos=(x, 2.0, z), size=(2.0, 0.1, 2.0), grid=0))
    grid = GridSpec(pos=(-22.0, 0.0, 28.0), size=(44.0, 4.0, 44.0), interval=(2.0, 2.0, 2.0), start_heliostat=0,
                    n_heliostats=len(hs), heliostat_type=0, receiver=0)
    return Workload("small-field 25 heliostats, 2.56M rays, 128x128", cfg, [recv], [grid], hs, 1, 0.15)


def ps10_layout(n: int = 624, r_min: float = 100.0, r_max: float = 750.0, span_deg: float = 40.0,
                seed: int = 1, n_rows: int = 33) -> np.ndarray:
    """Radial-staggered north field: rows every dR metres from r_min, azimuthal pitch chosen so that
    n heliostats fill the rows up to about r_max; odd rows are shifted by half a pitch; positions get
    a seeded +-0.5 m jitter. Returns (n, 2) float64 (x, z)."""
    rng = np.random.default_rng(seed)
    d_r = (r_max - r_min - 10.0) / (n_rows - 1)
    radii = r_min + d_r * np.arange(n_rows)
    span = math.radians(span_deg)
    pitch = 2 * span * radii.sum() / n
    pts = []
    for i, r in enumerate(radii):
        k = max(1, int(2 * span * r / pitch))
        dth = 2 * span / k
        shift = 0.5 * dth if i % 2 else 0.0
        for j in range(k):
            th = -span + (j + 0.5) * dth + shift
            if th > span:
                th -= 2 * span
            pts.append((r * math.sin(th), r * math.cos(th), r))
    pts.sort(key=lambda p: (p[2], p[0]))
    while len(pts) < n:  # pad the outermost row when integer rounding left a deficit
        r = radii[-1] + d_r
        th = -span + (len(pts) % 17) * (2 * span / 17)
        pts.append((r * math.sin(th), r * math.cos(th), r))
    arr = np.array([(p[0], p[1]) for p in pts[:n]], np.float64)
    arr += rng.uniform(-0.5, 0.5, size=arr.shape)
    return arr


def ps10_field(L: int = 16, groups: int = 4096, n: int = 624, seed: int = 1,
               sun_deg: tuple[float, float] = (180.0, 50.0)) -> Workload:
    """PS10-like field; `sun_deg` is the sun's (azimuth, elevation) in degrees."""
    cfg = make_config(sun_dir=sun_direction(*sun_deg), dni=1000.0, csr=0.05, num_sunshape_groups=groups,
                      num_sunshape_lights_per_group=L, inverse_transform_sampling_groups=64,
                      receiver_pixel_length=12.0 / 256.0, helio_disturb_std=0.0015, helio_pixel_length=0.05,
                      helio_reflected_rate=0.88)
    recv = ReceiverSpec(type=0, pos=(0.0, 100.0, 0.0), size=(12.0, 12.0, 0.0), normal=(0.0, 0.0, 1.0), face=0,
                        res_x=256, res_y=256, tilt_deg=12.5)
    xz = ps10_layout(n=n, seed=seed)
    hs = [HeliostatSpec(pos=(float(np.float32(x)), 5.5, float(np.float32(z))), size=(12.84, 0.1, 9.45), grid=0)
          for x, z in xz]
    cell = 5.0
    lo_x = math.floor((xz[:, 0].min() - 10.0) / cell) * cell
    hi_x = math.ceil((xz[:, 0].max() + 10.0) / cell) * cell
    lo_z = math.floor((xz[:, 1].min() - 10.0) / cell) * cell
    hi_z = math.ceil((xz[:, 1].max() + 10.0) / cell) * cell
    grid = GridSpec(pos=(lo_x, 0.0, lo_z), size=(hi_x - lo_x, 15.0, hi_z - lo_z), interval=(cell, cell, cell),
                    start_heliostat=0, n_heliostats=len(hs), heliostat_type=0, receiver=0)
    return Workload("PS10-like 624 heliostats, 483M rays, 256x256", cfg, [recv], [grid], hs, 1, 0.15)


def cylinder_field(L: int = 8, groups: int = 8192, n: int = 2650, seed: int = 2) -> Workload:
    """Config-3 field: 2650 focus-flat heliostats (10.95 x 9.75 m, 4 x 7 canted facets with 0.1 m gaps,
    focal length = slant range) in a seeded radial-staggered surround field from 90 m out to 1.2 km,
    cylinder receiver (radius 4 m, height 10.5 m) centred 140 m up, 512 x 256 map (circumference x
    height), sun at 35 deg elevation, CSR 0.1, 0.1 m micro-facets x 8 rays."""
    cfg = make_config(sun_dir=sun_direction(180.0, 35.0), dni=1000.0, csr=0.1, num_sunshape_groups=groups,
                      num_sunshape_lights_per_group=L, inverse_transform_sampling_groups=64,
                      receiver_pixel_length=2 * math.pi * 4.0 / 512.0, helio_disturb_std=0.0015,
                      helio_pixel_length=0.1, helio_reflected_rate=0.88)
    recv = ReceiverSpec(type=1, pos=(0.0, 140.0, 0.0), size=(4.0, 10.5, 4.0), normal=(0.0, 0.0, 1.0), face=0,
                        res_x=512, res_y=256)
    xz = ps10_layout(n=n, r_min=90.0, r_max=1200.0, span_deg=180.0, seed=seed, n_rows=45)
    hs = [HeliostatSpec(pos=(float(np.float32(x)), 5.5, float(np.float32(z))), size=(10.95, 0.1, 9.75), grid=0,
                        rows=4, cols=7, gap=(0.1, 0.1), surface_property=(-1.0, -1.0, -1.0, -1.0, -1.0, -1.0))
          for x, z in xz]
    cell = 10.0
    lo = -1220.0
    grid = GridSpec(pos=(lo, 0.0, lo), size=(-2 * lo, 15.0, -2 * lo), interval=(cell, 7.5, cell),
                    start_heliostat=0, n_heliostats=len(hs), heliostat_type=1, receiver=0)
    return Workload("config-3 field: 2650 focus-flat heliostats, cylinder receiver, 512x256", cfg, [recv], [grid],
                    hs, 5, 0.05)


def _fmt(x: float) -> str:
    return repr(float(np.float32(x)))


def write_scn(path: str | Path, wl: Workload) -> None:
    """Write the workload in the .scn scene format (Scene/SceneLoader.cpp:243-353). The res/tilt
    receiver extensions have no .scn keyword and are not written."""
    lines = ["# generated scene", "ground 1024.0 1024.0", f"ngrid {len(wl.grids)}", ""]
    written_recv = set()
    for gi, g in enumerate(wl.grids):
        if g.receiver not in written_recv:
            r = wl.receivers[g.receiver]
            lines += [f"Recv {r.type}", "pos " + " ".join(_fmt(v) for v in r.pos),
                      "size " + " ".join(_fmt(v) for v in r.size), "norm " + " ".join(_fmt(v) for v in r.normal),
                      f"face {r.face}", "end", ""]
            written_recv.add(g.receiver)
        lines += [f"Grid {g.type}", "pos " + " ".join(_fmt(v) for v in g.pos),
                  "size " + " ".join(_fmt(v) for v in g.size), "inter " + " ".join(_fmt(v) for v in g.interval),
                  f"n {g.n_heliostats}", f"type {g.heliostat_type}", "end", ""]
        for h in wl.heliostats[g.start_heliostat:g.start_heliostat + g.n_heliostats]:
            lines += [f"gap {_fmt(h.gap[0])} {_fmt(h.gap[1])}", f"matrix {h.rows} {h.cols}",
                      "surface_property " + " ".join(_fmt(v) for v in h.surface_property),
                      "helio " + " ".join(_fmt(v) for v in h.pos), "      " + " ".join(_fmt(v) for v in h.size)]
        lines.append("")
    Path(path).write_text("\n".join(lines) + "\n")


def write_json(path: str | Path, cfg) -> None:
    import json

    d = {
        "sun_dir": [float(v) for v in cfg.sun_dir],
        "dni": float(cfg.dni),
        "csr": float(cfg.csr),
        "num_of_sunshape_groups": int(cfg.num_sunshape_groups),
        "num_per_sunshape_group": int(cfg.num_sunshape_lights_per_group),
        "inverse_transform_sampling_groups": int(cfg.inverse_transform_sampling_groups),
        "receiver_pixel_length": float(cfg.receiver_pixel_length),
        "helio_disturb_std": float(cfg.helio_disturb_std),
        "helio_reflected_rate": float(cfg.helio_reflected_rate),
        "helio_pixel_length": float(cfg.helio_pixel_length),
    }
    Path(path).write_text(json.dumps(d, indent=2))


def sun_sweep(n: int, latitude_deg: float = 37.4, min_elevation_deg: float = 10.0, dni_range=(600.0, 1000.0),
              seed: int = 5) -> tuple[np.ndarray, np.ndarray]:
    """Seeded sun positions over a year of daylight: declination uniform on [-23.44, 23.44] deg, hour
    angle uniform over that day's daylight, rejected below the minimum elevation; DNI uniform on
    dni_range. Returns (sun_dir (n, 3) float32 unit vectors in the light's travel direction, dni (n,))."""
    rng = np.random.default_rng(seed)
    phi = math.radians(latitude_deg)
    dirs, dnis = [], []
    while len(dirs) < n:
        dec = math.radians(rng.uniform(-23.44, 23.44))
        h0 = math.acos(max(-1.0, min(1.0, -math.tan(phi) * math.tan(dec))))
        ha = rng.uniform(-h0, h0)
        east = -math.cos(dec) * math.sin(ha)
        north = math.sin(dec) * math.cos(phi) - math.cos(dec) * math.sin(phi) * math.cos(ha)
        up = math.sin(dec) * math.sin(phi) + math.cos(dec) * math.cos(phi) * math.cos(ha)
        if up <= math.sin(math.radians(min_elevation_deg)):
            continue
        v = -np.array([east, up, north])
        v /= np.linalg.norm(v)
        dirs.append(v.astype(np.float32))
        dnis.append(rng.uniform(*dni_range))
    return np.array(dirs, np.float32), np.array(dnis, np.float32)


def aim_strategies(n_heliostats: int, n_random: int = 3, span: float = 3.0, seed: int = 6) -> np.ndarray:
    """Aim-point offset sets (n_random + 1, heliostats, 2): set 0 aims every heliostat at the receiver
    centre, the others draw each heliostat's (across, up) offset uniformly in [-span, span] metres."""
    rng = np.random.default_rng(seed)
    sets = [np.zeros((n_heliostats, 2), np.float32)]
    for _ in range(n_random):
        sets.append(rng.uniform(-span, span, size=(n_heliostats, 2)).astype(np.float32))
    return np.stack(sets)


def batched_instances(n_suns: int = 4000, n_aims: int = 4, seed: int = 5):
    """The batched workload: n_suns sun positions x n_aims aim strategies, sun-major order.
    Returns (sun_dirs (n, 3), dni (n,), aim_set (n,))."""
    d, e = sun_sweep(n_suns, seed=seed)
    sun = np.repeat(d, n_aims, axis=0)
    dni = np.repeat(e, n_aims)
    aim = np.tile(np.arange(n_aims, dtype=np.int32), n_suns)
    return sun, dni, aim


ADJOINT_TARGET = 7.0e5   # W/m2, flat target flux of the aim-point optimisation
ADJOINT_PEAK = 1.0e6     # W/m2, peak-flux penalty threshold
ADJOINT_LAMBDA = 1.0


def adjoint_small_field(seed: int = 6) -> tuple[Workload, np.ndarray]:
    """Differentiated form on the small field: 25 heliostats x 1600 micro-facets x 50 rays = 2M rays,
    128x128 map, aim offsets drawn uniformly in [-2, 2] m (across, up). Returns (workload, offsets)."""
    wl = small_field(L=50, groups=1024)
    offs = np.random.default_rng(seed).uniform(-2.0, 2.0, size=(len(wl.heliostats), 2)).astype(np.float32)
    for h, o in zip(wl.heliostats, offs):
        h.aim_offset = (float(o[0]), float(o[1]))
    wl.name = "small field, differentiable: 25 heliostats, 2M rays, 128x128"
    return wl, offs


def ps10_sampled_field(sun_deg: tuple[float, float] = (180.0, 50.0)) -> Workload:
    """PS10-like field at instance sampling: 624 heliostats, 0.2 m micro-facets (3008 per heliostat)
    x 1 ray = 1.88M rays, 128x128 map, every heliostat aimed at the receiver centre. The field of
    the batched sun x aim sweep and (with aim offsets) of the differentiable trace. `sun_deg` is the
    sun's (azimuth, elevation) in degrees."""
    wl = ps10_field(L=1, groups=65536, sun_deg=sun_deg)
    wl.config.helio_pixel_length = 0.2
    r = wl.receivers[0]
    r.res_x = r.res_y = 128
    wl.name = "PS10-like field, sampled: 624 heliostats, 1.88M rays, 128x128"
    return wl


def adjoint_ps10_field(seed: int = 6, sun_deg: tuple[float, float] = (180.0, 50.0)) -> tuple[Workload, np.ndarray]:
    """Differentiated form on the PS10-like field: 624 heliostats, 0.2 m micro-facets (3008 per
    heliostat) x 1 ray = 1.88M rays per forward pass, 128x128 map, aim offsets in [-2, 2] m; `sun_deg`
    is the sun's (azimuth, elevation) in degrees."""
    wl = ps10_sampled_field(sun_deg)
    offs = np.random.default_rng(seed).uniform(-2.0, 2.0, size=(len(wl.heliostats), 2)).astype(np.float32)
    for h, o in zip(wl.heliostats, offs):
        h.aim_offset = (float(o[0]), float(o[1]))
    wl.name = "PS10-like field, differentiable: 624 heliostats, 1.88M rays, 128x128"
    return wl, offs
