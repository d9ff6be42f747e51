"""Scenes re-expressed from the reference's assertion-bearing tests (values, not code), shared by
the CPU KAT tests (oracle and host model) and the GPU facet tests."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from solarenergy_chier_amd import GridSpec, HeliostatSpec, ReceiverSpec, make_config

FIX = Path(__file__).resolve().parent / "golden" / "ref_fixtures"


def rectangle_grid_scene(sub_counts=None):
    """Tests/DataStructureTests/RectangleGridTest.cpp:29-71: grid at (-15,0,20), size (30,10,20),
    10 m cells; five flat heliostats of sizes 1.5, 0.75, 0.75, 1, 1 x (4, 0.1, 3)."""
    cfg = make_config(sun_dir=(0.0, -1.0, 0.0), receiver_pixel_length=0.5, helio_pixel_length=1.0)
    recv = ReceiverSpec(type=0, pos=(0.0, 50.0, 30.0), size=(2.0, 2.0, 0.0), normal=(0.0, 0.0, 1.0), face=0)
    base = np.array([4.0, 0.1, 3.0], np.float32)
    sizes = [base * np.float32(1.5), base * np.float32(0.75), base * np.float32(0.75), base, base]
    poss = [(-8.0, 5.0, 25.0), (10.0, 5.0, 22.5), (10.0, 5.0, 27.5), (-10.0, 5.0, 35.0), (0.0, 5.0, 35.0)]
    hs = []
    for i, (p, s) in enumerate(zip(poss, sizes)):
        hs.append(HeliostatSpec(pos=p, size=tuple(float(x) for x in s), grid=0,
                                sub_count_override=(sub_counts[i] if sub_counts else 0)))
    grid = GridSpec(pos=(-15.0, 0.0, 20.0), size=(30.0, 10.0, 20.0), interval=(10.0, 10.0, 10.0), start_heliostat=0,
                    n_heliostats=5, heliostat_type=0, receiver=0)
    return cfg, [recv], [grid], hs


def rectangle_helio_scene():
    """Tests/DataStructureTests/RectangleHelioTest.cpp:28-35: size (6, 0.1, 5), gap (2, 1), matrix 3x2,
    pixel length 1, normal (0,1,0) at the origin. The normal comes out as (0,1,0) from an aim point
    straight above and a vertical sun."""
    cfg = make_config(sun_dir=(0.0, -1.0, 0.0), receiver_pixel_length=0.5, helio_pixel_length=1.0)
    recv = ReceiverSpec(type=0, pos=(0.0, 30.0, 0.0), size=(2.0, 2.0, 0.0), normal=(0.0, 0.0, 1.0), face=0)
    h = HeliostatSpec(pos=(0.0, 0.0, 0.0), size=(6.0, 0.1, 5.0), grid=0, rows=3, cols=2, gap=(2.0, 1.0))
    grid = GridSpec(pos=(-10.0, -5.0, -10.0), size=(20.0, 10.0, 20.0), interval=(5.0, 5.0, 5.0), start_heliostat=0,
                    n_heliostats=1, heliostat_type=0, receiver=0)
    return cfg, [recv], [grid], [h]


RECT_HELIO_EXPECTED = np.array([(-2.5, 0.05, -2.0), (-1.5, 0.05, -2.0), (1.5, 0.05, -2.0), (2.5, 0.05, -2.0),
                                (-2.5, 0.05, 0.0), (-1.5, 0.05, 0.0), (1.5, 0.05, 0.0), (2.5, 0.05, 0.0),
                                (-2.5, 0.05, 2.0), (-1.5, 0.05, 2.0), (1.5, 0.05, 2.0), (2.5, 0.05, 2.0)], np.float32)


def focus_helio_scene():
    """Tests/DataStructureTests/FocusFlatRectangleHelioTest.cpp:39-57: focus-flat heliostat at
    (10, 2, 6), size (1.7, 0.1, 2.3), gap (0.1, 0.3), matrix 2x3, pixel length 0.5, focal length 1,
    normal (0,1,0)."""
    cfg = make_config(sun_dir=(0.0, -1.0, 0.0), receiver_pixel_length=0.5, helio_pixel_length=0.5)
    recv = ReceiverSpec(type=0, pos=(10.0, 40.0, 6.0), size=(2.0, 2.0, 0.0), normal=(0.0, 0.0, 1.0), face=0)
    h = HeliostatSpec(pos=(10.0, 2.0, 6.0), size=(1.7, 0.1, 2.3), grid=0, rows=2, cols=3, gap=(0.1, 0.3),
                      surface_property=(1.0, -1.0, -1.0, -1.0, -1.0, -1.0))
    grid = GridSpec(pos=(0.0, -5.0, -5.0), size=(20.0, 10.0, 20.0), interval=(5.0, 5.0, 5.0), start_heliostat=0,
                    n_heliostats=1, heliostat_type=1, receiver=0)
    return cfg, [recv], [grid], [h]


FOCUS_LOCAL_NORMALS = None
FOCUS_LOCAL_CENTERS = np.array([(-0.6, 0.245625, -0.65), (0.0, 0.155625, -0.65), (0.6, 0.245625, -0.65),
                                (-0.6, 0.245625, 0.65), (0.0, 0.155625, 0.65), (0.6, 0.245625, 0.65)], np.float64)


def focus_local_normals():
    n = np.array([(0.3, 1.0, 0.325), (0.0, 1.0, 0.325), (-0.3, 1.0, 0.325),
                  (0.3, 1.0, -0.325), (0.0, 1.0, -0.325), (-0.3, 1.0, -0.325)], np.float64)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def l2w(d, n):
    """Textbook local-to-world frame (u = (0,1,0) x n normalised, v = u x n normalised) in float64."""
    d = np.asarray(d, np.float64)
    n = np.asarray(n, np.float64)
    if abs(n[0]) < 1e-5 and abs(n[2]) < 1e-5:
        return d
    u = np.cross([0.0, 1.0, 0.0], n)
    u /= np.linalg.norm(u)
    v = np.cross(u, n)
    v /= np.linalg.norm(v)
    return d[0] * u + d[1] * n + d[2] * v


def focus_expected():
    """Expected sub-heliostat vertices and micro-facet origins of the focus KAT, built the way the
    reference test builds them (FocusFlatRectangleHelioTest.cpp:124-270): facet (0,0) and (0,1)
    transformed explicitly, the rest by the mirror symmetries x -> -x and z -> -z."""
    N_ = focus_local_normals()
    Cc = FOCUS_LOCAL_CENTERS
    loc = np.array([10.0, 2.0, 6.0])

    def tr(p, k):
        return l2w(p, N_[k]) + Cc[k]

    nx = lambda p: p * np.array([-1.0, 1.0, 1.0])
    nz = lambda p: p * np.array([1.0, 1.0, -1.0])
    LB, LU, RU, RB = [np.array(v) for v in ((-0.25, 0, 0.5), (-0.25, 0, -0.5), (0.25, 0, -0.5), (0.25, 0, 0.5))]
    s00 = {k: tr(v, 0) for k, v in zip(("LB", "LU", "RU", "RB"), (LB, LU, RU, RB))}
    s01 = {k: tr(v, 1) for k, v in zip(("LB", "LU", "RU", "RB"), (LB, LU, RU, RB))}
    verts = [s00["LB"], s00["LU"], s00["RU"], s01["LB"], s01["LU"], s01["RU"],
             nx(s00["RB"]), nx(s00["RU"]), nx(s00["LU"]),
             nz(s00["LU"]), nz(s00["LB"]), nz(s00["RB"]), nz(s01["LU"]), nz(s01["LB"]), nz(s01["RB"]),
             nx(nz(s00["RU"])), nx(nz(s00["RB"])), nx(nz(s00["LB"]))]
    verts = np.array(verts) + loc
    up, bot = np.array([0.0, 0.0, -0.25]), np.array([0.0, 0.0, 0.25])
    p00, p10 = tr(up, 0), tr(bot, 0)
    p01, p11 = tr(up, 1), tr(bot, 1)
    origins = np.array([p00, p01, nx(p00), p10, p11, nx(p10), nz(p10), nz(p11), nx(nz(p10)),
                        nz(p00), nz(p01), nx(nz(p00))]) + loc
    normals = np.array([N_[0], N_[1], N_[2], N_[0], N_[1], N_[2], N_[3], N_[4], N_[5], N_[3], N_[4], N_[5]])
    return verts, origins, normals


QMCRT_H0_VERTICES = np.array([(2.0, 1.5, 7.85), (2.0, -1.5, 7.85), (-2.0, -1.5, 7.85)], np.float32)
QMCRT_H2_VERTICES = np.array([(-3.40423, 2.09987, 20.1253), (-3.40541, -1.90012, 20.135),
                              (-6.38378, -1.90012, 19.7754)], np.float32)
QMCRT_H0_ORIGINS = np.array([(1.5, 1.0, 7.85), (0.5, 1.0, 7.85), (-0.5, 1.0, 7.85), (-1.5, 1.0, 7.85),
                             (1.5, 0.0, 7.85), (0.5, 0.0, 7.85), (-0.5, 0.0, 7.85), (-1.5, 0.0, 7.85),
                             (1.5, -1.0, 7.85), (0.5, -1.0, 7.85), (-0.5, -1.0, 7.85), (-1.5, -1.0, 7.85)],
                            np.float32)
