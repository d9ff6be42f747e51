"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from solarenergy_chier_amd import _native as N  # noqa: E402

LIB = ROOT / "oracle" / "liboracle.so"
_lib = None

_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from oracle import build as ob
            ob.build()
        _lib = load(LIB)
    return _lib


def use_library(path):
    """Make the oracle functions of this module call the library at `path` (a build of oracle.cpp, for
    instance one with a planted slip); returns the library in use before, for restoring."""
    global _lib
    before = lib()
    _lib = load(path)
    return before


def restore_library(before):
    global _lib
    _lib = before


def load(path):
    """Load a build of oracle.cpp and declare its signatures."""
    l = C.CDLL(str(path))
    sig = {
        "orc_scene_build": (_P, [C.POINTER(N.Config), C.POINTER(N.ReceiverDesc), C.c_int, C.POINTER(N.GridDesc),
                                 C.c_int, C.POINTER(N.HeliostatDesc), C.c_int, C.c_char_p, C.c_int]),
        "orc_scene_free": (None, [_P]),
        "orc_receiver_info": (C.c_int, [_P, C.c_int, C.POINTER(N.ReceiverInfo)]),
        "orc_heliostat_info": (C.c_int, [_P, C.c_int, C.POINTER(N.HeliostatInfo)]),
        "orc_grid_info": (C.c_int, [_P, C.c_int, C.POINTER(N.GridInfo)]),
        "orc_grid_cells": (C.c_int, [_P, C.c_int, _IP, _IP]),
        "orc_subheliostat_vertices": (C.c_int, [_P, C.c_int, _FP]),
        "orc_factor": (C.c_float, [_P, C.c_int]),
        "orc_facets": (C.c_int, [_P, C.c_int, _FP, _FP]),
        "orc_micro_count": (C.c_int, [_P, C.c_int]),
        "orc_sunshape_table": (None, [C.POINTER(N.Config), _FP, _FP, _FP]),
        "orc_pool_from_uniforms": (None, [C.POINTER(N.Config), _FP, _FP, _FP, _FP, C.c_int, _FP, _FP]),
        "orc_pool_generate": (None, [C.POINTER(N.Config), C.c_uint64, C.c_uint64, _FP, _FP, _FP, _FP, _FP, _FP]),
        "orc_table_theta": (None, [_DP, _DP, _DP, _DP, C.c_int, _DP, C.c_int, _DP]),
        "orc_pool_generate_buie": (None, [C.POINTER(N.Config), C.c_uint64, C.c_uint64, _DP, _DP, _DP, _DP,
                                          C.c_int, _FP, _FP]),
        "orc_offsets_generate": (None, [_P, C.c_uint64, C.c_int, C.c_int, _IP]),
        "orc_trace_heliostat": (C.c_int, [_P, C.c_int, _FP, _FP, C.c_int, C.c_int, _IP, _DP,
                                          C.POINTER(N.RayRecord), C.POINTER(N.TraceStats)]),
        "orc_trace_many": (C.c_int, [_P, _IP, C.c_int, _FP, _FP, C.c_int, C.c_int, _IP, _DP, C.c_int,
                                     C.c_int64, C.POINTER(N.TraceStats)]),
        "orc_smooth": (None, [_FP, C.c_int, C.c_int, C.c_int, C.c_float]),
        "orc_ray_parallelogram": (C.c_int, [_FP, _FP, _FP, _FP, _FP, _FP]),
        "orc_cylinder_intersect": (C.c_int, [_FP, _FP, _FP, _FP, _FP]),
        "orc_cylinder_focus": (None, [_FP, _FP, _FP, _FP]),
        "orc_cylinder_resolution": (C.c_int, [_FP, C.c_int, _IP]),
        "orc_dda": (C.c_int, [_P, C.c_int, _FP, _FP, C.c_int, C.c_int, _IP]),
        "orc_brute_any_hit": (C.c_int, [_P, C.c_int, _FP, _FP, C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    return l


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_FP)


def ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_IP)


class OracleScene:
    """The oracle's own processed scene, built from the same descriptors the product loads."""

    def __init__(self, config, receivers, grids, heliostats):
        self.config = config
        R = (N.ReceiverDesc * max(len(receivers), 1))(*[r if isinstance(r, N.ReceiverDesc) else r.to_c() for r in receivers])
        G = (N.GridDesc * max(len(grids), 1))(*[g if isinstance(g, N.GridDesc) else g.to_c() for g in grids])
        hd = [h if isinstance(h, N.HeliostatDesc) else h.to_c() for h in heliostats]
        self._grid_recv = [int(G[i].receiver) for i in range(len(grids))]
        H = (N.HeliostatDesc * max(len(hd), 1))(*hd)
        err = C.create_string_buffer(256)
        self.handle = lib().orc_scene_build(C.byref(config), R, len(receivers), G, len(grids), H, len(hd), err, 256)
        if not self.handle:
            raise RuntimeError("oracle scene: " + err.value.decode())
        self.n_heliostats = len(hd)

    @classmethod
    def from_workload(cls, wl):
        return cls(wl.config, wl.receivers, wl.grids, wl.heliostats)

    @classmethod
    def from_product_scene(cls, scene):
        cfg, R, nr, G, ng, H, nh = scene.specs()
        return cls(cfg, list(R)[:nr], list(G)[:ng], list(H)[:nh])

    def __del__(self):
        try:
            if self.handle:
                lib().orc_scene_free(self.handle)
        except Exception:
            pass

    def receiver_info(self, r):
        o = N.ReceiverInfo()
        assert lib().orc_receiver_info(self.handle, r, C.byref(o)) == 0
        return o

    def heliostat_info(self, h):
        o = N.HeliostatInfo()
        assert lib().orc_heliostat_info(self.handle, h, C.byref(o)) == 0
        return o

    def grid_info(self, g):
        o = N.GridInfo()
        assert lib().orc_grid_info(self.handle, g, C.byref(o)) == 0
        return o

    def grid_cells(self, g):
        gi = self.grid_info(g)
        cs = np.zeros(gi.nx * gi.ny * gi.nz + 1, np.int32)
        m = np.zeros(max(gi.n_match, 1), np.int32)
        lib().orc_grid_cells(self.handle, g, ip(cs), ip(m))
        return cs, m[: gi.n_match]

    def sub_vertices(self, g):
        n = lib().orc_subheliostat_vertices(self.handle, g, None)
        out = np.zeros((max(n, 1), 3), np.float32)
        lib().orc_subheliostat_vertices(self.handle, g, fp(out))
        return out[:n]

    def factor(self, h):
        return lib().orc_factor(self.handle, h)

    def micro_count(self, h):
        return lib().orc_micro_count(self.handle, h)

    def facets(self, h):
        F = self.micro_count(h)
        o = np.zeros((F, 3), np.float32)
        n = np.zeros((F, 3), np.float32)
        lib().orc_facets(self.handle, h, fp(o), fp(n))
        return o, n

    def offsets(self, seed, P, h):
        out = np.zeros(self.micro_count(h) + 1, np.int32)
        lib().orc_offsets_generate(self.handle, C.c_uint64(seed), P, h, ip(out))
        return out

    def all_offsets(self, seed, P):
        return np.concatenate([self.offsets(seed, P, h) for h in range(self.n_heliostats)]).astype(np.int32)

    def flux_shape(self, r=0):
        ri = self.receiver_info(r)
        return ri.res_y, ri.res_x

    def trace_heliostat(self, h, dirs, perts, L, offsets, flux=None, records=True, stats=True):
        dirs, perts = f32(dirs).reshape(-1, 3), f32(perts).reshape(-1, 3)
        P = dirs.shape[0]
        g = self.heliostat_info(h).grid
        r = self.grid_receiver(g)
        if flux is None:
            flux = np.zeros(self.flux_shape(r), np.float64)
        F = self.micro_count(h)
        rec = np.zeros(F * L, dtype=N.RECORD_DTYPE) if records else None
        st = N.TraceStats()
        off = np.ascontiguousarray(offsets, np.int32)
        rc = lib().orc_trace_heliostat(self.handle, h, fp(dirs), fp(perts), P, L, ip(off),
                                       flux.ctypes.data_as(_DP),
                                       rec.ctypes.data_as(C.POINTER(N.RayRecord)) if rec is not None else None,
                                       C.byref(st) if stats else None)
        assert rc == 0
        return flux, rec, st.as_dict()

    def trace_many(self, ids, dirs, perts, L, offsets_all, nthreads=1, stride=1, flux=None):
        dirs, perts = f32(dirs).reshape(-1, 3), f32(perts).reshape(-1, 3)
        P = dirs.shape[0]
        ids = np.ascontiguousarray(ids, np.int32)
        g = self.heliostat_info(int(ids[0])).grid
        if flux is None:
            flux = np.zeros(self.flux_shape(self.grid_receiver(g)), np.float64)
        st = N.TraceStats()
        off = np.ascontiguousarray(offsets_all, np.int32)
        rc = lib().orc_trace_many(self.handle, ip(ids), len(ids), fp(dirs), fp(perts), P, L, ip(off),
                                  flux.ctypes.data_as(_DP), nthreads, stride, C.byref(st))
        assert rc == 0, rc
        return flux, st.as_dict()

    def grid_receiver(self, g):
        return self._grid_recv[g]

    def dda(self, g, o, d, own_lo, own_hi):
        hit = C.c_int32(-1)
        r = lib().orc_dda(self.handle, g, fp(f32(o)), fp(f32(d)), own_lo, own_hi, C.byref(hit))
        return bool(r), hit.value

    def brute_any_hit(self, g, o, d, own_lo, own_hi):
        return bool(lib().orc_brute_any_hit(self.handle, g, fp(f32(o)), fp(f32(d)), own_lo, own_hi))


def pool_generate(config, seed_sun, seed_pert):
    P = config.num_sunshape_groups * config.num_sunshape_lights_per_group
    d = np.zeros((P, 3), np.float32)
    p = np.zeros((P, 3), np.float32)
    ut, up, gx, gz = (np.zeros(P, np.float32) for _ in range(4))
    lib().orc_pool_generate(C.byref(config), C.c_uint64(seed_sun), C.c_uint64(seed_pert), fp(d), fp(p), fp(ut),
                            fp(up), fp(gx), fp(gz))
    return d, p, (ut, up, gx, gz)


def pool_generate_buie(config, seed_sun, seed_pert, table):
    """Physical Buie pools from the table (theta, cdf, lo, hi) of qmc_sunshape_table_buie."""
    P = config.num_sunshape_groups * config.num_sunshape_lights_per_group
    d = np.zeros((P, 3), np.float32)
    p = np.zeros((P, 3), np.float32)
    th, cdf, lo, hi = (np.ascontiguousarray(a, np.float64) for a in table)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    lib().orc_pool_generate_buie(C.byref(config), C.c_uint64(seed_sun), C.c_uint64(seed_pert), dp(th), dp(cdf),
                                 dp(lo), dp(hi), len(lo), fp(d), fp(p))
    return d, p


def table_theta(table, u):
    th, cdf, lo, hi = (np.ascontiguousarray(a, np.float64) for a in table)
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(len(u))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    lib().orc_table_theta(dp(th), dp(cdf), dp(lo), dp(hi), len(lo), dp(u), len(u), dp(out))
    return out


def pool_from_uniforms(config, ut, up, gx, gz):
    P = len(ut)
    d = np.zeros((P, 3), np.float32)
    p = np.zeros((P, 3), np.float32)
    lib().orc_pool_from_uniforms(C.byref(config), fp(f32(ut)), fp(f32(up)), fp(f32(gx)), fp(f32(gz)), P, fp(d), fp(p))
    return d, p


def sunshape_table(config):
    n = config.inverse_transform_sampling_groups
    cdf = np.zeros(2 * (n + 1), np.float32)
    sl = np.zeros(n, np.float32)
    ab = np.zeros(5, np.float32)
    lib().orc_sunshape_table(C.byref(config), fp(cdf), fp(sl), fp(ab))
    return cdf.reshape(-1, 2), sl, ab


def smooth(img, radius, ratio):
    a = f32(img).copy()
    h, w = a.shape
    lib().orc_smooth(fp(a), w, h, radius, C.c_float(ratio))
    return a


def ray_parallelogram(o, d, A, B, C_):
    tuv = np.zeros(3, np.float32)
    ok = lib().orc_ray_parallelogram(fp(f32(o)), fp(f32(d)), fp(f32(A)), fp(f32(B)), fp(f32(C_)), fp(tuv))
    return bool(ok), tuv


def cylinder_intersect(pos, size, o, d):
    tuv = np.zeros(3, np.float32)
    ok = lib().orc_cylinder_intersect(fp(f32(pos)), fp(f32(size)), fp(f32(o)), fp(f32(d)), fp(tuv))
    return bool(ok), tuv


def cylinder_focus(pos, size, hpos):
    out = np.zeros(3, np.float32)
    lib().orc_cylinder_focus(fp(f32(pos)), fp(f32(size)), fp(f32(hpos)), fp(out))
    return out


def cylinder_resolution(size, ppm):
    res = np.zeros(2, np.int32)
    lib().orc_cylinder_resolution(fp(f32(size)), ppm, ip(res))
    return int(res[0]), int(res[1])


def adjoint_fp64(orc: OracleScene, grid, dirs, perts, L, offsets_all, target, peak, lam, fd_helio=-1, delta=None):
    dirs, perts = f32(dirs).reshape(-1, 3), f32(perts).reshape(-1, 3)
    P = dirs.shape[0]
    l = lib()
    fn = l.orc_adjoint_fp64
    fn.restype = C.c_int
    fn.argtypes = [_P, C.c_int, _FP, _FP, C.c_int, C.c_int, _IP, C.c_float, C.c_float, C.c_float, C.c_int, _DP, _DP,
                   _DP, _DP, _DP, C.POINTER(C.c_int64)]
    ri = orc.receiver_info(orc.grid_receiver(grid))
    flux = np.zeros((ri.res_y, ri.res_x), np.float64)
    loss = C.c_double(0)
    ga = np.zeros(2 * orc.n_heliostats, np.float64)
    gs = np.zeros(3, np.float64)
    nv = C.c_int64(0)
    dl = None
    if delta is not None:
        dl = np.ascontiguousarray(delta, np.float64)
    off = np.ascontiguousarray(offsets_all, np.int32)
    rc = fn(orc.handle, grid, fp(dirs), fp(perts), P, L, ip(off), target, peak, lam, fd_helio,
            dl.ctypes.data_as(_DP) if dl is not None else None, flux.ctypes.data_as(_DP), C.byref(loss),
            ga.ctypes.data_as(_DP), gs.ctypes.data_as(_DP), C.byref(nv))
    assert rc == 0, rc
    return flux, loss.value, ga.reshape(-1, 2), gs, nv.value
