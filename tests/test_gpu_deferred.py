"""GPU parity of the map path's deferred re-trace.

The map path decides a shadow or blocking traversal from the ray's tests of its wave's candidates and the
walk-free proof (dda_visit_proof); a traversal these leave open sends its ray to a queue, and
k_trace_deferred traces the queued rays again with the traversal's own walk (rectGridDDA.cu:31-119).
When a launch defers more rays than its queue holds, k_trace_deferred replays every ray of the launch
and outputs the deferred ones. The fields the bench and the other tests use defer almost nothing, so
this field is built to defer: focus-flat heliostats of 9 x 9 facets packed 7 m apart under a low sun,
whose neighbours put more candidate quads in a heliostat's cones (81 each) than its wave lists hold
(64). Every record of the map path's decisions is compared bit for bit with the oracle's, with the
queue at its own size and capped far below the deferred count, and the flux maps of qmc_trace and
qmc_trace_noalloc with the oracle's double-precision map.
"""
import numpy as np
import pytest

from solarenergy_chier_amd import workloads
from solarenergy_chier_amd import _native as N
from test_gpu_band_parity import FLUX_RTOL_L1, rel_l1, records_vs_oracle, upload

pytestmark = pytest.mark.gpu


def dense_facet_field():
    """5 x 5 focus-flat 6 x 6 m heliostats of 9 x 9 facets, 7 m apart, sun at azimuth 150 and elevation
    20 degrees, 4 x 4 m receiver 30 m up; 324 micro-facets x 4 rays per heliostat (32,400 rays)."""
    cfg = workloads.make_config(sun_dir=workloads.sun_direction(150.0, 20.0), dni=1000.0, csr=0.1,
                                num_sunshape_groups=256, num_sunshape_lights_per_group=4,
                                inverse_transform_sampling_groups=64, receiver_pixel_length=4.0 / 64,
                                helio_disturb_std=0.002, helio_pixel_length=0.25, helio_reflected_rate=0.88)
    recv = workloads.ReceiverSpec(type=0, pos=(0.0, 30.0, 0.0), size=(4.0, 4.0, 0.0), normal=(0.0, 0.0, 1.0),
                                  face=0, res_x=64, res_y=64)
    hs = [workloads.HeliostatSpec(pos=(x, 3.0, z), size=(6.0, 0.1, 6.0), grid=0, rows=9, cols=9, gap=(0.05, 0.05),
                                  surface_property=(-1.0, -1.0, -1.0, -1.0, -1.0, -1.0))
          for x in (-14.0, -7.0, 0.0, 7.0, 14.0) for z in (25.0, 32.0, 39.0, 46.0, 53.0)]
    grid = workloads.GridSpec(pos=(-20.0, 0.0, 18.0), size=(40.0, 8.0, 42.0), interval=(2.0, 2.0, 2.0),
                              start_heliostat=0, n_heliostats=len(hs), heliostat_type=1, receiver=0)
    return workloads.Workload("dense focus-flat field, 25 heliostats of 81 facets", cfg, [recv], [grid], hs, 1, 0.15)


@pytest.mark.parametrize("kernel", [N.KERNEL_PER_RAY, N.KERNEL_BANDED])
def test_deferred_rays_match_oracle(device, kernel):
    """Queue at its own size (65,536 entries for this launch): every deferred ray is traced again from the
    queue; every record equals the oracle's."""
    scene, orc, dirs, perts, L = upload(device, dense_facet_field())
    ids = list(range(25))
    try:
        device.set_defer_capacity(0)
        bad, n, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, kernel | N.KERNEL_MAP_DECISIONS)
        deferred = device.defer_count()
        print(f"kernel {kernel}: {deferred} of {n} rays deferred", flush=True)
        assert n == 25 * 324 * 4
        assert 1000 < deferred <= 65536
        assert bad == 0 and dep > 0
    finally:
        device.set_defer_capacity(0)


@pytest.mark.parametrize("kernel", [N.KERNEL_PER_RAY, N.KERNEL_BANDED])
def test_overflowing_queue_replays_the_launch(device, kernel):
    """Queue capped at 100 entries, far below the deferred count: the re-trace replays every ray of the
    launch and outputs only the deferred ones; every record still equals the oracle's."""
    scene, orc, dirs, perts, L = upload(device, dense_facet_field())
    ids = list(range(25))
    try:
        device.set_defer_capacity(100)
        bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, kernel | N.KERNEL_MAP_DECISIONS)
        assert device.defer_count() > 100
        assert bad == 0 and dep > 0
    finally:
        device.set_defer_capacity(0)


def test_deferred_flux_maps(device):
    """qmc_trace with the queue at its own size and overflowing, and qmc_trace_noalloc, against the
    oracle's double-precision map and against each other."""
    scene, orc, dirs, perts, L = upload(device, dense_facet_field())
    n = scene.counts.n_heliostats
    offs = np.concatenate([device.offsets_get(h, scene.heliostat_info(h).n_micro) for h in range(n)])
    ref, _ = orc.trace_many(list(range(n)), dirs, perts, L, offs, nthreads=4)
    maps = {}
    try:
        for name, cap, noalloc in (("queue", 0, False), ("overflow", 100, False), ("noalloc", 0, True)):
            device.set_defer_capacity(cap)
            device.flux_clear(0)
            if noalloc:
                ids = np.arange(n, dtype=np.int32)
                N.call("qmc_trace_noalloc", device.handle, N.iptr(ids), n)
            else:
                device.trace(range(n))
            maps[name] = device.flux(0).astype(np.float64)
            deferred = device.defer_count()
            err = rel_l1(maps[name], ref)
            print(f"{name}: {deferred} rays deferred, rel L1 {err:.2e}", flush=True)
            assert deferred > 1000
            assert err < FLUX_RTOL_L1
    finally:
        device.set_defer_capacity(0)
    assert rel_l1(maps["overflow"], maps["queue"]) < 1e-6
    assert rel_l1(maps["noalloc"], maps["queue"]) < 1e-6
