"""GPU parity of the kernel behind the headline rate, and of the trace-call semantics around it.

* The persistent LDS-banded kernel (k_trace_band) writes a record per ray, indexed by ray id, through
  its own slot, wave-mask and band indexing; every record is compared bit for bit with the oracle's
  (map_raytracing, RectangleReceiverRectangleGridRayTracing.cu:6-41) on the PS10-like field, under the
  bench's 50-degree sun and under low suns whose shadow traversals have candidates, and on the config-3
  cylinder field. The same holds for the records of the map path's own decisions (the path qmc_trace
  runs, which decides traversals from the ray's candidate tests and the walk-free proof and re-traces
  the rays these leave open), up to which hit quad a shadowed or blocked record names.
* The full 483M-ray PS10-like map and the full config-3 map are compared with the oracle's
  double-precision maps (relative L1 <= 1e-5: float atomics reorder sums).
* A heliostat listed twice is traced twice (one reference rayTracing call per occurrence), and with
  per-call resampling each trace draws fresh group offsets (QuasiMonteCarloRayTracer.cpp:135-161,
  Heliostat.cu:57-62), checked against the oracle fed each offset set.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from solarenergy_chier_amd import Scene, workloads
from solarenergy_chier_amd import _native as N

pytestmark = pytest.mark.gpu

FLUX_RTOL_L1 = 1e-5
THREADS = max(1, min(16, os.cpu_count() or 1))


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a).sum())


def mismatches(gpu, cpu, map_decisions=False):
    """Rays whose records differ. With the map path's decisions a shadowed or blocked record may name
    another quad than the traversal's first hit (the ray hits both), so its index is not compared."""
    index = gpu["index"] == cpu["index"]
    if map_decisions:
        index |= (cpu["status"] == N.RAY_SHADOWED) | (cpu["status"] == N.RAY_BLOCKED)
    same = ((gpu["status"] == cpu["status"]) & index &
            (gpu["energy"].view(np.uint32) == cpu["energy"].view(np.uint32)) & (gpu["flags"] == cpu["flags"]))
    return int((~same).sum())


def upload(device, wl, seeds=(2, 4, 3)):
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    device.pool_generate(seeds[0], seeds[1])
    device.offsets_generate(seeds[2])
    cfg = scene.config
    L = cfg.num_sunshape_lights_per_group
    dirs, perts = device.pool_get(cfg.num_sunshape_groups * L)
    return scene, O.OracleScene.from_product_scene(scene), dirs, perts, L


def records_vs_oracle(device, scene, orc, dirs, perts, L, ids, kernel):
    """One record trace of `ids` with `kernel`; returns (mismatching rays, rays, deposits, edge rays)."""
    r = scene.grid_desc(scene.heliostat_info(ids[0]).grid).receiver
    device.flux_clear(r)
    rec = device.trace_records_many(ids, kernel)
    flux = device.flux(r)
    at = bad = 0
    cflux = np.zeros(orc.flux_shape(r), np.float64)
    for h in ids:
        F = scene.heliostat_info(h).n_micro
        offs = device.offsets_get(h, F)
        _, crec, _ = orc.trace_heliostat(h, dirs, perts, L, offs, flux=cflux)
        bad += mismatches(rec[at:at + F * L], crec, bool(kernel & N.KERNEL_MAP_DECISIONS))
        at += F * L
    assert at == len(rec)
    dep = int((rec["status"] == N.RAY_DEPOSITED).sum())
    edge = int(((rec["flags"] & N.RAYF_EDGE_CLAMPED) != 0).sum())
    err = rel_l1(flux, cflux)
    print(f"kernel {kernel}: {len(ids)} heliostats, {at} rays, {bad} mismatches, {dep} deposits, "
          f"{edge} edge-clamped, flux rel L1 {err:.2e}", flush=True)
    assert err < FLUX_RTOL_L1
    return bad, at, dep, edge


def test_banded_kernel_records_ps10_sampled(device):
    """Innermost, mid-field and outermost heliostats: the banded kernel and the one-lane-per-ray kernel
    both give every ray the oracle's record."""
    scene, orc, dirs, perts, L = upload(device, workloads.ps10_field())
    ids = [0, 17, 300, 623]
    for kernel in (N.KERNEL_BANDED, N.KERNEL_PER_RAY, N.KERNEL_AUTO):
        bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, kernel)
        assert bad == 0 and dep > 0


def test_banded_kernel_records_ps10_every_eighth(device):
    """Every ray of every eighth heliostat of the bench's field (78 x 774,144 rays) through the banded
    kernel, with per-heliostat, per-wave and per-ray culling and the lookup walks on."""
    scene, orc, dirs, perts, L = upload(device, workloads.ps10_field())
    bad, n, _, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, list(range(0, 624, 8)), N.KERNEL_BANDED)
    assert n == 78 * 48384 * 16
    assert bad == 0


@pytest.mark.parametrize("sun", [(135.0, 15.0), (90.0, 10.0)])
def test_banded_kernel_records_low_sun(device, sun):
    """Low suns: neighbours shade flat heliostats, so the banded kernel's shadow-traversal wave masks
    and lookup walks run with candidates; every ray still equals the oracle's."""
    scene, orc, dirs, perts, L = upload(device, workloads.ps10_field(sun_deg=sun))
    ids = [0, 17, 100, 300, 450, 623]
    assert (device.cull_flags(ids) & 1).sum() >= 2
    bad, _, _, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, N.KERNEL_BANDED)
    assert bad == 0
    shadowed = device.trace_records_many([100], N.KERNEL_BANDED)
    assert (shadowed["status"] == N.RAY_SHADOWED).any()


def test_banded_kernel_records_cylinder_field(device):
    """Config-3 field (focus-flat heliostats, cylinder receiver; full-height windows of 160 of the 512x256
    map's columns, one per workgroup) through the banded kernel, every sixteenth heliostat."""
    scene, orc, dirs, perts, L = upload(device, workloads.cylinder_field())
    bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, list(range(0, 2650, 16)),
                                       N.KERNEL_BANDED)
    assert bad == 0 and dep > 0


@pytest.mark.parametrize("ids", [[1800], [7, 1320, 2649]])
def test_banded_column_windows_few_blocks(device, ids):
    """Column windows with fewer 64-facet blocks than workgroups (141 per config-3 heliostat, 256
    workgroups): with one heliostat most runs are empty and their waves only steal; with three, runs
    hold one or two blocks. Every ray's record equals the oracle's."""
    scene, orc, dirs, perts, L = upload(device, workloads.cylinder_field())
    bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, N.KERNEL_BANDED)
    assert bad == 0 and dep > 0


@pytest.mark.parametrize("L", [16, 32, 64])
def test_banded_facet_shared_order_ragged_blocks(device, L):
    """Flat heliostats with 16, 32 or 64 rays per micro-facet take the banded kernel's facet-shared
    order (each wave computes 64 facets' origins and group offsets once, then traces their rays in L
    steps). Heliostats 12.79 m wide have 255 x 189 facets, not a multiple of 64, so every heliostat ends
    in a partial block; records of both record modes equal the oracle's, under the bench's sun and a low
    sun with shadow candidates."""
    import dataclasses

    for sun in ((180.0, 50.0), (135.0, 15.0)):
        wl = workloads.ps10_field(L=L, groups=1024, n=80, sun_deg=sun)
        wl.heliostats = [dataclasses.replace(h, size=(12.79, 0.1, 9.45)) for h in wl.heliostats]
        scene, orc, dirs, perts, L_ = upload(device, wl)
        assert L_ == L
        F = scene.heliostat_info(0).n_micro
        assert F % 64 != 0
        ids = [0, 9, 40, 79]
        for kernel in (N.KERNEL_BANDED, N.KERNEL_BANDED | N.KERNEL_MAP_DECISIONS):
            bad, n, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, kernel)
            assert n == len(ids) * F * L
            assert bad == 0 and dep > 0


@pytest.mark.parametrize("L", [2, 32])
def test_banded_facet_shared_order_focus_flat(device, L):
    """Focus-flat heliostats take the facet-shared order for any power-of-two L up to 64 (their slots
    are their ray ids): each lane computes its micro-facet's origin and canted-facet index once and the
    rays read them from it. The config-3 heliostats have 9016 micro-facets, not a multiple of 64, so
    every heliostat ends in a partial block; records of both record modes equal the oracle's."""
    wl = workloads.cylinder_field(L=L, groups=max(64, 8192 * 8 // (L * 8)), n=120)
    scene, orc, dirs, perts, L_ = upload(device, wl)
    assert L_ == L
    F = scene.heliostat_info(0).n_micro
    assert F % 64 != 0
    ids = [0, 7, 60, 119]
    for kernel in (N.KERNEL_BANDED, N.KERNEL_BANDED | N.KERNEL_MAP_DECISIONS):
        bad, n, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids, kernel)
        assert n == len(ids) * F * L
        assert bad == 0 and dep > 0


@pytest.mark.parametrize("kernel", [N.KERNEL_BANDED, N.KERNEL_PER_RAY])
def test_map_path_decisions_ps10(device, kernel):
    """The decisions of the path qmc_trace runs, where a traversal the walk-free proof (dda_visit_proof)
    shows to reach a hit quad's cells runs no walk: every ray's status, pixel, energy and flags equal the
    oracle's, on the sampled heliostats, every sixteenth heliostat and under a low sun with shadow
    candidates."""
    for sun, ids in (((180.0, 50.0), [0, 17, 300, 623]), ((180.0, 50.0), list(range(4, 624, 16))),
                     ((135.0, 15.0), [0, 17, 100, 300, 450, 623])):
        scene, orc, dirs, perts, L = upload(device, workloads.ps10_field(sun_deg=sun))
        bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, ids,
                                           kernel | N.KERNEL_MAP_DECISIONS)
        assert bad == 0 and dep > 0


def test_map_path_decisions_cylinder_and_small(device):
    """The same on the config-3 field (focus-flat heliostats, cylinder receiver) and the small field."""
    scene, orc, dirs, perts, L = upload(device, workloads.cylinder_field())
    for kernel in (N.KERNEL_AUTO, N.KERNEL_BANDED):
        bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, list(range(3, 2650, 16)),
                                           kernel | N.KERNEL_MAP_DECISIONS)
        assert bad == 0 and dep > 0
    scene, orc, dirs, perts, L = upload(device, workloads.small_field())
    bad, _, dep, _ = records_vs_oracle(device, scene, orc, dirs, perts, L, list(range(25)),
                                       N.KERNEL_AUTO | N.KERNEL_MAP_DECISIONS)
    assert bad == 0 and dep > 0


def full_map_vs_oracle(device, wl):
    scene, orc, dirs, perts, L = upload(device, wl)
    n = scene.counts.n_heliostats
    device.flux_clear(0)
    device.trace(range(n))
    got = device.flux(0).astype(np.float64)
    offs = np.concatenate([device.offsets_get(h, scene.heliostat_info(h).n_micro) for h in range(n)])
    ref, st = orc.trace_many(list(range(n)), dirs, perts, L, offs, nthreads=THREADS)
    err = rel_l1(got, ref)
    big = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30)
    print(f"{wl.name}: {st['rays']} rays, {st['deposits']} deposits; full map rel L1 {err:.2e}, "
          f"max pixel error {big:.2e} of the peak", flush=True)
    return err, st


def test_ps10_full_map_vs_oracle(device):
    """The bench's workload in full (624 x 48,384 x 16 = 483M rays; the banded kernel) against the
    oracle's double-precision map."""
    err, st = full_map_vs_oracle(device, workloads.ps10_field())
    assert st["rays"] == 624 * 48384 * 16
    assert err < FLUX_RTOL_L1


def test_ps10_low_sun_full_map_vs_oracle(device):
    """The bench's field under its 15-degree sun in full (483M rays): 578 heliostats run shadow
    traversals and a third of the rays are shadowed, through the facet-shared banded kernel's walk-free
    decisions and deferred re-trace; the map against the oracle's double-precision map."""
    err, st = full_map_vs_oracle(device, workloads.ps10_field(sun_deg=(180.0, 15.0)))
    assert st["rays"] == 624 * 48384 * 16
    assert st["unshadowed"] < 0.7 * st["rays"]
    assert err < FLUX_RTOL_L1


def test_config3_full_map_vs_oracle(device):
    """Config 3 in full (2650 focus-flat heliostats, 191M rays, 512x256 cylinder map) against the
    oracle's double-precision map."""
    err, st = full_map_vs_oracle(device, workloads.cylinder_field())
    assert st["deposits"] > 0
    assert err < FLUX_RTOL_L1


def test_repeated_heliostats_are_traced_twice(device):
    """A heliostat listed twice in one call is traced twice (the reference calls rayTracing once per
    listed heliostat). Lists longer than the scene's heliostat count are accepted."""
    scene, _, _, _, _ = upload(device, workloads.small_field())
    device.flux_clear(0)
    device.trace([7])
    one = device.flux(0).astype(np.float64)
    device.flux_clear(0)
    device.trace([7, 7])
    assert rel_l1(device.flux(0), 2 * one) < 1e-6
    scene, _, _, _, _ = upload(device, workloads.ps10_field())
    n = scene.counts.n_heliostats
    device.flux_clear(0)
    device.trace(range(n))
    full = device.flux(0).astype(np.float64)
    device.flux_clear(0)
    device.trace(list(range(n)) * 2)  # 1248 entries for 624 heliostats
    assert rel_l1(device.flux(0), 2 * full) < FLUX_RTOL_L1


def test_per_call_offset_resampling_matches_oracle(device):
    """With resampling on, trace k (counted per pass since it was switched on) first draws the
    offsets of its heliostats with seed + k: two consecutive calls, and one call listing a heliostat
    twice, equal the oracle fed each of those offset sets."""
    scene, orc, dirs, perts, L = upload(device, workloads.small_field())
    cfg = scene.config
    P = cfg.num_sunshape_groups * L
    seed = 100
    cflux = np.zeros(orc.flux_shape(0), np.float64)
    device.flux_clear(0)
    device.offsets_resample(True, seed)
    try:
        device.trace([3])                                   # pass seed + 0
        o1 = device.offsets_get(3, scene.heliostat_info(3).n_micro)
        assert np.array_equal(o1, orc.offsets(seed, P, 3))
        orc.trace_heliostat(3, dirs, perts, L, o1, flux=cflux, records=False)
        device.trace([3])                                   # pass seed + 1
        o2 = device.offsets_get(3, scene.heliostat_info(3).n_micro)
        assert np.array_equal(o2, orc.offsets(seed + 1, P, 3)) and not np.array_equal(o1, o2)
        orc.trace_heliostat(3, dirs, perts, L, o2, flux=cflux, records=False)
        device.trace([3, 5, 3])                             # passes seed + 2 (3 and 5), seed + 3 (3)
        for h, s in ((3, seed + 2), (5, seed + 2), (3, seed + 3)):
            orc.trace_heliostat(h, dirs, perts, L, orc.offsets(s, P, h), flux=cflux, records=False)
        assert np.array_equal(device.offsets_get(5, scene.heliostat_info(5).n_micro), orc.offsets(seed + 2, P, 5))
        err = rel_l1(device.flux(0), cflux)
        print("resampled traces vs oracle rel L1", err, flush=True)
        assert err < FLUX_RTOL_L1
    finally:
        device.offsets_resample(False)
    # off again: the offsets stay as the last trace left them
    before = device.offsets_get(3, scene.heliostat_info(3).n_micro)
    device.trace([3])
    assert np.array_equal(device.offsets_get(3, scene.heliostat_info(3).n_micro), before)


def test_record_trace_rejects_repeats_and_bad_kernels(device):
    scene, _, _, _, _ = upload(device, workloads.small_field(L=8, groups=16))
    with pytest.raises(N.QmcError) as e:
        device.trace_records_many([1, 1])
    assert e.value.status == N.QMC_E_INVALID_ARG
    with pytest.raises(N.QmcError) as e:
        device.trace_records_many([1], kernel=7)
    assert e.value.status == N.QMC_E_INVALID_ARG
