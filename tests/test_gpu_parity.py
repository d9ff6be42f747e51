"""GPU parity: the HIP path against the CPU oracle on the same scene descriptors and sample arrays.

Bar: per-ray status and hit index (sub-heliostat id or pixel) bit-exact; deposited energies
bit-exact; flux maps within 1e-5 relative L1 (float atomics reorder the sums; the oracle sums in
double)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from solarenergy_chier_amd import Scene, workloads
from solarenergy_chier_amd import _native as N

pytestmark = pytest.mark.gpu

FLUX_RTOL_L1 = 1e-5


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a).sum())


def setup_device(device, wl, seeds=(2, 4, 3)):
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    device.pool_generate(seeds[0], seeds[1])
    device.offsets_generate(seeds[2])
    return scene


def compare_records(gpu, cpu):
    same_status = gpu["status"] == cpu["status"]
    same_index = gpu["index"] == cpu["index"]
    same_energy = gpu["energy"].view(np.uint32) == cpu["energy"].view(np.uint32)
    same_flags = gpu["flags"] == cpu["flags"]
    bad = ~(same_status & same_index & same_energy & same_flags)
    return int(bad.sum())


def test_small_field_per_ray_records(device):
    wl = workloads.small_field()
    scene = setup_device(device, wl)
    orc = O.OracleScene.from_product_scene(scene)
    cfg = scene.config
    L = cfg.num_sunshape_lights_per_group
    P = cfg.num_sunshape_groups * L
    dirs, perts = device.pool_get(P)
    device.flux_clear(0)
    total_bad = 0
    flux_cpu = np.zeros(orc.flux_shape(0), np.float64)
    for h in range(scene.counts.n_heliostats):
        F = scene.heliostat_info(h).n_micro
        offs = device.offsets_get(h, F)
        rec = device.trace_records(h, F * L)
        _, crec, _ = orc.trace_heliostat(h, dirs, perts, L, offs, flux=flux_cpu)
        bad = compare_records(rec, crec)
        print(f"heliostat {h}: {bad} mismatching rays of {F * L}", flush=True)
        total_bad += bad
    assert total_bad == 0
    flux = device.flux(0)
    err = rel_l1(flux, flux_cpu)
    print("flux rel L1", err)
    assert err < FLUX_RTOL_L1


def test_small_field_batched_trace_equals_records(device):
    wl = workloads.small_field()
    scene = setup_device(device, wl)
    n = scene.counts.n_heliostats
    device.flux_clear(0)
    device.trace(range(n))
    a = device.flux(0)
    device.flux_clear(0)
    st = device.trace_stats(range(n))
    b = device.flux(0)
    assert rel_l1(a, b) < FLUX_RTOL_L1
    assert st["rays"] == n * 1600 * 64
    orc = O.OracleScene.from_product_scene(scene)
    cfg = scene.config
    L = cfg.num_sunshape_lights_per_group
    P = cfg.num_sunshape_groups * L
    dirs, perts = device.pool_get(P)
    offs = np.concatenate([device.offsets_get(h, scene.heliostat_info(h).n_micro) for h in range(n)])
    fc, cst = orc.trace_many(list(range(n)), dirs, perts, L, offs, nthreads=8)
    assert st == cst
    assert rel_l1(a, fc) < FLUX_RTOL_L1


def scene_from_files(scn, cfg, L=16, groups=256, helio_pixel_length=0.1):
    """Load a reference scene + configuration, then cap the ray count for the per-ray comparison
    (the reference inputs imply up to 10^8 rays per heliostat): L rays per facet, `groups` sun-shape
    groups and the micro-facet pitch are overridden; everything else is the file's."""
    import kat_scenes as K
    from solarenergy_chier_amd import _native as Nn
    s = Scene()
    s.load_config(str(K.FIX / cfg))
    c = s.config
    c.num_sunshape_lights_per_group = L
    c.num_sunshape_groups = groups
    c.helio_pixel_length = max(c.helio_pixel_length, helio_pixel_length)
    s.set_config(c)
    Nn.call("qmc_scene_load_file", s.handle, str(K.FIX / scn).encode())
    s.process()
    return s


def parity_on_scene(device, scene, heliostats, pools=None, seeds=(2, 4, 3)):
    """Trace each heliostat on the GPU with records and compare to the oracle; returns mismatches."""
    device.upload(scene)
    cfg = scene.config
    L = cfg.num_sunshape_lights_per_group
    P = cfg.num_sunshape_groups * L
    if pools is None:
        device.pool_generate(seeds[0], seeds[1])
    else:
        device.pool_set(pools[0], pools[1], L)
    device.offsets_generate(seeds[2])
    dirs, perts = device.pool_get(P)
    orc = O.OracleScene.from_product_scene(scene)
    bad = 0
    heliostats = list(heliostats)
    flags = dict(zip(heliostats, device.cull_flags(heliostats)))
    for h in heliostats:
        g = scene.heliostat_info(h).grid
        r = scene.grid_desc(g).receiver
        F = scene.heliostat_info(h).n_micro
        offs = device.offsets_get(h, F)
        device.flux_clear(r)
        rec = device.trace_records(h, F * L)
        flux = device.flux(r)
        cflux, crec, _ = orc.trace_heliostat(h, dirs, perts, L, offs)
        b = compare_records(rec, crec)
        print(f"heliostat {h}: {b} mismatches of {F * L}; deposited {(rec['status'] == 3).sum()}; "
              f"cull flags {flags[h]}", flush=True)
        bad += b
        # a traversal the cull kernel skips must be one in which the oracle's DDA finds no hit
        if not flags[h] & 1:
            assert not (crec["status"] == 0).any(), f"heliostat {h}: shadow traversal culled but shadowed rays"
        if not flags[h] & 2:
            assert not (crec["status"] == 1).any(), f"heliostat {h}: blocking traversal culled but blocked rays"
        if cflux.sum() > 0:
            assert rel_l1(flux, cflux) < FLUX_RTOL_L1
        else:
            assert flux.sum() == 0
    return bad


def test_cylinder_reference_scene_parallel_and_sampled(device):
    # Tests/QuasiMonteCarloRayTracingTests/cylinderReceiverIntersectionText.cpp: three grids on one
    # cylinder receiver; heliostat 2 of grid 2 shadows heliostat 1
    scene = scene_from_files("qmcrt_test_scene_cylinder_receiver.scn", "qmcrt_test_configuration2.json")
    cfg = scene.config
    P = cfg.num_sunshape_groups * cfg.num_sunshape_lights_per_group
    par = np.tile([0.0, 1.0, 0.0], (P, 1)).astype(np.float32)
    n = scene.counts.n_heliostats
    assert parity_on_scene(device, scene, range(n), pools=(par, par)) == 0
    assert parity_on_scene(device, scene, range(n)) == 0


def test_rectangle_reference_scene(device):
    scene = scene_from_files("qmcrt_test_scene.scn", "qmcrt_test_configuration.json")
    assert parity_on_scene(device, scene, range(scene.counts.n_heliostats)) == 0


def test_focus_flat_reference_scenes(device):
    """The reference's focus-flat inputs, analytic facets (read method -1) and facets read from the
    file's subhelio blocks (read method 1: PS10_2_typical_4x7, focusflat_example_2x4)."""
    for scn, cfg in [("inputs/PS10_typical/PS10_typical_4x7_scene.scn",
                      "inputs/PS10_typical/PS10_typical_4x7_configuration.json"),
                     ("inputs/FocusFlatRectangleHeliostatInputs/focusflat_example_scene.scn",
                      "inputs/FocusFlatRectangleHeliostatInputs/focusflat_example_configuration.json"),
                     ("inputs/PS10_typical/PS10_2_typical_4x7_scene.scn",
                      "inputs/PS10_typical/PS10_2_typical_configuration.json"),
                     ("inputs/FocusFlatRectangleHeliostatInputs/focusflat_example_2x4_scene.scn",
                      "inputs/FocusFlatRectangleHeliostatInputs/focusflat_example_configuration.json")]:
        scene = scene_from_files(scn, cfg)
        assert parity_on_scene(device, scene, range(scene.counts.n_heliostats)) == 0


def test_two_receiver_scene(device):
    scene = scene_from_files("inputs/CompleteInputFilesExample/twoReceiverScene.scn",
                             "inputs/CompleteInputFilesExample/twoReceiverConfiguration.json")
    n = scene.counts.n_heliostats
    assert parity_on_scene(device, scene, range(min(n, 6))) == 0


def test_ps10_field_sampled_heliostats(device):
    wl = workloads.ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    # innermost, a mid-field and the outermost heliostats (blocking varies with radius)
    assert parity_on_scene(device, scene, [0, 17, 300, 623]) == 0


def test_ps10_field_every_eighth_heliostat(device):
    """Every ray of every eighth heliostat of the bench's PS10-like field (78 heliostats x 774,144 rays)
    equals the oracle's, with all culling on: blocked and unblocked heliostats across the field."""
    wl = workloads.ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    assert parity_on_scene(device, scene, range(0, scene.counts.n_heliostats, 8)) == 0


def test_cylinder_field_every_sixteenth_heliostat(device):
    """Every ray of every sixteenth heliostat of the config-3 field (166 focus-flat heliostats x 72,128
    rays) equals the oracle's, with all culling on."""
    wl = workloads.cylinder_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    assert parity_on_scene(device, scene, range(0, scene.counts.n_heliostats, 16)) == 0


@pytest.mark.parametrize("L", [1, 4])
def test_ps10_field_tiled_ray_order(device, L):
    """PS10-like field with 1 or 4 rays per micro-facet: flat heliostats then take their rays in 8x8 or
    4x4 facet tiles, so each wave's candidate mask is bounded from a tile of facets."""
    wl = workloads.ps10_field(L=L)
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    assert parity_on_scene(device, scene, [0, 17, 300, 623]) == 0


@pytest.mark.parametrize("sun, L", [((135.0, 15.0), 16), ((90.0, 10.0), 16), ((90.0, 10.0), 1)])
def test_ps10_field_low_sun_shadowing(device, sun, L):
    """PS10-like field under a low sun (azimuth, elevation in degrees), where neighbours shade flat
    heliostats: the shadow traversal, its wave masks (per facet row and, at 1 ray per facet, per 8x8
    facet tile) and its lookup walk run with candidates, and every ray still equals the oracle's."""
    wl = workloads.ps10_field(L=L, sun_deg=sun)
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    ids = [0, 17, 100, 300, 450, 623]
    assert parity_on_scene(device, scene, ids) == 0
    assert (device.cull_flags(ids) & 1).sum() >= 2  # the shadow traversal is needed on several of them


def test_wave_masks_spanning_rows_and_facets(device):
    """Wave masks whose rays span several micro-facet rows (flat heliostats, 3 rays per facet: no facet
    tiles, about 43 facets per mask lane of two waves) and several canted facets (focus-flat, 1 ray per
    facet: 128 facets per lane, up to the parallelogram limit and its keep-every-candidate fallback)
    still trace every ray as the oracle does."""
    wl = workloads.ps10_field(L=3)
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    assert parity_on_scene(device, scene, [0, 300, 623]) == 0
    wl = workloads.cylinder_field(L=1)
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    assert parity_on_scene(device, scene, [0, 1000, 2649]) == 0


def test_cull_flags_on_full_fields(device):
    """How much the culling bound removes on the two large fields (printed for the record); that no
    culled traversal has a hit is checked ray by ray against the oracle in parity_on_scene."""
    for wl in (workloads.ps10_field(), workloads.cylinder_field()):
        scene = setup_device(device, wl)
        n = scene.counts.n_heliostats
        f = device.cull_flags(range(n))
        print(f"{wl.name}: {n} heliostats, shadow traversal needed {int((f & 1).sum())}, "
              f"blocking needed {int(((f >> 1) & 1).sum())}", flush=True)
        assert f.max() <= 3
        assert (f & 3).sum() < 2 * n  # the bound removes something on both fields
    with pytest.raises(RuntimeError):
        device.cull_flags([n])


def test_cylinder_field_sampled_heliostats(device):
    """Config-3 field: canted focus-flat facets, 28 blocking quads per heliostat, cylinder receiver."""
    wl = workloads.cylinder_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    assert parity_on_scene(device, scene, [0, 1000, 2649]) == 0


def test_facets_kernel_bit_exact(device):
    import kat_scenes as K
    for specs in (K.rectangle_helio_scene(), K.focus_helio_scene()):
        scene = Scene.from_specs(*specs)
        device.upload(scene)
        orc = O.OracleScene(*specs)
        F = scene.heliostat_info(0).n_micro
        o, n = device.facets(0, F)
        co, cn = orc.facets(0)
        assert np.array_equal(o.view(np.uint32), co.view(np.uint32))
        assert np.array_equal(n.view(np.uint32), cn.view(np.uint32))
    np.testing.assert_allclose(device.facets(0, 12)[0], K.focus_expected()[1], atol=1e-3)
    wl = workloads.ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    orc = O.OracleScene.from_workload(wl)
    for h in (0, 311, 623):
        F = scene.heliostat_info(h).n_micro
        o, n = device.facets(h, F)
        co, cn = orc.facets(h)
        assert np.array_equal(o.view(np.uint32), co.view(np.uint32)), h
        assert np.array_equal(n.view(np.uint32), cn.view(np.uint32)), h


def test_pool_and_offsets_generation_match_oracle(device):
    wl = workloads.small_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    device.pool_generate(2, 4)
    cfg = scene.config
    P = cfg.num_sunshape_groups * cfg.num_sunshape_lights_per_group
    d, p = device.pool_get(P)
    cd, cp, (ut, up, gx, gz) = O.pool_generate(cfg, 2, 4)
    # transcendentals are double-evaluated and rounded once on both sides: expect identity, allow 1 ulp
    for a, b in ((d, cd), (p, cp)):
        ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
        print("pool max ulp", int(ulp.max()), "differing", int((ulp > 0).sum()))
        assert ulp.max() <= 1
    device.pool_from_uniforms(ut, up, gx, gz)
    d2, p2 = device.pool_get(P)
    assert np.abs(d2.view(np.int32).astype(np.int64) - cd.view(np.int32).astype(np.int64)).max() <= 1
    assert np.array_equal(p2.view(np.uint32), cp.view(np.uint32))
    device.offsets_generate(3)
    orc = O.OracleScene.from_workload(wl)
    for h in (0, 24):
        F = scene.heliostat_info(h).n_micro
        assert np.array_equal(device.offsets_get(h, F), orc.offsets(3, P, h))


def test_physical_buie_pool_matches_oracle(device):
    """QMC_SUNSHAPE_BUIE pools on the device equal the oracle's inversion of the same table (the table
    itself is pinned by quadrature on the CPU); an unknown model is rejected."""
    wl = workloads.small_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    device.pool_generate(2, 4, sunshape_model=N.SUNSHAPE_BUIE)
    cfg = scene.config
    P = cfg.num_sunshape_groups * cfg.num_sunshape_lights_per_group
    d, p = device.pool_get(P)
    cd, cp = O.pool_generate_buie(cfg, 2, 4, scene.sunshape_table_buie())
    ulp = np.abs(d.view(np.int32).astype(np.int64) - cd.view(np.int32).astype(np.int64))
    print("buie pool max ulp", int(ulp.max()), "differing", int((ulp > 0).sum()))
    assert ulp.max() <= 1
    assert np.array_equal(p.view(np.uint32), cp.view(np.uint32))
    bad = N.PoolParams(2, 4, 5, 0)
    assert N.status_of("qmc_pool_generate", device.handle, C.byref(bad))[0] == N.QMC_E_INVALID_ARG
    device.pool_generate(2, 4)  # back to the reference sampler for later users of this device


@pytest.mark.parametrize("shape,r,ratio", [((128, 128), 1, 0.15), ((256, 512), 5, 0.05), ((37, 53), 2, 0.1),
                                            ((7, 5), 4, 0.2)])
def test_smoother_bit_exact(device, shape, r, ratio):
    from solarenergy_chier_amd import ReceiverSpec
    wl = workloads.small_field()
    h, w = shape
    rs = ReceiverSpec(type=0, pos=(0.0, 40.0, 0.0), size=(4.0, 4.0, 0.0), res_x=w, res_y=h)
    device.upload(Scene.from_specs(wl.config, [rs], wl.grids, wl.heliostats))
    rng = np.random.default_rng(7)
    img = rng.gamma(2.0, 300.0, size=shape).astype(np.float32)
    img[rng.random(shape) < 0.3] = 0.0
    device.flux_write(0, img)
    device.smooth(0, r, ratio)
    got = device.flux(0)
    want = O.smooth(img, r, ratio)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_ps10_full_size_properties(device):
    """Full 483M-ray field: counts consistent, trace is additive over heliostat subsets and repeatable."""
    wl = workloads.ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    device.pool_generate(2, 4)
    device.offsets_generate(3)
    n = scene.counts.n_heliostats
    device.flux_clear(0)
    st = device.trace_stats(range(n))
    full = device.flux(0).astype(np.float64)
    assert st["rays"] == n * 48384 * 16 and st["facets"] == n * 48384
    assert st["rays"] >= st["unshadowed"] >= st["unblocked"] >= st["deposits"] > 0
    device.flux_clear(0)
    device.trace(range(0, n, 2))
    a = device.flux(0).astype(np.float64)
    device.flux_clear(0)
    device.trace(range(1, n, 2))
    b = device.flux(0).astype(np.float64)
    assert rel_l1(a + b, full) < FLUX_RTOL_L1
    device.flux_clear(0)
    device.trace(range(n))
    device.trace(range(n))
    assert rel_l1(device.flux(0) / 2.0, full) < FLUX_RTOL_L1
    st2 = device.trace_stats(range(n))
    assert st2 == st


def test_reference_example_full_map_vs_oracle(device):
    """The reference's example inputs in full (InputFiles/example: 3 heliostats x 120,000 micro-facets
    x 2048 rays = 737M rays onto 320x200 pixels) against the oracle's double-precision map. Here a
    pixel receives tens of thousands of nearly equal deposits; one float accumulator per pixel would
    round each against a sum thousands of times larger, a bias near 3e-4. The trace sums per-workgroup
    LDS bands and folds them, which keeps the map within the 1e-5 bar."""
    import os
    import kat_scenes as K
    from solarenergy_chier_amd import _native as Nn
    ex = K.FIX / "inputs" / "example"
    s = Scene()
    s.load_config(str(ex / "example_configuration.json"))
    Nn.call("qmc_scene_load_file", s.handle, str(ex / "example_scene.scn").encode())
    s.process()
    device.upload(s)
    device.pool_generate(2, 4)
    device.offsets_generate(3)
    cfg = s.config
    L = cfg.num_sunshape_lights_per_group
    P = cfg.num_sunshape_groups * L
    n = s.counts.n_heliostats
    device.flux_clear(0)
    device.trace(range(n))
    got = device.flux(0).astype(np.float64)
    dirs, perts = device.pool_get(P)
    offs = np.concatenate([device.offsets_get(h, s.heliostat_info(h).n_micro) for h in range(n)])
    orc = O.OracleScene.from_product_scene(s)
    print("oracle: tracing", n * s.heliostat_info(0).n_micro * L, "rays", flush=True)
    ref, _ = orc.trace_many(list(range(n)), dirs, perts, L, offs, nthreads=min(16, os.cpu_count() or 1))
    err = rel_l1(got, ref)
    print("example full map rel L1 vs oracle", err, flush=True)
    assert err < FLUX_RTOL_L1


def test_cylinder_field_full_trace_equals_counting_pass(device):
    """Config-3 field in full (191M rays, cylinder receiver): the map of the plain trace (per-heliostat,
    per-wave and per-ray culling, lookup walks) equals the counting pass's map (never culled, every
    traversal run with its own tests), up to float summation order."""
    wl = workloads.cylinder_field()
    scene = setup_device(device, wl)
    n = scene.counts.n_heliostats
    device.flux_clear(0)
    st = device.trace_stats(range(n))
    ref = device.flux(0).astype(np.float64)
    device.flux_clear(0)
    device.trace(range(n))
    got = device.flux(0).astype(np.float64)
    err = rel_l1(got, ref)
    print("config-3 full map rel L1", err, "deposits", st["deposits"], flush=True)
    assert err < FLUX_RTOL_L1 and got.sum() > 0


def test_atomic_rate_probe(device):
    """The capability probe returns positive rates, and LDS atomics outrun device-scope ones (the
    reason the trace accumulates its deposits in LDS bands)."""
    r = device.probe_atomic_rates()
    print("atomic rates (G ops/s)", r, flush=True)
    assert all(v > 0 for v in r.values())
    assert r["lds_f32"] > r["global_f32"]


def test_fast_sqrt_sequences_exhaustive(device):
    """The tracer's normalisations take a short square-root and reciprocal sequence for arguments in
    [2^-8, 2^8] (no exponent scaling or special-value steps). For every one of the 134,217,729 floats
    in that range both results equal the general correctly rounded sqrtf and 1.0f / sqrtf bit for bit,
    and those equal the double-precision computation rounded to float (an independent reference: a
    float square root or quotient rounded through double is exact)."""
    n, bad = device.check_fast_sqrt()
    print(f"fast sqrt check: {n} arguments, mismatches {bad}", flush=True)
    assert n == (0x43800000 - 0x3B800000) + 1
    assert bad == [0, 0, 0]


def test_kernel_events_bracket_the_trace_kernel(device):
    """qmc_kernel_events: the library's events around the trace kernel of a plain trace measure a
    positive time within the whole call's (which adds the culling kernels and folds); bad slots are
    refused and (-1, -1) turns the events off."""
    wl = workloads.ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    device.pool_generate(2, 4)
    device.offsets_generate(3)
    ids = list(range(scene.counts.n_heliostats))
    device.kernel_events(2, 3)
    try:
        for _ in range(2):
            device.flux_clear(0)
            device.event_record(0)
            device.trace(ids)
            device.event_record(1)
            k, total = device.event_elapsed_ms(2, 3), device.event_elapsed_ms(0, 1)
            print(f"trace kernel {k:.3f} ms of a {total:.3f} ms call", flush=True)
            assert 0.0 < k <= total
    finally:
        device.kernel_events(-1, -1)
    for bad in ((2, 2), (-1, 3), (0, 16)):
        with pytest.raises(N.QmcError):
            device.kernel_events(*bad)


def test_offsets_set_validates_range(device):
    """Group offsets index the sample pool, so values outside [0, P] are refused before any launch."""
    wl = workloads.small_field(L=8, groups=16)
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    device.upload(scene)
    F = scene.heliostat_info(0).n_micro
    with pytest.raises(N.QmcError) as e:
        device.offsets_set(0, np.zeros(F + 1, np.int32))
    assert e.value.status == N.QMC_E_STATE
    device.pool_generate(2, 4)
    P = 16 * 8
    good = np.full(F + 1, P, np.int32)
    device.offsets_set(0, good)
    assert np.array_equal(device.offsets_get(0, F), good)
    for bad_value in (-1, P + 1):
        bad = good.copy()
        bad[F // 2] = bad_value
        with pytest.raises(N.QmcError) as e:
            device.offsets_set(0, bad)
        assert e.value.status == N.QMC_E_RANGE
