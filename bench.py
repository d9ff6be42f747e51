"""Benchmark: rays/s of the plain tracer on one GPU, PS10-like field (624 heliostats, 483M rays per
step, 256x256 flux map), with the trace kernel's vector-issue roofline and the CPU oracle timed on a
bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ps10|small] [--no-cpu-baseline]

A step = clear the flux map + trace every heliostat of the field in one trace call: fresh micro-facet
group offsets for every heliostat (as the reference draws them on each rayTracing call), the culling
kernels, the trace kernel and its folds, with the scene and sample pools resident in HBM. Prints ONE
JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# MI355X vector issue: 256 CUs x 4 SIMDs, a wave64 VALU instruction occupies a 16-lane SIMD for 4
# cycles, 2.4 GHz peak engine clock -> 614.4 G wave instructions per second
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(st: dict, pool_entries: int = 0) -> int:
    """Bytes the reference control flow reads/writes with no on-chip reuse (per trace launch):
    28 B per ray (offset 4 + pool sample 12 + facet origin 12), 28 B per unshadowed ray (facet normal
    12 + next offset 4 + slope sample 12), 8 B per deposit (atomic read-modify-write), 8 B per DDA
    cell (two cell-list offsets), 4 B per cell-list entry read, 36 B per non-own quad tested."""
    return (28 * st["rays"] + 28 * st["unshadowed"] + 8 * st["deposits"] + 8 * st["cells"]
            + 4 * st["candidates"] + 36 * st["tests"])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ps10", choices=["ps10", "small"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched sun x aim measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    ap.add_argument("--no-cylinder", action="store_true", help="skip the config-3 (cylinder) field measurement")
    ap.add_argument("--no-low-sun", action="store_true", help="skip the PS10 point under a 15-degree sun")
    ap.add_argument("--batched-suns", type=int, default=4000)
    ap.add_argument("--loop-instances", type=int, default=200)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (calibrates HBM byte counters), then exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build runs on a single MI355X; --gpus {args.gpus} is not "
              "supported (the field would shard by heliostat ranges with one flux all-reduce).", file=sys.stderr)
        return 2

    import numpy as np

    from solarenergy_chier_amd import Device, Scene, workloads

    wl = workloads.ps10_field() if args.workload == "ps10" else workloads.small_field()
    log(f"workload: {wl.name}")
    t0 = time.time()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev = Device(0)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    dev.sync()
    n = scene.counts.n_heliostats
    ids = list(range(n))
    cfg = scene.config
    L = cfg.num_sunshape_lights_per_group
    P = cfg.num_sunshape_groups * L
    F = [scene.heliostat_info(h).n_micro for h in ids]
    rays_per_step = sum(F) * L
    log(f"scene ready in {time.time() - t0:.1f}s: {n} heliostats, {rays_per_step} rays per step")

    # algorithmic work of one launch (instrumented kernel, untimed)
    st = dev.trace_stats(ids)
    alg_bytes = algorithmic_bytes(st)
    log(f"counted work: {st}; algorithmic bytes per launch {alg_bytes / 1e9:.2f} GB")

    def timed(steps: int) -> tuple[float, float, float]:
        """K steps between events on the trace stream, and the library's events around each launch's
        trace kernel alone (the offsets, culling kernels and folds of a call are in the step time, not
        in the kernel time): (ms per step, trace-kernel ms per step, host wall seconds)."""
        kernel_ms = 0.0
        dev.sync()
        dev.kernel_events(2, 3)
        wall0 = time.perf_counter()
        dev.event_record(0)
        for _ in range(steps):
            dev.flux_clear(0)
            dev.trace(ids)
            kernel_ms += dev.event_elapsed_ms(2, 3)
        dev.event_record(1)
        dev.sync()
        dev.kernel_events(-1, -1)
        return dev.event_elapsed_ms(0, 1) / steps, kernel_ms / steps, time.perf_counter() - wall0

    # the reference's semantics: every trace call draws fresh group offsets for its heliostats
    dev.offsets_resample(True, 1000)
    for _ in range(args.warmup):
        dev.flux_clear(0)
        dev.trace(ids)
    dev.sync()
    log(f"warmup done ({args.warmup} steps)")
    ms_per_step, kernel_avg_ms, wall = timed(args.steps)
    rays_per_s = rays_per_step / (ms_per_step / 1e3)
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step (wall {wall:.3f}s), trace kernel {kernel_avg_ms:.3f} ms, "
        f"{rays_per_s / 1e9:.3f} G rays/s")
    flux = dev.flux(0)
    log(f"flux map sum {float(flux.sum(dtype=np.float64)):.6e}, max {float(flux.max()):.3f}")
    # the same steps with the offsets kept (no resampling kernel in the step)
    dev.offsets_resample(False)
    ms_fixed, kernel_fixed_ms, _ = timed(args.steps)
    log(f"fixed offsets: {ms_fixed:.3f} ms/step, trace kernel {kernel_fixed_ms:.3f} ms")

    roofline = valu_roofline(args.workload, kernel_avg_ms, alg_bytes)
    # deposit rates of the trace against the device's measured atomic ceilings (capability probe)
    ceil = dev.probe_atomic_rates()
    log(f"atomic ceilings (G ops/s): {ceil}")
    roofline["atomics"] = atomic_rates(args.workload, kernel_avg_ms, ceil)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, dev, ids, L, P, F, args.cpu_seconds)
    batched = batched_ps10 = None
    if not args.no_batched:
        # the sweep's own field is the 25-heliostat config-1 field (2.56M launched rays per instance);
        # the 624-heliostat field at instance sampling is measured beside it
        batched = bench_batched(dev, "small", args.batched_suns, args.loop_instances)
        batched_ps10 = bench_batched(dev, "ps10", args.batched_suns, args.loop_instances)
    adjoint = None
    if not args.no_adjoint:
        adjoint = bench_adjoint(dev)
    cylinder = None
    if not args.no_cylinder:
        cylinder = bench_cylinder(dev)
    low_sun = None
    if args.workload == "ps10" and not args.no_low_sun:
        low_sun = bench_low_sun(dev)

    line = {
        "metric": "rays/s on 1 GPU (PS10-like); batched scenes/s; adjoint:forward cost",
        "value": rays_per_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded field layout, Sobol sun-shape pool, counter-hash slope errors)",
        "config": {
            "workload": wl.name,
            "heliostats": n,
            "rays_per_step": rays_per_step,
            "micro_facets_per_heliostat": F[0],
            "rays_per_facet": L,
            "pool": P,
            "flux_map": list(flux.shape),
        },
        "roofline": roofline,
        "fixed_offsets": {"ms_per_step": ms_fixed, "kernel_avg_ms": kernel_fixed_ms,
                          "rays_per_s": rays_per_step / (ms_fixed / 1e3),
                          "note": "the same step without the per-call resampling of the group offsets"},
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_ps10_sampled": batched_ps10,
        "adjoint": adjoint,
        "config3_cylinder_field": cylinder,
        "ps10_low_sun": low_sun,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


def profile_json(name: str):
    """A committed rocprofv3 summary (profiles/<name>.json) and whether it was measured on the
    sources this library is built from."""
    from solarenergy_chier_amd import build as b

    f = ROOT / "profiles" / f"{name}.json"
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    here = b.source_digest()
    tag = (f"profiles/{name}.json: rocprofv3 run recorded there (sources {d.get('source_digest')}; "
           + ("the sources of this build)" if d.get("source_digest") == here else f"this build's sources are {here}: stale)"))
    return d, tag


def valu_roofline(workload: str, kernel_ms: float, alg_bytes: int) -> dict:
    """Roofline of the trace kernel against vector issue, the resource that binds it: achieved = its
    VALU wave instructions per launch (rocprofv3 SQ_INSTS_VALU, a count fixed by the code and the
    workload) / this run's kernel time, peak = 614.4 G/s. HBM and the reference-equivalent
    algorithmic-byte rate are reported beside it, labelled."""
    sec = kernel_ms / 1e3
    out = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINST, "unit": "G VALU wave64 instructions/s",
           "frac": None, "traffic": None,
           "kernel": "k_trace_band<0,1024> (persistent, LDS band of map rows)" if workload == "ps10" else "trace kernel",
           "kernel_avg_ms": kernel_ms, "kernel_time_source": "HIP events on the trace stream around each launch",
           "algorithmic_bytes_per_launch": alg_bytes,
           "algorithmic_gbs_reference_equivalent": alg_bytes / sec / 1e9,
           "algorithmic_gbs_note": "bytes the reference's control flow would move with no on-chip reuse and no "
                                   "culling, over this kernel's time; not a roofline (culling skips that work)",
           "provenance": {}}
    v, vtag = profile_json(f"trace_{workload}_valu")
    if v and "valu_insts_per_launch" in v:
        out["valu_insts_per_launch"] = v["valu_insts_per_launch"]
        out["valu_insts_per_64_rays"] = v.get("valu_insts_per_64_rays")
        out["salu_insts_per_64_rays"] = v.get("salu_insts_per_64_rays")
        out["achieved"] = v["valu_insts_per_launch"] / sec / 1e9
        out["frac"] = out["achieved"] / VALU_PEAK_GINST
        out["valu_busy_profiled"] = v.get("valu_busy")
        out["provenance"]["valu_insts_per_launch"] = vtag
    t, ttag = profile_json(f"trace_{workload}_hbm_traffic")
    if t and "bytes_per_launch" in t:
        out["traffic"] = t["bytes_per_launch"]
        out["hbm_gbs"] = t["bytes_per_launch"] / sec / 1e9
        out["hbm_frac"] = out["hbm_gbs"] / HBM_PEAK_GBS
        out["provenance"]["traffic"] = ttag
        out["provenance"]["hbm_gbs"] = "profiled bytes per launch / this run's kernel time"
    return out


def atomic_rates(workload: str, kernel_ms: float, ceil: dict) -> dict:
    """The trace kernel's deposit rates (profiled atomic counts / this run's kernel time) against the
    device's atomic ceilings measured in this run (qmc_probe_atomic_rates)."""
    out = {"ceiling_gops": ceil, "ceiling_source": "qmc_probe_atomic_rates in this run"}
    a, tag = profile_json(f"trace_{workload}_atomics")
    if a and "lds_atomic_lanes_per_launch" in a:
        sec = kernel_ms / 1e3
        lds_rate = a["lds_atomic_lanes_per_launch"] / sec / 1e9
        glob_rate = a["global_atomics_per_launch"] / sec / 1e9
        out.update({"lds_atomic_lanes_per_launch": a["lds_atomic_lanes_per_launch"],
                    "global_atomics_per_launch": a["global_atomics_per_launch"],
                    "lds_gops": lds_rate, "lds_frac": lds_rate / ceil["lds_f32"],
                    "global_gops": glob_rate, "global_frac": glob_rate / ceil["global_f64"],
                    "provenance": tag})
    return out


def bench_low_sun(dev, steps: int = 3) -> dict:
    """The PS10-like field under a 15-degree sun (azimuth 180): neighbours shade and block each other,
    so shadow traversals run (the 50-degree headline sun culls every one of them)."""
    from solarenergy_chier_amd import Scene, workloads

    sun = (180.0, 15.0)
    wl = workloads.ps10_field(sun_deg=sun)
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    ids = list(range(scene.counts.n_heliostats))
    rays = sum(scene.heliostat_info(h).n_micro for h in ids) * scene.config.num_sunshape_lights_per_group
    st = dev.trace_stats(ids)
    flags = dev.cull_flags(ids)
    dev.offsets_resample(True, 2000)
    dev.flux_clear(0)
    dev.trace(ids)
    dev.sync()
    total = kern = 0.0
    dev.kernel_events(2, 3)
    for _ in range(steps):
        dev.flux_clear(0)
        dev.event_record(12)
        dev.trace(ids)
        dev.event_record(13)
        dev.sync()
        total += dev.event_elapsed_ms(12, 13)
        kern += dev.event_elapsed_ms(2, 3)
    dev.kernel_events(-1, -1)
    dev.offsets_resample(False)
    ms, kms = total / steps, kern / steps
    out = {"workload": f"{wl.name}, sun azimuth {sun[0]:.0f} deg, elevation {sun[1]:.0f} deg", "rays": rays,
           "ms_per_step": ms, "kernel_avg_ms": kms, "rays_per_s": rays / (ms / 1e3),
           "shadowed_fraction": 1.0 - st["unshadowed"] / st["rays"],
           "blocked_fraction": (st["unshadowed"] - st["unblocked"]) / st["rays"],
           "heliostats_needing_shadow_traversal": int((flags & 1).sum()),
           "heliostats_needing_blocking_traversal": int(((flags >> 1) & 1).sum()),
           "algorithmic_bytes_per_launch": algorithmic_bytes(st), "counted_work": st}
    log(f"PS10 low sun {sun}: {ms:.3f} ms/step (kernel {kms:.3f} ms) = {rays / ms / 1e6:.2f} G rays/s; "
        f"shadowed {out['shadowed_fraction']:.3f}, blocked {out['blocked_fraction']:.3f}, shadow traversals on "
        f"{out['heliostats_needing_shadow_traversal']} heliostats")
    return out


def cpu_baseline(scene, dev, ids, L, P, F, seconds):
    """The oracle (CPU restatement) on a strided sample of the same rays: every host thread of the
    process's CPU share (the GPU box gives one GPU's process 16, whatever os.cpu_count() reports), and
    one thread on a smaller sample. The stride is set from one calibration run (every 64th ray) so the
    all-thread sample takes about `seconds`."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the checker / CPU baseline only

    host_cpus = os.cpu_count() or 1
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = host_cpus
    threads = max(1, min(16, share))
    orc = O.OracleScene.from_product_scene(scene)
    dirs, perts = dev.pool_get(P)
    offs = np.concatenate([dev.offsets_get(h, F[h]) for h in ids])
    total = sum(F) * L

    def run(stride, nthreads):
        t = time.perf_counter()
        _, st = orc.trace_many(ids, dirs, perts, L, offs, nthreads=nthreads, stride=stride)
        dt = time.perf_counter() - t
        return st["rays"], dt

    r0, t0 = run(64, threads)
    rate = r0 / max(t0, 1e-6)
    stride = max(1, int(total / max(rate * seconds, 1.0)))
    log(f"cpu baseline: calibration {rate / 1e6:.2f} M rays/s on {threads} threads (every 64th ray: {r0} rays "
        f"in {t0:.2f}s); sample stride {stride}")
    rays, dt = run(stride, threads)
    value = rays / dt
    log(f"cpu baseline: {rays} rays in {dt:.2f}s = {value / 1e6:.3f} M rays/s on {threads} threads")
    stride1 = max(stride, int(total / max(value / threads * 2.0, 1.0)))
    rays1, dt1 = run(stride1, 1)
    log(f"cpu baseline: one thread {rays1} rays in {dt1:.2f}s = {rays1 / dt1 / 1e6:.3f} M rays/s")
    return {"value": value, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}-th ray of all {len(ids)} heliostats ({rays} rays, {dt:.1f} s)",
            "host_cpus": host_cpus, "cpu_share": share,
            "single_thread": {"value": rays1 / dt1, "sample": f"every {stride1}-th ray ({rays1} rays, {dt1:.1f} s)"}}


def bench_batched(dev, field: str, n_suns: int, n_loop: int) -> dict:
    """One field over n_suns sun positions x 4 aim strategies in one persistent launch (schedule from
    the autotuner), against looping the plain call on the first n_loop instances: re-process the scene
    for the instance's sun/DNI/aims on the host, refresh the device state, clear, trace, read the map
    back. The loop's trace-kernel time alone (HIP events) is reported beside it."""
    import numpy as np

    from solarenergy_chier_amd import Scene, workloads

    wl = workloads.ps10_sampled_field() if field == "ps10" else workloads.small_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    nh = scene.counts.n_heliostats
    aims = workloads.aim_strategies(nh)
    dev.batch_set_aims(scene, aims)
    sun, dni, aim = workloads.batched_instances(n_suns=n_suns)
    inst = dev.instances(sun, dni, aim)
    n = len(inst)
    shape = dev.flux_shape(0)
    rays = sum(scene.heliostat_info(h).n_micro for h in range(nh)) * scene.config.num_sunshape_lights_per_group
    t = time.perf_counter()
    sched = dev.batch_autotune(0, inst)
    tune_s = time.perf_counter() - t
    log(f"batched {field}: {n} instances x {rays} rays, autotuned schedule {sched} in {tune_s:.2f}s")
    dev.trace_batched(0, inst, sched)  # warm-up
    dev.sync()
    dev.event_record(4)
    dev.trace_batched(0, inst, sched)
    dev.event_record(5)
    dev.sync()
    ms = dev.event_elapsed_ms(4, 5)
    batched_sps = n / (ms / 1e3)
    maps = dev.batch_read(0, min(n, 8), shape)
    log(f"batched {field}: {ms:.1f} ms for {n} instances = {batched_sps:.0f} scenes/s "
        f"({n * rays / (ms / 1e3) / 1e9:.2f} G rays/s)")
    # looped plain calls, timed on the host clock (host re-processing is part of the plain path)
    cfg = scene.config
    k = min(n_loop, n)
    dev.sync()
    kernel_ms = 0.0
    t = time.perf_counter()
    for i in range(k):
        cfg.sun_dir[:] = [float(x) for x in sun[i]]
        cfg.dni = float(dni[i])
        scene.set_config(cfg)
        scene.set_aim_offsets(aims[aim[i]])
        scene.process()
        dev.update(scene)
        dev.flux_clear(0)
        dev.event_record(6)
        dev.trace(range(nh))
        dev.event_record(7)
        m = dev.flux(0)
        kernel_ms += dev.event_elapsed_ms(6, 7)
        if i < len(maps):
            d = np.abs(m.astype(np.float64) - maps[i]).sum() / max(np.abs(m).sum(), 1e-30)
            if d > 1e-5:
                raise AssertionError(f"batched instance {i} differs from the plain call: rel L1 {d:.2e}")
    loop_s = time.perf_counter() - t
    loop_sps = k / loop_s
    kernel_sps = k / (kernel_ms / 1e3)
    log(f"plain loop {field}: {k} instances in {loop_s:.3f}s = {loop_sps:.0f} scenes/s (trace kernels alone "
        f"{kernel_sps:.0f} scenes/s); speedup {batched_sps / loop_sps:.2f}x over the loop, "
        f"{batched_sps / kernel_sps:.2f}x over its kernels")
    return {"metric": "scenes/s", "workload": f"{wl.name} x {n_suns} suns x 4 aims", "instances": n,
            "rays_per_instance": rays, "batched_ms": ms, "batched_scenes_per_s": batched_sps,
            "plain_loop_scenes_per_s": loop_sps, "plain_loop_instances": k,
            "speedup_vs_plain_loop": batched_sps / loop_sps,
            "plain_kernel_only_scenes_per_s": kernel_sps, "speedup_vs_plain_kernels": batched_sps / kernel_sps,
            "schedule": sched, "autotune_s": tune_s}


def bench_cylinder(dev, steps: int = 3) -> dict:
    """Config-3 field: 2650 focus-flat heliostats x 9016 micro-facets x 8 rays = 191M rays per launch
    onto a 512x256 cylinder map, then the pipeline's trimmed-mean smoothing (radius 5, ratio 0.05)."""
    from solarenergy_chier_amd import Scene, workloads

    wl = workloads.cylinder_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    ids = list(range(scene.counts.n_heliostats))
    rays = sum(scene.heliostat_info(h).n_micro for h in ids) * scene.config.num_sunshape_lights_per_group
    st = dev.trace_stats(ids)
    dev.flux_clear(0)
    dev.trace(ids)
    dev.sync()
    trace_ms = 0.0
    for _ in range(steps):
        dev.flux_clear(0)
        dev.event_record(12)
        dev.trace(ids)
        dev.event_record(13)
        dev.sync()
        trace_ms += dev.event_elapsed_ms(12, 13)
    trace_ms /= steps
    # smoothing of the traced map: one untimed call (first launch of the kernel), then the average
    # of five, each smoothing the previous result in place
    dev.smooth(0, wl.smooth_radius, wl.smooth_ratio)
    dev.sync()
    dev.event_record(12)
    for _ in range(5):
        dev.smooth(0, wl.smooth_radius, wl.smooth_ratio)
    dev.event_record(13)
    dev.sync()
    smooth_ms = dev.event_elapsed_ms(12, 13) / 5
    alg = algorithmic_bytes(st)
    log(f"config-3 field: {rays} rays in {trace_ms:.3f} ms = {rays / trace_ms / 1e6:.2f} G rays/s; "
        f"smoothing 512x256 r=5 {smooth_ms:.3f} ms; counted {st}")
    return {"workload": wl.name, "rays": rays, "trace_ms": trace_ms, "rays_per_s": rays / (trace_ms / 1e3),
            "smooth_ms": smooth_ms, "algorithmic_bytes_per_launch": alg,
            "algorithmic_gbs": alg / (trace_ms / 1e3) / 1e9, "counted_work": st}


def bench_adjoint(dev, steps: int = 64) -> dict:
    """Differentiated form on the PS10-like field (624 heliostats x 2 aim offsets + sun direction,
    1.88M rays, 128x128 bilinear map): cost of the adjoint against a plain forward trace of the same
    rays, then `steps` gradient steps of aim-point optimisation."""
    import numpy as np

    from solarenergy_chier_amd import Scene, workloads

    T, PK, LAM = workloads.ADJOINT_TARGET, workloads.ADJOINT_PEAK, workloads.ADJOINT_LAMBDA
    wl, offs = workloads.adjoint_ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    nh = scene.counts.n_heliostats
    ids = list(range(nh))
    rays = sum(scene.heliostat_info(h).n_micro for h in ids) * scene.config.num_sunshape_lights_per_group
    buf = dev.flux_device_ptr(0)
    npix = 128 * 128
    # warm-up of every kernel involved
    dev.trace(ids)
    dev.adjoint_forward(0, buf)
    dev.flux_loss(buf, npix, T, PK, LAM)
    dev.adjoint_backward(scene, 0)
    # medians over the repetitions: the backward returns host arrays, so its span includes a stream
    # synchronisation, which the host's scheduler now and then stretches by milliseconds (one call in
    # about twelve on the measured boxes; the kernels themselves vary by a few percent)
    reps = 9
    tp, tf, tb = [], [], []
    for _ in range(reps):
        dev.sync()
        dev.event_record(6)
        dev.trace(ids)
        dev.event_record(7)
        dev.adjoint_forward(0, buf)
        dev.event_record(8)
        loss = dev.flux_loss(buf, npix, T, PK, LAM)
        dev.event_record(9)
        ga, gs = dev.adjoint_backward(scene, 0)
        dev.event_record(10)
        dev.sync()
        tp.append(dev.event_elapsed_ms(6, 7))
        tf.append(dev.event_elapsed_ms(7, 8))
        tb.append(dev.event_elapsed_ms(9, 10))
    t_plain, t_fwd, t_bwd = statistics.median(tp), statistics.median(tf), statistics.median(tb)
    log(f"adjoint: {rays} rays; plain trace {t_plain:.3f} ms, differentiable forward {t_fwd:.3f} ms, "
        f"backward {t_bwd:.3f} ms -> adjoint:forward {t_bwd / t_plain:.2f}, gradient (forward + backward) "
        f"{(t_fwd + t_bwd) / t_plain:.2f}")
    fd = crn_finite_differences(dev, scene, offs, buf, npix, T, PK, LAM)
    # aim-point optimisation: normalised gradient steps, offsets kept within +-2 m
    o = offs.copy()
    losses = []
    t = time.perf_counter()
    for step in range(steps):
        scene.set_aim_offsets(o)
        scene.process()
        dev.update(scene)
        dev.adjoint_forward(0, buf)
        losses.append(dev.flux_loss(buf, npix, T, PK, LAM))
        ga, gs = dev.adjoint_backward(scene, 0)
        g = ga / max(float(np.abs(ga).max()), 1e-30)
        o = np.clip(o - 0.1 * g, -2.0, 2.0).astype(np.float32)
    opt_s = time.perf_counter() - t
    log(f"aim-point optimisation: {steps} steps in {opt_s:.2f}s, loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    return {"metric": "adjoint:forward cost", "workload": wl.name, "parameters": 2 * nh + 3, "rays": rays,
            "plain_forward_ms": t_plain, "differentiable_forward_ms": t_fwd, "backward_ms": t_bwd,
            "adjoint_to_forward": t_bwd / t_plain, "adjoint_plus_forward_to_forward": (t_bwd + t_fwd) / t_plain,
            "cost_note": "adjoint_to_forward = backward / plain forward trace of the same rays; one gradient "
                         "evaluation costs the differentiable forward plus the backward; each time is the "
                         f"median of {reps} repetitions (HIP events; the backward's includes its host part)",
            "backward_ms_max": max(tb),
            "crn_finite_differences": fd,
            "optimisation_steps": steps, "loss_first": losses[0], "loss_last": losses[-1],
            "optimisation_s": opt_s}


def crn_finite_differences(dev, scene, offs, buf, npix, T, PK, LAM, n_check: int = 8, nodes: int = 128) -> dict:
    """Central differences with common random numbers (same pools, offsets, frozen visibility; the
    masked forward accumulates in double) of the loss w.r.t. the n_check largest aim-offset gradient
    components, against the adjoint, at the configuration's 1e-3 m step and at 1e-4 and 1e-6 m.

    The bilinear splat is piecewise linear in each hit position, kinked where a hit crosses a pixel
    centre, so the loss is continuous and piecewise smooth along a coordinate and its central
    difference over [x - h, x + h] equals the mean of the derivative over that interval (fundamental
    theorem of calculus), not the derivative at x. Beside the pointwise comparison, the difference
    quotient is therefore also compared with the adjoint averaged over the interval: the adjoint of
    the frozen-visibility model at `nodes` midpoint nodes (dL/dF from the double masked map at each
    node, the visibility of the base forward)."""
    import numpy as np

    scene.set_aim_offsets(offs)
    scene.process()
    dev.update(scene)
    dev.adjoint_forward(0, buf)
    dev.flux_loss(buf, npix, T, PK, LAM)
    ga, _ = dev.adjoint_backward(scene, 0)

    def dloss(fp, fm):
        return float(((fp - fm) * (fp + fm - 2 * T)).sum()
                     + LAM * (np.maximum(fp - PK, 0) ** 2 - np.maximum(fm - PK, 0) ** 2).sum())

    def flux_at(o):
        scene.set_aim_offsets(o)
        scene.process()
        dev.update(scene)
        return dev.adjoint_forward_masked(0)

    def adjoint_at(o):
        f = flux_at(o)
        dev.flux_write(0, (2.0 * (f - T) + 2.0 * LAM * np.maximum(f - PK, 0.0)).astype(np.float32))
        g, _ = dev.adjoint_backward(scene, 0, d_grad=buf)
        return g

    order = np.argsort(-np.abs(ga).ravel())[:n_check]
    out = {"components": int(n_check), "bar": 1e-3, "interval_mean_nodes": int(nodes)}
    t0 = time.perf_counter()
    for h in (1e-3, 1e-4, 1e-6):
        fd_v, ad_v, im_v = [], [], []
        for flat in order:
            hh, k = divmod(int(flat), 2)
            op, om = offs.copy(), offs.copy()
            op[hh, k] += h
            om[hh, k] -= h
            lo, hi = float(om[hh, k]), float(op[hh, k])  # the float32 offsets actually used
            fd_v.append(dloss(flux_at(op), flux_at(om)) / (hi - lo))
            ad_v.append(ga[hh, k])
            if h >= 1e-4:
                acc = 0.0
                for m in range(nodes):
                    o = offs.copy()
                    o[hh, k] = np.float32(lo + (m + 0.5) * (hi - lo) / nodes)
                    acc += float(adjoint_at(o)[hh, k])
                im_v.append(acc / nodes)
        fd_v, ad_v = np.array(fd_v), np.array(ad_v)
        err = float(np.linalg.norm(fd_v - ad_v) / np.linalg.norm(ad_v))
        out[f"rel_err_step_{h:g}m"] = err
        msg = f"adjoint: CRN central differences, step {h:g} m: relative error {err:.2e} over {n_check} components"
        if im_v:
            im_v = np.array(im_v)
            e_im = float(np.linalg.norm(fd_v - im_v) / np.linalg.norm(im_v))
            out[f"rel_err_step_{h:g}m_interval_mean_adjoint"] = e_im
            msg += f"; against the adjoint averaged over the step interval {e_im:.2e}"
        log(msg)
    out["seconds"] = time.perf_counter() - t0
    flux_at(offs)
    return out


def copy_probe(mib: int) -> int:
    """Device-to-device copy of a known size (HBM counter calibration for rocprofv3 --pmc runs)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: {mib} MiB read + {mib} MiB written")
    return 0


if __name__ == "__main__":
    sys.exit(main())
