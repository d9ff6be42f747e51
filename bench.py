"""Benchmark: rays/s of the plain tracer on one GPU, PS10-like field (624 heliostats, 483M rays per
step, 256x256 flux map), with the trace kernel's algorithmic-byte roofline and the CPU oracle timed
on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ps10|small] [--no-cpu-baseline]

A step = clear the flux map + trace every heliostat of the field (one trace call: the culling
kernel, the trace kernel and its folds) with the sample pools and micro-facet offsets resident in
HBM. Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


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

    for _ in range(args.warmup):
        dev.flux_clear(0)
        dev.trace(ids)
    dev.sync()
    log(f"warmup done ({args.warmup} steps)")

    # timed region: K steps, events around the whole region and, recorded by the library on the
    # trace stream, around each launch's trace kernel alone (the culling kernels and folds of the
    # call are in the step time, not in the kernel time)
    kernel_ms = 0.0
    dev.sync()
    dev.kernel_events(2, 3)
    wall0 = time.perf_counter()
    dev.event_record(0)
    for k in range(args.steps):
        dev.flux_clear(0)
        dev.trace(ids)
        kernel_ms += dev.event_elapsed_ms(2, 3)
    dev.event_record(1)
    dev.sync()
    dev.kernel_events(-1, -1)
    wall = time.perf_counter() - wall0
    total_ms = dev.event_elapsed_ms(0, 1)
    ms_per_step = total_ms / args.steps
    kernel_avg_ms = kernel_ms / args.steps
    rays_per_s = rays_per_step / (ms_per_step / 1e3)
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step (wall {wall:.3f}s), trace kernel {kernel_avg_ms:.3f} ms, "
        f"{rays_per_s / 1e9:.3f} G rays/s")
    flux = dev.flux(0)
    log(f"flux map sum {float(flux.sum(dtype=np.float64)):.6e}, max {float(flux.max()):.3f}")

    achieved = alg_bytes / (kernel_avg_ms / 1e3) / 1e9
    # HBM bytes per launch and VALU busy fraction of this kernel, from the committed rocprofv3 PMC
    # summaries of the same workload (profiles/trace_<workload>_{hbm_traffic,valu}.md)
    traffic = valu_busy = atom = None
    tfile = ROOT / "profiles" / f"trace_{args.workload}_hbm_traffic.json"
    vfile = ROOT / "profiles" / f"trace_{args.workload}_valu.json"
    afile = ROOT / "profiles" / f"trace_{args.workload}_atomics.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
    if vfile.exists():
        valu_busy = json.loads(vfile.read_text()).get("valu_busy")
    if afile.exists():
        atom = json.loads(afile.read_text())
    # deposit rates of the trace against the device's measured atomic ceilings (capability probe)
    ceil = dev.probe_atomic_rates()
    log(f"atomic ceilings (G ops/s): {ceil}")
    atomics = {"ceiling_gops": ceil}
    if atom:
        lds_rate = atom["lds_atomic_lanes_per_launch"] / (kernel_avg_ms / 1e3) / 1e9
        glob_rate = atom["global_atomics_per_launch"] / (kernel_avg_ms / 1e3) / 1e9
        atomics.update({"lds_atomic_lanes_per_launch": atom["lds_atomic_lanes_per_launch"],
                        "global_atomics_per_launch": atom["global_atomics_per_launch"],
                        "lds_gops": lds_rate, "lds_frac": lds_rate / ceil["lds_f32"],
                        "global_gops": glob_rate, "global_frac": glob_rate / ceil["global_f64"]})

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
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_trace_band (rectangle receiver; persistent, LDS band of map rows)",
            "kernel_avg_ms": kernel_avg_ms,
            "algorithmic_bytes_per_launch": alg_bytes,
            "hbm_gbs_measured": traffic / (kernel_avg_ms / 1e3) / 1e9 if traffic else None,
            "binding_resource": "vector issue (VALU about 80% busy) at 4 waves per SIMD, the occupancy the LDS band allows",
            "valu_busy": valu_busy,
            "atomics": atomics,
        },
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_ps10_sampled": batched_ps10,
        "adjoint": adjoint,
        "config3_cylinder_field": cylinder,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


def cpu_baseline(scene, dev, ids, L, P, F, seconds):
    """The oracle (CPU restatement) on a strided sample of the same rays, all host threads (<= 16)."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O  # the checker / CPU baseline only

    threads = max(1, min(16, os.cpu_count() or 1))
    orc = O.OracleScene.from_product_scene(scene)
    dirs, perts = dev.pool_get(P)
    offs = np.concatenate([dev.offsets_get(h, F[h]) for h in ids])
    # two calibration samples separate the per-heliostat facet set-up from the per-ray cost, then
    # the stride is chosen for about `seconds` of CPU work
    total = sum(F) * L
    runs = []
    for stride in (4096, 1024):
        t = time.perf_counter()
        _, st = orc.trace_many(ids, dirs, perts, L, offs, nthreads=threads, stride=stride)
        runs.append((st["rays"], time.perf_counter() - t))
    (r1, t1), (r2, t2) = runs
    rate = (r2 - r1) / max(t2 - t1, 1e-6)
    setup = max(t1 - r1 / rate, 0.0)
    want = max(rate * max(seconds - setup, 1.0), 1.0)
    stride = max(1, int(total / want))
    log(f"cpu baseline: calibration {rate / 1e6:.2f} M rays/s on {threads} threads (+{setup:.2f}s set-up); "
        f"sample stride {stride}")
    t = time.perf_counter()
    _, st = orc.trace_many(ids, dirs, perts, L, offs, nthreads=threads, stride=stride)
    dt = time.perf_counter() - t
    value = st["rays"] / dt
    log(f"cpu baseline: {st['rays']} rays in {dt:.2f}s = {value / 1e6:.3f} M rays/s")
    return {"value": value, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}-th ray of all {len(ids)} heliostats ({st['rays']} rays, {dt:.1f} s)"}


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
    reps = 5
    t_plain = t_fwd = t_bwd = 0.0
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
        t_plain += dev.event_elapsed_ms(6, 7)
        t_fwd += dev.event_elapsed_ms(7, 8)
        t_bwd += dev.event_elapsed_ms(9, 10)
    t_plain, t_fwd, t_bwd = t_plain / reps, t_fwd / reps, t_bwd / reps
    log(f"adjoint: {rays} rays; plain trace {t_plain:.3f} ms, differentiable forward {t_fwd:.3f} ms, "
        f"backward {t_bwd:.3f} ms -> adjoint:forward {t_bwd / t_plain:.2f}")
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
            "optimisation_steps": steps, "loss_first": losses[0], "loss_last": losses[-1],
            "optimisation_s": opt_s}


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
