"""Differentiable trace: the FP64 host adjoint (dual numbers) is pinned by finite differences of its
own model; the GPU adjoint is checked against it and against common-random-number finite
differences of the GPU's masked forward; the bilinear forward matches the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from solarenergy_chier_amd import Scene, workloads

T, PK, LAM = workloads.ADJOINT_TARGET, workloads.ADJOINT_PEAK, workloads.ADJOINT_LAMBDA


def small_case(L=8, groups=64):
    wl, offs = workloads.adjoint_small_field()
    wl.config.num_sunshape_lights_per_group = L
    wl.config.num_sunshape_groups = groups
    return wl, offs


def oracle_inputs(wl):
    orc = O.OracleScene.from_workload(wl)
    d, p, _ = O.pool_generate(wl.config, 2, 4)
    offs = orc.all_offsets(3, d.shape[0])
    return orc, d, p, offs


def dloss(fp, fm):
    """L(F+) - L(F-) differenced per pixel before summing (no cancellation against the 1e16 loss)."""
    return float(((fp - fm) * (fp + fm - 2 * T)).sum()
                 + LAM * (np.maximum(fp - PK, 0) ** 2 - np.maximum(fm - PK, 0) ** 2).sum())


def test_fp64_adjoint_matches_its_own_finite_differences():
    """Central differences of the dual model's own value. The bilinear splat is piecewise linear in each
    hit position (kinks where a hit crosses a pixel centre), so steps stay far below the 3 cm pixel:
    1e-6 m for aim offsets, 1e-7 for the sun direction, which moves every ray."""
    wl, _ = small_case()
    orc, d, p, offs = oracle_inputs(wl)
    L = wl.config.num_sunshape_lights_per_group
    flux, loss, ga, gs, nv = O.adjoint_fp64(orc, 0, d, p, L, offs, T, PK, LAM)
    assert nv > 0 and flux.sum() > 0
    h = 1e-6
    for helio, k in [(3, 0), (12, 1), (20, 0)]:
        delta = np.zeros(5)
        delta[k] = h
        fp_ = O.adjoint_fp64(orc, 0, d, p, L, offs, T, PK, LAM, helio, delta)[0]
        fm_ = O.adjoint_fp64(orc, 0, d, p, L, offs, T, PK, LAM, helio, -delta)[0]
        assert dloss(fp_, fm_) / (2 * h) == pytest.approx(ga[helio, k], rel=1e-6)
    h = 1e-7
    for k in range(3):
        delta = np.zeros(5)
        delta[2 + k] = h
        fp_ = O.adjoint_fp64(orc, 0, d, p, L, offs, T, PK, LAM, 0, delta)[0]
        fm_ = O.adjoint_fp64(orc, 0, d, p, L, offs, T, PK, LAM, 0, -delta)[0]
        assert dloss(fp_, fm_) / (2 * h) == pytest.approx(gs[k], rel=1e-3)


def test_fp64_forward_is_the_reference_energy_bilinearly_splatted():
    """At zero perturbation the dual forward deposits the same total energy as the reference trace
    (bilinear weights sum to one away from the map border)."""
    wl, _ = small_case()
    orc, d, p, offs = oracle_inputs(wl)
    L = wl.config.num_sunshape_lights_per_group
    flux, loss, ga, gs, nv = O.adjoint_fp64(orc, 0, d, p, L, offs, T, PK, LAM)
    ref, st = orc.trace_many(list(range(25)), d, p, L, offs, nthreads=4)
    assert nv == st["deposits"]
    # mass lost at the border only: the nearest-pixel map has no energy on the outermost ring
    inner = ref[1:-1, 1:-1].sum() / ref.sum()
    assert flux.sum() / ref.sum() == pytest.approx(1.0, abs=1.0 - inner + 1e-6)


@pytest.mark.gpu
class TestGpuAdjoint:
    @staticmethod
    def setup(device, wl):
        scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
        device.upload(scene)
        device.pool_generate(2, 4)
        device.offsets_generate(3)
        return scene

    def common(self, device, wl):
        scene = self.setup(device, wl)
        cfg = scene.config
        L = cfg.num_sunshape_lights_per_group
        P = cfg.num_sunshape_groups * L
        dirs, perts = device.pool_get(P)
        n = scene.counts.n_heliostats
        offs = np.concatenate([device.offsets_get(h, scene.heliostat_info(h).n_micro) for h in range(n)])
        return scene, L, dirs, perts, offs

    @pytest.mark.parametrize("case", ["small", "ps10", "ps10_low_sun"])
    def test_gradient_matches_fp64_adjoint(self, device, case):
        """The GPU model runs in double like the host adjoint, so the gradients agree far inside
        1e-3 even on the PS10-like field, whose gradient moves by 1e-2 under a 1e-8 rad sun shift;
        the forward map differs only by float atomic summation. Under the low sun (azimuth 90,
        elevation 10 degrees) neighbours shade flat heliostats, so the recorded visibility also
        drops shadowed rays."""
        if case == "small":
            wl, _ = workloads.adjoint_small_field()
        else:
            wl, _ = workloads.adjoint_ps10_field(sun_deg=(90.0, 10.0) if case == "ps10_low_sun" else (180.0, 50.0))
        scene, L, dirs, perts, offs = self.common(device, wl)
        buf = device.flux_device_ptr(0)
        device.adjoint_forward(0, buf)
        flux = device.flux(0)
        loss = device.flux_loss(buf, flux.size, T, PK, LAM)
        ga, gs = device.adjoint_backward(scene, 0)
        masked = device.adjoint_forward_masked(0)
        orc = O.OracleScene.from_product_scene(scene)
        f64, loss64, ga64, gs64, nv = O.adjoint_fp64(orc, 0, dirs, perts, L, offs, T, PK, LAM)
        e_flux = np.abs(flux - f64).sum() / np.abs(f64).sum()
        e_masked = np.abs(masked - f64).sum() / np.abs(f64).sum()
        e_aim = np.linalg.norm(ga - ga64) / np.linalg.norm(ga64)
        e_sun = np.linalg.norm(gs - gs64) / np.linalg.norm(gs64)
        print(f"{case}: visible {nv}, loss {loss:.9e} vs {loss64:.9e}, flux L1 {e_flux:.2e}, masked L1 "
              f"{e_masked:.2e}, aim grad {e_aim:.2e}, sun grad {e_sun:.2e}", flush=True)
        assert e_flux < 1e-5
        assert e_masked < 1e-9
        assert loss == pytest.approx(loss64, rel=1e-5)
        assert e_aim < 1e-3 and e_sun < 1e-3

    def test_masked_forward_equals_forward_at_base(self, device):
        wl, _ = workloads.adjoint_small_field()
        self.setup(device, wl)
        buf = device.flux_device_ptr(0)
        device.adjoint_forward(0, buf)
        a = device.flux(0).astype(np.float64)
        b = device.adjoint_forward_masked(0)
        assert np.abs(a - b).sum() / np.abs(b).sum() < 1e-6

    def test_gradient_matches_crn_finite_differences(self, device):
        """Central differences of the GPU's masked forward (double accumulation) with common random
        numbers: same pools, offsets and frozen visibility. The splat is piecewise linear in each hit
        position, kinked where a hit crosses a pixel centre (3 cm pixels here), so the derivative is
        checked with steps far below the pixel (1e-6 m for aim offsets, 1e-7 for the sun direction,
        which moves every ray) to 1e-3; at a 1e-3 m step the kinks inside the interval bias the
        difference quotient by about 1e-2 against the derivative at the centre, which is printed and
        bounded loosely. At that step the quotient equals the mean of the derivative over the interval
        (the loss is continuous and piecewise smooth along the coordinate), so it is also compared,
        to 1e-3, with the adjoint averaged over the interval: the adjoint at 16 midpoint nodes, each
        with dL/dF of the node's own masked map and the base forward's visibility."""
        wl, offs0 = workloads.adjoint_small_field()
        scene = self.setup(device, wl)
        buf = device.flux_device_ptr(0)
        device.adjoint_forward(0, buf)
        device.flux_loss(buf, 128 * 128, T, PK, LAM)
        ga, gs = device.adjoint_backward(scene, 0)

        def flux_at(offsets, sun=None):
            scene.set_aim_offsets(offsets)
            if sun is not None:
                cfg = scene.config
                cfg.sun_dir[:] = [float(x) for x in sun]
                scene.set_config(cfg)
            scene.process()
            device.update(scene)
            return device.adjoint_forward_masked(0)

        order = np.argsort(-np.abs(ga).ravel())[:8]
        for h, tol in ((1e-6, 1e-3), (1e-3, 3e-2)):
            fd_v, ad_v = [], []
            for flat in order:
                hh, k = divmod(int(flat), 2)
                op, om = offs0.copy(), offs0.copy()
                op[hh, k] += h
                om[hh, k] -= h
                step = float(op[hh, k]) - float(om[hh, k])  # the float32 offsets actually used
                fd_v.append(dloss(flux_at(op), flux_at(om)) / step)
                ad_v.append(ga[hh, k])
            fd_v, ad_v = np.array(fd_v), np.array(ad_v)
            err = np.linalg.norm(fd_v - ad_v) / np.linalg.norm(ad_v)
            print(f"aim h={h}: adjoint {ad_v} fd {fd_v} rel {err:.2e}", flush=True)
            assert err < tol
            if h == 1e-3:
                nodes, im_v = 16, []
                for flat in order:
                    hh, k = divmod(int(flat), 2)
                    lo, hi = float(np.float32(offs0[hh, k] - h)), float(np.float32(offs0[hh, k] + h))
                    acc = 0.0
                    for m in range(nodes):
                        o = offs0.copy()
                        o[hh, k] = np.float32(lo + (m + 0.5) * (hi - lo) / nodes)
                        f = flux_at(o)
                        device.flux_write(0, (2.0 * (f - T) + 2.0 * LAM * np.maximum(f - PK, 0.0)).astype(np.float32))
                        g, _ = device.adjoint_backward(scene, 0, d_grad=buf)
                        acc += float(g[hh, k])
                    im_v.append(acc / nodes)
                im_v = np.array(im_v)
                e_im = np.linalg.norm(fd_v - im_v) / np.linalg.norm(im_v)
                print(f"aim h={h}: adjoint averaged over the interval {im_v} rel {e_im:.2e}", flush=True)
                assert e_im < 1e-3
                flux_at(offs0)
        sun0 = np.array(scene.config.sun_dir[:], np.float32)
        fd_s = []
        for k in range(3):
            sp, sm = sun0.copy(), sun0.copy()
            sp[k] = np.float32(sp[k] + 1e-7)
            sm[k] = np.float32(sm[k] - 1e-7)
            step = float(sp[k]) - float(sm[k])
            fd_s.append(dloss(flux_at(offs0, sp), flux_at(offs0, sm)) / step)
        flux_at(offs0, sun0)
        fd_s = np.array(fd_s)
        err = np.linalg.norm(fd_s - gs) / np.linalg.norm(gs)
        print(f"sun: adjoint {gs} fd {fd_s} rel {err:.2e}", flush=True)
        assert err < 1e-3

    def test_aim_point_optimisation_reduces_loss(self, device):
        wl, offs = workloads.adjoint_small_field()
        scene = self.setup(device, wl)
        buf = device.flux_device_ptr(0)
        losses = []
        o = offs.copy()
        lr = None
        for step in range(64):
            scene.set_aim_offsets(o)
            scene.process()
            device.update(scene)
            device.adjoint_forward(0, buf)
            losses.append(device.flux_loss(buf, 128 * 128, T, PK, LAM))
            ga, _ = device.adjoint_backward(scene, 0)
            if lr is None:
                lr = 0.05 / max(np.abs(ga).max(), 1e-30)
            o = np.clip(o - lr * ga, -2.0, 2.0).astype(np.float32)
        print("loss first/last", losses[0], losses[-1])
        assert losses[-1] < losses[0]


@pytest.mark.gpu
def test_torch_autograd_binding_matches_adjoint():
    """The torch.autograd binding composes a loss written in torch with the hand-written adjoint: for
    the library's own loss its gradients equal qmc_adjoint_backward's. Runs in a child process, which
    initialises torch's GPU before loading the library (this session already holds the library)."""
    import subprocess
    import sys
    from pathlib import Path

    script = Path(__file__).resolve().parent / "gpu_scripts" / "torch_binding_check.py"
    out = subprocess.run([sys.executable, "-u", str(script)], capture_output=True, text=True, timeout=240)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0
    assert "-> ok" in out.stdout
