"""qmc_trace_noalloc (no device allocation: scratch in a file-scope device array) against qmc_trace,
called through the C ABI: two calls in a row, the second over fewer heliostats, each compared with the
allocating entry point on the same inputs (relative L1 <= 1e-5: float atomics reorder sums)."""
import numpy as np
import pytest

from solarenergy_chier_amd import Device, Scene, workloads
from solarenergy_chier_amd import _native as N

pytestmark = pytest.mark.gpu

FLUX_RTOL_L1 = 1e-5


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def trace_noalloc(dev, ids):
    a = np.ascontiguousarray(list(ids), np.int32)
    N.call("qmc_trace_noalloc", dev.handle, N.iptr(a), len(a))


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    yield d
    d.close()


def both(dev, ids):
    dev.flux_clear(0)
    dev.trace(ids)
    want = dev.flux(0).astype(np.float64)
    dev.flux_clear(0)
    trace_noalloc(dev, ids)
    got = dev.flux(0).astype(np.float64)
    return got, want


def test_noalloc_equals_trace_two_calls(dev):
    wl = workloads.ps10_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    n = scene.counts.n_heliostats
    # the whole field: its wave masks (121 MB) exceed what the scratch leaves, so the call is split
    got, want = both(dev, list(range(n)))
    err = rel_l1(got, want)
    print(f"PS10 full field: noalloc vs trace rel L1 {err:.2e}", flush=True)
    assert err < FLUX_RTOL_L1 and want.sum() > 0
    # a smaller second call (and a heliostat listed twice) on the same scratch
    ids = list(range(5, n, 7)) + [5]
    got, want = both(dev, ids)
    err = rel_l1(got, want)
    print(f"PS10 subset of {len(ids)}: noalloc vs trace rel L1 {err:.2e}", flush=True)
    assert err < FLUX_RTOL_L1 and want.sum() > 0


def test_noalloc_cylinder_field_and_limits(dev):
    wl = workloads.cylinder_field()
    scene = Scene.from_specs(wl.config, wl.receivers, wl.grids, wl.heliostats)
    dev.upload(scene)
    dev.pool_generate(2, 4)
    dev.offsets_generate(3)
    ids = list(range(0, scene.counts.n_heliostats, 3))
    got, want = both(dev, ids)
    err = rel_l1(got, want)
    print(f"config-3 field, {len(ids)} heliostats: noalloc vs trace rel L1 {err:.2e}", flush=True)
    assert err < FLUX_RTOL_L1
    with pytest.raises(N.QmcError) as e:
        trace_noalloc(dev, [0] * 65537)
    assert e.value.status == N.QMC_E_RANGE
