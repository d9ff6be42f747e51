"""The reference's host C++ call surface (include/qmcrt_facade.hpp) compiled and driven the way the
reference's callers drive it (tests/facade_check.cpp):

* CPU: SceneConfiguration, SceneLoader (false on malformed scenes), object views, host processing,
  QuasiMonteCarloRayTracer::checkValidHeliostatIndex's messages (QuasiMonteCarloRayTracer.cpp:96-127),
  TaskLoader — the program is compiled here from the header with g++.
* GPU: the QMCRTracerTestFixture pattern (QuasiMonteCarloRayTracerTest.cpp:19-47,189-208):
  processScene, rayTracing(solarScene, 0), the receiver's device image; its sum equals the library's
  own trace of the same heliostat with the same seeds (pools 2, 4, offsets 3, resampled from 1000)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kat_scenes as K
from solarenergy_chier_amd import Scene
from solarenergy_chier_amd import build as B

ROOT = Path(__file__).resolve().parents[1]


def run_check(binary, mode):
    r = subprocess.run([str(binary), mode, str(K.FIX)], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    return r


def test_facade_compiles_and_host_pattern_runs(tmp_path):
    exe = tmp_path / "facade_check"
    cmd = B.facade_program_cmd(ROOT / "tests" / "facade_check.cpp", exe) + ["-Werror", "-Wl,-rpath," + str(B.LIBDIR)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    r = run_check(exe, "host")
    assert r.returncode == 0 and r.stdout.rstrip().endswith("all ok")
    assert "FAILED" not in r.stdout


@pytest.mark.gpu
def test_facade_device_pattern_equals_library(device):
    r = run_check(B.FACADE_CHECK, "device")
    assert r.returncode == 0 and "all ok" in r.stdout
    s1 = float(re.search(r"sum after rayTracing\(0\): (\S+)", r.stdout).group(1))
    s = Scene.from_files(str(K.FIX / "qmcrt_test_scene.scn"), str(K.FIX / "qmcrt_test_configuration.json"))
    device.upload(s)
    device.pool_generate(2, 4)
    device.offsets_generate(3)
    device.offsets_resample(True, 1000)
    try:
        device.flux_clear(0)
        device.trace([0])
        want = float(device.flux(0).astype(np.float64).sum())
    finally:
        device.offsets_resample(False)
    print("facade", s1, "library", want, flush=True)
    assert want > 0 and abs(s1 - want) <= 1e-5 * want
