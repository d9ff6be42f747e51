"""The walk-free blocking proof (dda_visit_proof, qmc_math.h) is sound: whenever it holds for a quad a ray
hits, the oracle's restatement of the reference traversal (rectGridDDA.cu:31-119) reports a hit for that
ray. dda_proof_check.cpp is built here with g++ against the kernels' own header and the oracle, and run on
oracle scenes of the fields the trace is measured on (PS10-like sampled field, the config-3 cylinder field,
the small field) and on grids whose cells are not multiples of the heliostat spacing, with many rays aimed
at points pushed onto cell faces and edges and with direction components next to the traversal's
never-step threshold. The trace uses the proof only where it holds (the traversal decides the rest), so
this is what keeps its statuses equal to the reference's."""
import ctypes as C
import dataclasses
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from solarenergy_chier_amd import workloads

ROOT = Path(__file__).resolve().parents[1]


# the margins: the box's face margin (1 mm and rounding slack) and the per-ray term
FACE_MARGIN = "const double margin = 1e-3 + 4.0 * ext * 0x1p-22 + proof_slack(ext);"
RAY_MARGIN = "(fabsf(rp - mid[a]) < lim[a] - S)"


def build_checker(d, csrc):
    O.lib()  # the oracle is built and loaded first
    so = d / "libdda_proof_check.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-I", str(csrc),
                    "-I", str(ROOT / "oracle"), "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "dda_proof_check.cpp"), "-L", str(O.LIB.parent), "-loracle",
                    "-Wl,-rpath," + str(O.LIB.parent), "-o", str(so)], check=True)
    fn = C.CDLL(str(so)).dda_proof_check
    fp3 = C.POINTER(C.c_float)
    fn.argtypes = [C.c_void_p, C.c_int, fp3, fp3, C.c_int, C.c_uint64, C.POINTER(C.c_longlong)]
    fn.restype = C.c_int
    return fn


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    return build_checker(tmp_path_factory.mktemp("dda_proof"), ROOT / "solarenergy_chier_amd" / "csrc")


def run(checker, wl, n_rays, seed=11, grid=0):
    orc = O.OracleScene.from_workload(wl)
    g = wl.grids[grid]
    pos = (C.c_float * 3)(*g.pos)
    iv = (C.c_float * 3)(*g.interval)
    out = (C.c_longlong * 8)()
    rc = checker(orc.handle, grid, pos, iv, n_rays, seed, out)
    assert rc == 0, rc
    rays, hit, proofs, dda, bad, notfull, unvisited, single = list(out)
    print(f"{wl.name[:44]}: rays {rays}, with a hit {hit}, proofs {proofs}, oracle hits {dda}, "
          f"violations {bad}, boxes not full {notfull}, proven boxes not visited {unvisited}, "
          f"single-cell proofs {single}", flush=True)
    return rays, hit, proofs, dda, bad, notfull, unvisited, single


def odd_grid(wl, interval):
    """The same field on a grid of other cell sizes (cells no longer line up with the heliostat rows)."""
    g = wl.grids[0]
    size = tuple(float(np.ceil(s / i) * i) for s, i in zip(g.size, interval))
    wl.grids[0] = dataclasses.replace(g, interval=tuple(interval), size=size)
    wl.name = wl.name + f" on {interval} m cells"
    return wl


@pytest.mark.parametrize("case", ["ps10", "cylinder", "small", "ps10_odd", "small_odd"])
def test_proof_never_claims_a_miss(checker, case):
    wl = {"ps10": lambda: workloads.ps10_field(L=1),
          "cylinder": workloads.cylinder_field,
          "small": workloads.small_field,
          "ps10_odd": lambda: odd_grid(workloads.ps10_field(L=1), (3.7, 2.9, 4.3)),
          "small_odd": lambda: odd_grid(workloads.small_field(), (1.3, 0.7, 2.1))}[case]()
    rays, hit, proofs, dda, bad, notfull, unvisited, single = run(checker, wl, 30000)
    assert notfull == 0  # every cell of a quad's cell box lists it (the proof's premise)
    assert bad == 0 and unvisited == 0
    # the check means something only if the proof holds for many rays that do hit, and for many
    # single cells
    assert hit > rays // 10 and proofs > hit // 2 and single > hit // 2


def test_proof_without_its_margin_is_caught(tmp_path):
    """The same check rejects a planted slip: the proof with its margins dropped (any point strictly inside
    the box) claims cells that the traversal passes by at an edge, where it steps two axes at once."""
    d = tmp_path / "slip"
    shutil.copytree(ROOT / "solarenergy_chier_amd" / "csrc", d)
    h = (d / "qmc_math.h").read_text()
    assert h.count(FACE_MARGIN) == 1 and h.count(RAY_MARGIN) == 1
    h = h.replace(FACE_MARGIN, "const double margin = 0.0;").replace(RAY_MARGIN, "(fabsf(rp - mid[a]) < lim[a])")
    (d / "qmc_math.h").write_text(h)
    slip = build_checker(tmp_path, d)
    *_, unvisited, single = run(slip, workloads.small_field(), 60000)
    assert unvisited > 0
