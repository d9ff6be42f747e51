"""The traversal-culling bound is conservative: a quad it rules out for a heliostat's source region
and ray cone is missed by every sampled ray of that region and cone (cull_bound_check.cpp, built
here with g++ against the kernels' own header). Culling only skips traversals this bound proves
empty; the GPU parity tests show that ray records are unchanged with it."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_cone_bound_never_culls_a_hit(tmp_path):
    exe = tmp_path / "cull_bound_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(ROOT / "solarenergy_chier_amd" / "csrc"),
                    str(ROOT / "tests" / "cull_bound_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "60000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    cases, culled, rays, bad = (int(x) for x in r.stdout.split()[-4:])
    assert r.returncode == 0 and bad == 0
    # the check is only meaningful if a good share of the cases is culled and traced
    assert culled > cases // 10 and rays == 400 * culled


def test_group_ball_rejection_covers_its_quads(tmp_path):
    """The culling scan tests a heliostat's group of quads as one ball before the quads: the ball test
    (sphere_may_hit) is monotone, so a group ball that fails it proves every quad ball inside it fails
    too (ball_group_check.cpp: random cones and groups of 1-28 balls, tight and loose group balls)."""
    exe = tmp_path / "ball_group_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(ROOT / "solarenergy_chier_amd" / "csrc"),
                    str(ROOT / "tests" / "ball_group_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "200000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    groups, rejected, members, bad = (int(x) for x in r.stdout.split()[-4:])
    assert r.returncode == 0 and bad == 0
    assert rejected > groups // 10 and members > rejected
