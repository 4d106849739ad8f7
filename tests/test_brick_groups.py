"""The brick-list builders' two-level plane test, on the CPU.

The list kernels test each section's plane against a 2x2x2 group of bricks
first and against the member bricks only where the group's box is hit.
tests/host/brick_groups.hip runs the same group and member functions on the
host for every brick shape at several box sizes and for random, axis-aligned,
diagonal and corner-tangent plane normals, and checks that the groups
partition the brick grid, that the member test is the per-brick geometry
(expanded box crossed by the plane, box reaching into the sample disk,
restated in double precision), and that a missed group never holds a member
the per-brick test keeps.  A planted shrink of the group box (-0.3 voxel) and
a planted swap of the member index's x and y bits each fail thousands of the
35 million checks."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_brick_groups_never_drop_a_pair(tmp_path):
    exe = str(tmp_path / "brick_groups")
    subprocess.check_call([HIPCC, "-std=c++17", "-O2", "--offload-arch=gfx950", "-ffp-contract=off",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                           "-I", os.path.join(ROOT, "spider_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                           "-o", exe, os.path.join(HERE, "host", "brick_groups.hip")],
                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "brick group checks passed" in p.stdout, p.stdout + p.stderr
