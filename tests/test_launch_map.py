"""Host check of the launch mapping the kernels use (no GPU).

``launch_map_check.cpp`` compiles the library's own ``csrc/msw_launch.hpp``
for the CPU and walks every hardware thread of every striped launch family
(entity, edge/cell and coefficient launches; stripes 0-7, blocks of 64-512
threads; 1-256 members) on the level 0, 3, 5 and 7 meshes and on the uneven
64 x L5 + 16 x L6 + 4 x L7 + 1 x L9 batch: items covered exactly once, none
past the end, whole wavefronts per entity with B >= 64 and whole 16-lane
groups in the coefficient launch.
"""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_launch_mapping_covers_items_exactly_once(tmp_path):
    exe = tmp_path / "launch_map_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", str(exe),
                    os.path.join(HERE, "launch_map_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "launch map ok" in r.stdout
