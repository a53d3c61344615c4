"""The library's host runtime on the CPU, under AddressSanitizer (no GPU).

``msw_runtime.cpp`` is compiled with g++ against ``hostcheck/``: a CPU
stand-in for the HIP runtime and for the kernel launchers that checks every
buffer a launch or copy would touch against the live allocations, enforces
HIP's stream-capture rules, orders work with vector clocks per stream (so a
missing event between the main and the side stream shows up as a race) and
can make a chosen copy, launch or allocation fail.  ``hostcheck/drive.py``
then runs, through the C ABI:

* the launch-mapping test's sequence (stripes 0-7 with blocks of 64-512
  threads, graphs captured and replayed, buffers freed between shapes) and
  the autotuner's, for FP64 and FP32 members at B = 1, 4, 64 and 256;
* every other entry point, eagerly and through graphs, with dissipation,
  drag, wind stress, mesh scaling and the frozen-wind case;
* a failure of each copy (eager, replayed and while a graph is captured), of
  a launch and of each allocation in msw_create, each of which must come
  back as MSW_EHIP / MSW_ENOMEM with the context still usable and nothing
  leaked;
* 40 distinct member buffer pairs through the graph cache, which must stay
  bounded.
"""

import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "mpas-sw-tl-ad_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_runtime_under_asan(tmp_path):
    lib = tmp_path / "libmsw_hostcheck.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-Wall", "-fsanitize=address",
                    "-fno-omit-frame-pointer", "-I" + os.path.join(HERE, "hostcheck"), "-o", str(lib),
                    os.path.join(CSRC, "msw_runtime.cpp"), os.path.join(HERE, "hostcheck", "mock_device.cpp")],
                   check=True)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(HERE, "hostcheck", "drive.py"), str(lib)],
                       capture_output=True, text=True, env=env, timeout=900)
    print(r.stdout)
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "host check ok" in r.stdout
