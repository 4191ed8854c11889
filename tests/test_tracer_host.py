"""The height-field tracer's grid walk against brute force, on the CPU.

``csrc/hf_device.hpp`` is plain C++ apart from its HIP include; compiled for
the host with g++, its ``trace`` (cell walk with exact plane skips and the
per-cell height cull) must return exactly the closest hit that testing every
triangle of the scene with the same ``mt_test`` returns -- same triangle,
same t and barycentrics, bit for bit -- for rays from above, inside, the
bottom face and the sides, in random, vertical and grazing directions.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "shapefromcaustics_amd", "csrc", "hf_device.hpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_trace_equals_brute_force(tmp_path):
    with open(HEADER) as f:
        text = f.read()
    assert "#include <hip/hip_runtime.h>" in text
    (tmp_path / "hf_device_host.hpp").write_text(text.replace("#include <hip/hip_runtime.h>", ""))
    exe = tmp_path / "trace_bruteforce"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(tmp_path), "-I", os.path.join(HERE, "..", "include"),
                    os.path.join(HERE, "host", "trace_bruteforce.cpp"), "-o", str(exe)], check=True)
    res = subprocess.run([str(exe), "4000"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    total, hits, bad = (int(x) for x in res.stdout.split())
    assert total == 16000 and hits > 4000 and bad == 0
