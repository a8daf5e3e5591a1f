"""The host-side parallel copy into the caller's array (sodecl_amd/csrc/host_copy.hpp),
compiled on its own and run with several thread counts."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("threads", ["1", "3", "8"])
def test_parallel_copy_covers_every_byte(tmp_path, threads):
    exe = tmp_path / "host_copy_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-Wall", "-Werror", os.path.join(HERE, "host_copy_check.cpp"),
                    "-o", str(exe)], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=threads)
    out = subprocess.run([str(exe)], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")
    assert f"workers {min(int(threads), len(os.sched_getaffinity(0)))}," in out.stdout
