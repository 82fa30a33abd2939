"""The persistent host worker pool of the tree passes (CPU only): built from its
source with g++ and run on ragged jobs from two caller threads at once."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_tree_pool_runs_every_index_once(tmp_path):
    exe = tmp_path / "tree_pool_stress"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(HERE, "tree_pool_stress.cpp"), "-o", str(exe)], check=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "wrong counts 0" in out.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_tree_pool_thread_cap(tmp_path):
    """QZ_TREE_THREADS=k leaves k - 1 workers beside the caller (at least one
    thread, the caller alone), and the jobs still run every index once."""
    exe = tmp_path / "tree_pool_stress"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                    os.path.join(HERE, "tree_pool_stress.cpp"), "-o", str(exe)], check=True, timeout=300)
    for cap, want in (("3", 2), ("1", 0)):
        env = dict(os.environ, QZ_TREE_THREADS=cap)
        out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
        assert out.returncode == 0, out.stdout + out.stderr
        assert f"workers {min(want, max(0, (os.cpu_count() or 1) - 1))}," in out.stdout, out.stdout
        assert "wrong counts 0" in out.stdout
