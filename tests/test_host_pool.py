"""The planners' host thread pool, compiled on its own with g++ (no GPU):
tests/host/pool_check.cpp, once plain and, where the toolchain has it, under
ThreadSanitizer; with the default worker count and with QLG_PLAN_THREADS=1
(serial)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "host", "pool_check.cpp"), os.path.join(ROOT, "ultradmrg_amd", "csrc", "host_pool.cpp")]


def _build(tmp_path, extra):
    exe = str(tmp_path / ("pool_check" + ("_tsan" if extra else "")))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", *extra, f"-I{os.path.join(ROOT, 'ultradmrg_amd', 'csrc')}",
           *SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return exe if r.returncode == 0 else None, r


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.parametrize("threads", [None, "1"])
def test_host_pool(tmp_path, threads):
    exe, r = _build(tmp_path, [])
    assert exe, r.stderr
    env = dict(os.environ)
    env.pop("QLG_PLAN_THREADS", None)
    if threads:
        env["QLG_PLAN_THREADS"] = threads
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    if threads == "1":
        assert "workers 0 " in out.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_pool_thread_sanitizer(tmp_path):
    exe, r = _build(tmp_path, ["-fsanitize=thread"])
    if exe is None:
        pytest.skip("g++ cannot build with -fsanitize=thread here")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and "ThreadSanitizer" not in out.stderr, out.stdout + out.stderr
