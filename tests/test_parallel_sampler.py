"""ParallelSampler (the C++ facade's request batching, parallel_sampler.cpp:
42-211) against a CPU test double of the solver: per-task results, batch cap,
failure fan-out, default axis, drain on shutdown (tests/sampler_check.cpp).
Built with g++ against include/ and the library; the second build compiles
the facade into the test with ThreadSanitizer, so that its locking is checked
too (no GPU needed: the double implements sample() itself)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build_and_run(tmp_path, extra, name):
    from deeprwcap_amd import _lib
    libdir = os.path.dirname(_lib.LIB_PATH)
    assert os.path.exists(_lib.LIB_PATH), "libdnnsolver.so is not built"
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", f"-I{os.path.join(ROOT, 'include')}", *extra,
           os.path.join(ROOT, "tests", "sampler_check.cpp"), f"-L{libdir}", "-ldnnsolver",
           f"-Wl,-rpath,{libdir}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all cases passed" in r.stdout
    return r


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_parallel_sampler_behaviour(tmp_path):
    _build_and_run(tmp_path, [], "sampler_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_parallel_sampler_thread_sanitizer(tmp_path):
    r = _build_and_run(tmp_path, ["-fsanitize=thread", os.path.join(ROOT, "deeprwcap_amd", "csrc", "facade.cpp")],
                       "sampler_check_tsan")
    assert "ThreadSanitizer" not in r.stderr, r.stderr
