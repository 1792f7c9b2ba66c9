"""Offline readers of the C++ facade (readStructureFile, readBinaryFile;
inference_cpp/src/dnn_solver.cpp:89-250) on files written by the check
program (tests/readers_check.cpp).  CPU only: the readers need no device."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_offline_readers(tmp_path):
    from deeprwcap_amd import _lib
    libdir = os.path.dirname(_lib.LIB_PATH)
    assert os.path.exists(_lib.LIB_PATH), "libdnnsolver.so is not built"
    exe = str(tmp_path / "readers_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", f"-I{os.path.join(ROOT, 'include')}",
                    os.path.join(ROOT, "tests", "readers_check.cpp"), f"-L{libdir}", "-ldnnsolver",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all cases passed" in r.stdout
