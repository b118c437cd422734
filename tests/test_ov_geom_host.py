"""The SCF engine's overlap partial-sum buffer against every overlap launch
the engine makes (host only: the geometry header compiled with g++ into a
small checker, tests/cpp/ov_geom_check.cpp).  An undersized buffer here
once let a 300-band Hermitian overlap write past its end on the GPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_engine_overlap_buffer_covers_every_launch(tmp_path):
    exe = str(tmp_path / "ov_geom_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "cpp", "ov_geom_check.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 do not fit" in r.stdout
