"""Host builds of device-header code: the compact Morton key round trip
(compact_key -> cell_of_key) over every cell of small levels and a lattice of
cells up to level 9."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_compact_key_round_trip(tmp_path):
    exe = tmp_path / "key_roundtrip"
    src = os.path.join(ROOT, "tests", "cpp", "key_roundtrip.cpp")
    inc = os.path.join(ROOT, "fluidsolver_amd", "csrc")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I", inc, src, "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "bad 0" in r.stdout
