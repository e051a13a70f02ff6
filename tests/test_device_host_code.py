"""Host builds of device-header code: the cell key order of the neighbor lists
(compact_key against the reference's libmorton key, and the neighbor order
code against the key order of a cell's 27 neighbors) over every cell of levels
1-4 and a lattice of cells up to level 9."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cell_key_and_neighbor_order(tmp_path):
    exe = tmp_path / "key_order"
    src = os.path.join(ROOT, "tests", "cpp", "key_order.cpp")
    inc = os.path.join(ROOT, "fluidsolver_amd", "csrc")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I", inc, src, "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert " bad 0 " in r.stdout


def test_box_rank_is_the_key_order_within_the_box(tmp_path):
    """box_rank (the rank table of the compressed search's sort keys) against
    sorting every cell of random boxes by compact_key, boxes straddling zero
    included, and a sample of the breaking dam's cell box at level 8."""
    exe = tmp_path / "box_rank"
    src = os.path.join(ROOT, "tests", "cpp", "box_rank.cpp")
    inc = os.path.join(ROOT, "fluidsolver_amd", "csrc")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-I", inc, src, "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert " bad 0 " in r.stdout
