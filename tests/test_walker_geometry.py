"""Walker geometry on the CPU (deeprwcap_amd/csrc/walker_geometry.h, built
with g++ into tests/walker_geometry_check.cpp), on the reference's ten test
structures (testcases/cap3d/case1-10.cap3d, fixtures under
tests/golden/cap3d) and on a made-up net with coplanar faces and a hole:

* the grid queries equal scans over every box at every point (nearest
  conductor and its net, the cube limit from layers and the dielectric boxes
  that show in the cube, the boxes listed in the cube's description);
* every fourth cube's description voxelizes (N = 23, first listed box wins,
  as in cuda_kernels.cu:44-77) to the permittivity of the structure itself at
  every voxel centre: the last listed dielectric box holding it (later boxes
  override earlier ones), else its planar layer;
* the Gaussian surface: no grown master box overlaps another conductor;
  accepted draws lie on the boundary of the union of the grown boxes and
  rejected ones do not (except on a coplanar face claimed by an earlier box),
  and the mean draw weight (faces drawn in proportion to area / margin,
  weighed by Z margin; zero when rejected) equals the union's boundary area
  computed exactly on the grid of box coordinates, within 4 standard
  errors.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from deeprwcap_amd import cap3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "cap3d")


def case_path(c):
    p = os.path.join(GOLD, f"case{c}.cap3d")
    return p if os.path.exists(p) else p + ".gz"


def reference_masters(c):
    ref = os.path.join(GOLD, f"case{c}.dspf" if c < 10 else "case10.spice")
    return list(cap3d.read_reference(ref))


def write_geometry(path, s, master_net, frac=0.5):
    order, ids = cap3d.nets(s)
    bx = np.ascontiguousarray(s["boxes"], np.float64)
    win = np.asarray(s["window"], np.float64)
    mb = np.ascontiguousarray(s["media_boxes"], np.float64).reshape(-1, 6)
    with open(path, "wb") as f:
        f.write(np.array([len(s["layer_z"]), len(bx), len(mb), order.index(master_net)], np.int32).tobytes())
        f.write(win.tobytes())
        f.write(np.asarray(s["layer_z"], np.float64).tobytes())
        f.write(np.asarray(s["layer_eps"], np.float64).tobytes())
        f.write(bx.tobytes())
        f.write(ids.astype(np.int32).tobytes())
        f.write(mb.tobytes())
        f.write(np.asarray(s["media_eps"], np.float64).tobytes())
        scale = float(np.min(bx[:, 3:] - bx[:, :3]))
        f.write(np.array([frac, float(np.max(win[3:] - win[:3])), 1e-3 * scale], np.float64).tobytes())


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = str(tmp_path_factory.mktemp("wgeo") / "walker_geometry_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", os.path.join(ROOT, "tests", "walker_geometry_check.cpp"),
                    "-o", exe], check=True, timeout=300)
    return exe


def run_check(exe, tmp_path, s, master, points=3000, seed=7):
    geo = str(tmp_path / "geo.bin")
    write_geometry(geo, s, master)
    r = subprocess.run([exe, geo, str(points), str(seed)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        out.setdefault(k, v)
    print(r.stdout.replace("\n", " | "))
    return out


def assert_geometry(out):
    for k in ("bad_nearest", "bad_bounded", "bad_fit", "bad_hits", "bad_describe", "bad_voxels"):
        assert int(out[k]) == 0, (k, out)
    assert int(out["voxels"]) > 100 * int(out["voxels_at_interfaces"])
    assert "gauss_error" not in out, out["gauss_error"]
    assert int(out["gauss_bad_accept"]) == 0 and int(out["gauss_bad_reject"]) == 0
    assert int(out["gauss_bad_clearance"]) == 0 and float(out["gauss_min_margin"]) > 0
    exact, mc, se = (float(out[k]) for k in ("gauss_exact_area", "gauss_mc_area", "gauss_mc_stderr"))
    assert abs(mc - exact) <= 4 * se + 1e-12 * exact, (mc, exact, se)


@pytest.mark.parametrize("case", range(1, 11))
def test_reference_structures(checker, tmp_path, case):
    s = cap3d.read(case_path(case))
    masters = reference_masters(case)
    for m in masters[:2]:
        out = run_check(checker, tmp_path, s, m, points=3000 if case > 3 else 1500)
        assert int(out["points"]) >= 1500
        assert_geometry(out)
        if len(cap3d.nets(s)[0]) > 1 and int(out["gauss_boxes"]) > 1:
            # several boxes: the union's boundary is smaller than the faces' sum
            assert float(out["gauss_exact_area"]) < float(out["gauss_face_area"])


def _ring_structure(extra):
    ring = [[0, 0, 0, 4, 1, 1], [0, 3, 0, 4, 4, 1], [0, 1, 0, 1, 3, 1], [3, 1, 0, 4, 3, 1]]
    other = [[1.6, 1.6, 0, 2.4, 2.4, 1]]
    boxes = np.array(ring + extra + other, np.float64)
    ids = np.array([0] * (4 + len(extra)) + [1], np.int32)
    return {"window": np.array([-10, -10, -10, 20, 20, 20.0]), "layer_z": np.array([-1e30, 0.5]),
            "layer_eps": np.array([3.9, 2.5]), "media_boxes": np.zeros((0, 6)), "media_eps": np.zeros(0),
            "boxes": boxes, "ids": ids, "names": ["M", "X"], "masters": ["M"]}


def test_net_with_hole_and_coplanar_faces(checker, tmp_path):
    """A square ring of four unit-thick boxes around another conductor 0.6
    away: every ring box grows by 0.3, the grown top and bottom faces are
    coplanar, and the union is the square [-0.3, 4.3]^2 with the hole
    (1.3, 2.7)^2, 1.6 high, of boundary area 2 (4.6^2 - 1.4^2) + 4 (4.6 +
    1.4) 1.6 = 76.8.  Then two overlapping boxes added on one side."""
    out = run_check(checker, tmp_path, _ring_structure([]), "M", points=2000)
    assert_geometry(out)
    assert float(out["gauss_exact_area"]) == pytest.approx(76.8, rel=1e-12)
    assert int(out["gauss_coplanar"]) > 0  # the coplanar faces were met and claimed once
    out = run_check(checker, tmp_path, _ring_structure([[4.5, 0, 0, 6, 1, 1], [5, 0.5, 0, 6.5, 1, 1]]), "M",
                    points=2000)
    assert_geometry(out)


def test_grid_beats_scan_on_large_structure(checker, tmp_path):
    """Case 2 (8,754 conductor boxes, 5,838 dielectric boxes): the grid's
    per-point query costs a small fraction of the scan's."""
    s = cap3d.read(case_path(2))
    out = run_check(checker, tmp_path, s, reference_masters(2)[0], points=1000)
    assert_geometry(out)
    assert float(out["grid_us"]) < 0.2 * float(out["scan_us"]), out
