"""eik_cli: the command line of the reference's example programs (src/examples/example1.cu:39-79
for tet meshes, example2.cu:34-84 for triangle meshes) on the C++ facade.  Flag parsing is
checked on the CPU (--print-options stops before any mesh is read); the solves run on the GPU."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

CLI = os.path.join(ROOT, "sci-solver_eikonal_amd", "eik_cli")


def _cli(*args, cwd=None, timeout=120):
    if not os.path.exists(CLI):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "sci-solver_eikonal_amd")], check=True)
    print("eik_cli", *args, flush=True)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)


def _opts(r):
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(line.split("=", 1) for line in r.stdout.strip().splitlines())


def test_3d_flags(tmp_path):
    a = tmp_path / "a.txt"
    b = tmp_path / "b.txt"
    a.write_text("1 2 3\n4 5 6\n")
    b.write_text("7\n8\n")
    o = _opts(_cli("3d", "-v", "-i", "mymesh", "-b", "12", "-m", "7", "-x", str(a), str(b),
                   "--print-options"))
    assert o == {"mesh": "3d", "input": "mymesh", "verbose": "1", "max_blocks": "12",
                 "max_verts_per_block": "64", "max_iterations": "7", "speed_type": "ONE",
                 "speed_values": "8"}


def test_3d_defaults():
    o = _opts(_cli("3d", "--print-options"))
    # example1's default mesh and the class defaults (src/Eikonal.cu:4-24)
    assert o["input"] == "../src/test/test_data/sphere334"
    assert (o["max_blocks"], o["max_verts_per_block"], o["max_iterations"]) == ("16", "64", "100")


@pytest.mark.parametrize("stype,want", [("CURVATURE", "CURVATURE"), ("NOISE", "NOISE"),
                                        ("MANUAL", "NOISE"), ("ONE", "ONE")])
def test_2d_flags(stype, want):
    """-b is vertices per block and -n the iteration cap in example2; MANUAL_SPEED has NOISE's
    value there (TriMesh.h:21-26)."""
    o = _opts(_cli("2d", "-i", "sq.ply", "-b", "32", "-n", "9", "-s", stype, "--print-options"))
    assert (o["mesh"], o["input"], o["max_verts_per_block"], o["max_iterations"],
            o["speed_type"]) == ("2d", "sq.ply", "32", "9", want)
    # the 2D default mesh is the class default (sphere_74verts.ply)
    assert _opts(_cli("2d", "--print-options"))["input"] == "../src/test/test_data/sphere_74verts.ply"


@pytest.mark.parametrize("kind", ["3d", "2d"])
def test_help_exits_zero(kind):
    r = _cli(kind, "-h")
    assert r.returncode == 0 and r.stdout.startswith("Usage: eik_cli %s" % kind)
    assert "-x" in r.stdout


def test_missing_mesh_exits_one(tmp_path):
    r = _cli("3d", "-i", str(tmp_path / "nothere"), cwd=tmp_path)
    assert r.returncode == 1 and "File open failed" in r.stdout


def test_bad_mode():
    assert _cli("4d").returncode == 64


def _vtk_values(path):
    lines = open(path).read().splitlines()
    i = lines.index("LOOKUP_TABLE default")
    return [float(x) for x in lines[i + 1:] if x.strip()]


@pytest.mark.gpu
def test_3d_solve_writes_vtk(tmp_path):
    r = _cli("3d", "-i", os.path.join(GOLDEN, "sphere334"), "-v", cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Converged to" in r.stdout  # the error graph (example1 prints it with -v)
    vals = _vtk_values(tmp_path / "result0.vtk")
    assert len(vals) == 334 and vals[0] == 0.0 and max(vals) < 100.0


@pytest.mark.gpu
def test_2d_solve_with_speeds_writes_vtk(tmp_path):
    r = _cli("2d", "-i", os.path.join(GOLDEN, "SquareMesh_size32.ply"), "-x",
             os.path.join(GOLDEN, "square_scalars.txt"), cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = _vtk_values(tmp_path / "result0.vtk")
    assert len(vals) == 1024 and vals[0] == 0.0
    # the example's own speed file: 20 inside the disc of radius 6 around (16, 16)
    s = [float(x) for x in open(tmp_path / "square_scalars.txt").read().split()]
    assert len(s) == 1922 and set(s) == {1.0, 20.0}
