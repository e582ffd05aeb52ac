"""The reference's own test cases (src/test/3d/*.cc, src/test/2d/*.cc) compiled against the C++
facade (sci-solver_eikonal_amd/csrc/Eikonal.h) and run as a separate program, one case per
process (tests/cpp/facade_test.cpp).  The missing-file cases need no device: the facade exits
with status 1 before any device call, which is what Sanity3D / Sanity2D assert."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "sci-solver_eikonal_amd", "_build", "facade_test")


def _run(case, cwd, timeout=120):
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "sci-solver_eikonal_amd")],
                       check=True)
    print("facade case", case, flush=True)
    return subprocess.run([BIN, case, GOLDEN], capture_output=True, text=True, timeout=timeout,
                          cwd=str(cwd))


@pytest.mark.parametrize("case", ["missing3d", "missing2d"])
def test_missing_mesh_exits_with_status_1(case, tmp_path):
    r = _run(case, tmp_path)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "File open failed" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small3d", "basic3d", "converge3d", "improve3d", "sanity3d",
                                  "basic2d", "converge2d", "improve2d", "sanity2d",
                                  "seeds_speeds", "async_final", "vtk"])
def test_reference_case_through_facade(case, tmp_path):
    r = _run(case, tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
