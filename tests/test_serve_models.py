"""CPU checks of the serving path's concurrency (no device).

* The serving loop's host code (csrc/serve.hip, compiled for the host only with ThreadSanitizer)
  against a model device whose streams run their operations in order on threads of their own
  (tests/cpp/serve_host_model.cpp): requests submitted from several threads at random gaps and
  waited for from others, with launches in flight and a memory budget holding requests back.
  ThreadSanitizer must report no race, every request must come back with its own values, and
  the stream must finish.
* The persistent kernels' work-queue protocol with several launches in flight on one device
  (tests/cpp/queue_model.cpp): workgroup slots shared by launches whose grids each fill the
  device, partial residency, per-XCD in-order or interleaved dispatch, the next take issued by a
  second wave while the first queues dependents, the empty-queue check on entry.  Every launch
  must end, every entry be consumed, and every solve reach its fixed point.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"
CSRC = os.path.join(ROOT, "sci-solver_eikonal_amd", "csrc")
INC = os.path.join(ROOT, "include")


@pytest.fixture(scope="module")
def serve_model(tmp_path_factory):
    if not (os.path.exists(HIPCC) and os.path.exists(CLANG)):
        pytest.skip("hipcc is not installed")
    d = tmp_path_factory.mktemp("serve_model")
    cc = [HIPCC, "-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "--cuda-host-only",
          "-Xarch_host", "-fsanitize=thread", "-I", INC, "-I", CSRC]
    subprocess.run(cc + ["-c", os.path.join(CSRC, "serve.hip"), "-o", str(d / "serve.o")],
                   check=True, capture_output=True)
    subprocess.run(cc + ["-x", "hip", "-c", os.path.join(ROOT, "tests", "cpp", "serve_host_model.cpp"),
                         "-o", str(d / "model.o")], check=True, capture_output=True)
    exe = d / "serve_model"
    subprocess.run([CLANG, "-fsanitize=thread", str(d / "serve.o"), str(d / "model.o"), "-o", str(exe),
                    "-lpthread"], check=True, capture_output=True)
    return str(exe)


@pytest.mark.parametrize("args", [("1500", "4", "4", "32", "2"), ("800", "1", "4", "32", "3"),
                                  ("800", "6", "8", "4", "4"), ("400", "2", "1", "1", "5")])
def test_serving_loop_is_race_free_and_finishes(serve_model, args):
    """requests, submitting threads, launches in flight, solves per launch, seed"""
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([serve_model, *args], capture_output=True, text=True, timeout=240, env=env)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "failures 0" in r.stdout


@pytest.fixture(scope="module")
def queue_model(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no C++ compiler")
    d = tmp_path_factory.mktemp("queue_model")
    exe = d / "queue_model"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "queue_model.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    return str(exe)


@pytest.mark.parametrize("args", [("1500", "4", "11"), ("1500", "4", "12", "-interleave"),
                                  ("800", "8", "13", "-interleave"), ("800", "1", "14")])
def test_queue_protocol_with_launches_in_flight(queue_model, args):
    """schedules, most launches in flight, seed, dispatch order"""
    r = subprocess.run([queue_model, *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout
    assert "every launch ended" in r.stdout
