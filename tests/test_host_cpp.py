"""The C++ host layer (libyafaray_amd/host/ymx_host.hpp), through its C++ test
driver tests/cpp/test_host.cpp.

The driver builds scenes with the reference's param maps (accelerator type,
materials, lights, camera, integrator), runs the single-ray host queries, the
batched GPU queries and the wavefront path tracer, and checks each against the
CPU oracle bit for bit (closest hit, shadow, transparent-shadow casters, the
path-traced film).
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "bin" / "test_host"


def _run(*args, timeout=300):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (run __graft_entry__.build())")
    res = subprocess.run([str(BIN), *args], capture_output=True, text=True, timeout=timeout)
    print(res.stdout)
    print(res.stderr)
    assert res.returncode == 0, f"test_host {' '.join(args)} failed:\n{res.stdout}\n{res.stderr}"
    assert "ALL PASSED" in res.stdout


def test_host_layer_param_maps_and_errors():
    # unknown accelerator / material / light types and missing objects are
    # refused with a message; param reads keep their types and defaults
    _run("--cpu")


@pytest.mark.gpu
def test_host_layer_matches_oracle():
    _run()
