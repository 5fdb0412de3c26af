"""The model.cpp-style C++ driver runs the substep through the drop-in
GPU_Implementation5 shim (icempm_2d_sp_amd/csrc/gpu_implementation5.hpp)."""
import os
import re
import subprocess

import pytest

from icempm_2d_sp_amd import build

pytestmark = pytest.mark.gpu


def test_model_driver_runs_frames():
    print("test_model_driver_runs_frames", flush=True)
    assert os.path.exists(build.DRIVER), "model_driver not built"
    res = subprocess.run([build.DRIVER, "128", "64", "3", "50", "0.1"], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    frames = re.findall(r"frame (\d+) step (\d+) .* err 0x0 .* mean vx ([-0-9.e+]+)", res.stdout)
    assert len(frames) == 3
    assert [int(f[1]) for f in frames] == [50, 100, 150]
    v = [float(f[2]) for f in frames]
    assert 0 < v[0] < v[1] < v[2]     # the wind keeps accelerating the sheet
    arrays = re.findall(r"frame \d+ arrays: ice nodes (\d+), mass on water nodes ([-0-9.e+]+)", res.stdout)
    assert len(arrays) == 3 and all(int(a[0]) > 0 and float(a[1]) > 0 for a in arrays)
    assert "model_driver done" in res.stdout
