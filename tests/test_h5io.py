"""CPU tests of the HDF5 formats (libicempm_io.so): snapshots
(SnapshotManager::SaveSnapshot / ReadPointsFromSnapshot,
snapshotmanager.cpp:308-396) and the static current field
(WindAndCurrentInterpolator::ReadStaticFlowData,
windandcurrentinterpolator.cpp:18-63).

The layout is checked with HDF5's own h5dump (an independent reader): dataset
names, shapes, element types, chunking, deflate level, attribute names and
types, and values.  The reference's own input file flow_velocities.h5 is read
when the reference tree is present (its sub-region against h5dump's values)."""
import os
import re
import subprocess

import numpy as np
import pytest

from icempm_2d_sp_amd import abi, h5io

H5DUMP = os.path.join(os.environ.get("ICE_HDF5_ROOT", "/opt/conda"), "bin", "h5dump")
REF_FLOW = "/root/reference/input/n2k/flow_velocities.h5"

pytestmark = pytest.mark.skipif(not os.path.exists(h5io.LIB_PATH), reason="libicempm_io.so not built (no HDF5)")


def h5dump(*args) -> str:
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


def dumped_values(text: str) -> np.ndarray:
    body = text[text.index("DATA {"):]
    body = body[:body.index("}")]
    vals = []
    for line in body.splitlines():
        if line.strip().startswith("("):
            vals += [float(v) for v in re.findall(r"[-+0-9.eEnaif]+(?=,|$)", line.split(":", 1)[1])]
    return np.array(vals)


def random_soa(n, lines=20, seed=0):
    rng = np.random.default_rng(seed)
    soa = rng.standard_normal((lines, n)).astype(np.float32)
    u = soa.view(np.uint32)
    u[abi.IDX_UTILITY] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)   # any bit pattern
    u[abi.IDX_CELL] = rng.integers(0, 4096, n) | (rng.integers(0, 2048, n) << 16)
    u[abi.IDX_POINT] = rng.permutation(n)
    return soa


@pytest.mark.parametrize("precision", [32, 64])
def test_snapshot_round_trip_bits(tmp_path, precision):
    soa = random_soa(3000)
    path = str(tmp_path / "s.h5")
    h5io.snapshot_save(path, soa, 123, 4.5, precision)
    info = h5io.snapshot_info(path)
    assert info == {"n": 3000, "step": 123, "time": 4.5, "precision": precision, "cohesion": True}
    back, step, t = h5io.snapshot_load(path, abi.N_LINES_MAX)
    assert (step, t) == (123, 4.5)
    assert np.array_equal(back.view(np.uint32), soa.view(np.uint32))   # every line, integer lines bit for bit
    back19, _, _ = h5io.snapshot_load(path, abi.N_LINES)
    assert np.array_equal(back19.view(np.uint32), soa[:19].view(np.uint32))


def test_empty_snapshot_round_trip(tmp_path):
    path = str(tmp_path / "e.h5")
    h5io.snapshot_save(path, np.zeros((20, 0), np.float32), 3, 0.5, 32)
    assert h5io.snapshot_info(path)["n"] == 0
    soa, step, t = h5io.snapshot_load(path, abi.N_LINES_MAX)
    assert soa.shape == (20, 0) and (step, t) == (3, 0.5)


def test_snapshot_layout_matches_the_reference(tmp_path):
    soa = random_soa(300000, seed=1)
    path = str(tmp_path / "s.h5")
    h5io.snapshot_save(path, soa, 7, 0.25, 32)
    hdr = h5dump("-p", "-H", path)
    pts = hdr[hdr.index('DATASET "pts_data"'):]
    assert "H5T_IEEE_F32LE" in pts.split("DATASPACE")[0]
    assert "SIMPLE { ( 19, 300000 ) / ( 19, 300000 ) }" in pts
    assert "CHUNKED ( 19, 262144 )" in pts                       # min(256 Ki, n) points per chunk
    assert "COMPRESSION DEFLATE { LEVEL 5 }" in pts
    for name, typ in [("SimulationStep", "H5T_STD_I32LE"), ("SimulationTime", "H5T_IEEE_F64LE"),
                      ("HSSOA_size", "H5T_STD_U32LE"), ("nPtsArrays", "H5T_STD_I32LE")]:
        block = pts[pts.index(f'ATTRIBUTE "{name}"'):]
        assert typ in block.split("}")[0], name
    att = h5dump("-a", "/pts_data/nPtsArrays", path)
    assert dumped_values(att)[0] == 19
    sub = h5dump("-d", "/pts_data", "-s", "6,100", "-c", "2,3", "-m", "%.9g", path)
    assert np.array_equal(dumped_values(sub).astype(np.float32), soa[6:8, 100:103].reshape(-1))


def test_double_snapshot_stores_integers_in_low_bits(tmp_path):
    """t_PointReal = double: numeric lines as doubles, integer lines as the low
    32 bits of the double's bit pattern (proxypoint2d.cpp:74-89)."""
    soa = random_soa(50, lines=19, seed=2)
    path = str(tmp_path / "d.h5")
    h5io.snapshot_save(path, soa, 1, 0.0, 64)
    assert "H5T_IEEE_F64LE" in h5dump("-H", "-d", "/pts_data", path)
    vals = dumped_values(h5dump("-d", "/pts_data", "-s", "8,0", "-c", "1,5", "-m", "%.17g", path))
    assert np.array_equal(vals, soa[8, :5].astype(np.float64))
    raw = h5dump("-d", "/pts_data", "-s", "2,0", "-c", "1,5", "-m", "%.17g", path)
    as_double = dumped_values(raw)
    low = np.frombuffer(as_double.tobytes(), np.uint64) & 0xFFFFFFFF
    assert np.array_equal(low.astype(np.uint32), soa[2, :5].view(np.uint32))


def test_snapshot_errors(tmp_path):
    flow = str(tmp_path / "f.h5")
    h5io.write_static_flow(flow, np.zeros((3, 4, 2), np.float32))
    with pytest.raises(h5io.IoError):
        h5io.snapshot_info(flow)                               # no pts_data
    with pytest.raises(h5io.IoError):
        h5io.snapshot_info(str(tmp_path / "missing.h5"))


def test_static_flow_mapping(tmp_path):
    """Grid node (i, j) takes pixel (i + ox, j + oy) of the [height][width][2] field."""
    h, w = 7, 9
    v = np.arange(h * w * 2, dtype=np.float32).reshape(h, w, 2)
    path = str(tmp_path / "flow.h5")
    h5io.write_static_flow(path, v)
    assert "SIMPLE { ( 7, 9, 2 ) / ( 7, 9, 2 ) }" in h5dump("-H", path)
    gx, gy, ox, oy = 5, 4, 2, 1
    cur = h5io.read_static_flow(path, w, h, ox, oy, gx, gy)
    for i in range(gx):
        for j in range(gy):
            assert cur[0, i * gy + j] == v[j + oy, i + ox, 0] and cur[1, i * gy + j] == v[j + oy, i + ox, 1]
    with pytest.raises(h5io.IoError):
        h5io.read_static_flow(path, w + 1, h, 0, 0, 2, 2)      # dimension mismatch
    with pytest.raises(h5io.IoError):
        h5io.read_static_flow(path, w, h, 6, 0, 5, 2)          # grid outside the image


@pytest.mark.skipif(not os.path.exists(REF_FLOW), reason="the reference's input files are not present")
def test_reference_flow_file():
    """The reference's own input (1323 x 2000 pixels): a sub-region read through
    the grid mapping equals h5dump's values of the same pixels."""
    gx, gy, ox, oy = 6, 4, 821, 445   # a sub-region with flow
    cur = h5io.read_static_flow(REF_FLOW, 2000, 1323, ox, oy, gx, gy)
    vals = dumped_values(h5dump("-d", "/velocities", "-s", f"{oy},{ox},0", "-c", f"{gy},{gx},2", "-m", "%.9g",
                                REF_FLOW)).reshape(gy, gx, 2)
    for i in range(gx):
        for j in range(gy):
            assert cur[0, i * gy + j] == np.float32(vals[j, i, 0]) and cur[1, i * gy + j] == np.float32(vals[j, i, 1])
    assert np.all(np.abs(cur[0]) > 0)
