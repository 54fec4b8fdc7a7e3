"""The GFM program's file contract (GFM_GPU.cu:24-34, TFM.cpp:90-94) on the CPU:
every malformed input is rejected before the GPU is touched, with a non-zero
exit status and no G.dat left behind."""
import os
import subprocess

import numpy as np
import pytest

from ztm_amd import gfm

EXE = os.path.join(os.path.dirname(gfm.LIB_PATH), "bin", "GFM")


def run(d, **env):
    e = dict(os.environ, ZTM_GFM_DIR=str(d), **{k: str(v) for k, v in env.items()})
    return subprocess.run([EXE], env=e, capture_output=True, text=True, timeout=60)


@pytest.fixture
def shm(tmp_path):
    # a stale result from an earlier call must never survive a failed call
    (tmp_path / "G.dat").write_bytes(b"\0" * 16)
    return tmp_path


def test_missing_inputs(shm):
    p = run(shm)
    assert p.returncode == 3 and "R.dat" in p.stderr
    assert not (shm / "G.dat").exists()


def test_station_file_size_checked(shm):
    np.zeros(61 * 61 * 65, np.float32).tofile(shm / "R.dat")
    np.zeros(61 * 61 - 1, np.float32).tofile(shm / "S.dat")
    p = run(shm)
    assert p.returncode == 3 and "S.dat" in p.stderr
    assert not (shm / "G.dat").exists()


def test_density_file_not_whole_layers(shm):
    np.zeros(61 * 61 * 65 + 7, np.float32).tofile(shm / "R.dat")
    np.zeros(61 * 61, np.float32).tofile(shm / "S.dat")
    p = run(shm)
    assert p.returncode == 3 and "multiple of ii*jj" in p.stderr


def test_density_layers_must_match_fv(shm):
    # 66 layers cannot be (dk-1)*2+1
    np.zeros(61 * 61 * 66, np.float32).tofile(shm / "R.dat")
    np.zeros(61 * 61, np.float32).tofile(shm / "S.dat")
    p = run(shm)
    assert p.returncode == 3 and "kk" in p.stderr


def test_odd_byte_count(shm):
    (shm / "R.dat").write_bytes(b"\0" * 10)
    np.zeros(61 * 61, np.float32).tofile(shm / "S.dat")
    p = run(shm)
    assert p.returncode == 3 and "whole number" in p.stderr


def test_bad_geometry_env(shm):
    p = run(shm, ZTM_GFM_DI="abc")
    assert p.returncode == 2


def test_no_device_fails_loudly(shm):
    """Valid files but no usable GPU (none in this process's view): exit status
    4 (ZTM_ENODEV), the reason on stderr, and no G.dat. TFM ignores the status
    (TFM.cpp:93-94), so the missing file is what makes the failure visible."""
    np.zeros(61 * 61 * 65, np.float32).tofile(shm / "R.dat")
    np.zeros(61 * 61, np.float32).tofile(shm / "S.dat")
    p = run(shm, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    assert p.returncode == 4, (p.returncode, p.stderr)
    assert "GFM" in p.stderr
    assert not (shm / "G.dat").exists()
