"""The C++ component pair on the host: it builds, links the C ABI, and reports errors (no device
here: HIPAccelerator's pg_device_create fails with a message instead of aborting)."""
import os

import pytest

from piper2_amd import scenes
from plugin_io import DRIVER, run_driver, write_scene


def test_driver_reports_errors_without_aborting(tmp_path):
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    r = run_driver()
    assert r.returncode == 2 and "usage" in r.stderr
    r = run_driver(str(tmp_path / "missing.pgsc"), str(tmp_path / "out"))
    assert r.returncode == 3 and "cannot open" in r.stderr
    bad = tmp_path / "bad.pgsc"
    bad.write_bytes(b"XXXX\x01\x00\x00\x00")
    r = run_driver(str(bad), str(tmp_path / "out"))
    assert r.returncode == 3 and "not a scene file" in r.stderr
    sc = scenes.cornell(width=16, height=16, spp=1)
    path = tmp_path / "cornell.pgsc"
    write_scene(sc, path)
    r = run_driver(str(path), str(tmp_path / "out"))
    if r.returncode == 0:
        pytest.skip("a GPU is present: the device path is covered by test_gpu_plugin.py")
    assert r.returncode == 3 and "pg_device_create" in r.stderr, (r.returncode, r.stderr)
