"""The register march's dispatch order (edge tiles first, then interior tiles) is a bijection onto
the launch's tiles for many grids, group mixes and tile shapes: a host-only build of
tests/native/tile_order_check.cpp against the library's own header checks every block."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_edge_first_dispatch_is_a_bijection(tmp_path):
    exe = tmp_path / "tile_order_check"
    subprocess.run(["hipcc", "--offload-host-only", "-O1", "-std=c++17", "-I", str(ROOT / "his_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "tile_order_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
