"""The LDS exchange position maps of the FFT passes (cc_device.hpp, xswz),
checked on the host from the header itself: linear bijections of [0, N) for
every shape and exchange, and conflict-free FP64 Stockham writes under the
gfx950 write model (8 lanes of 16 bytes per LDS cycle, 32 banks)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not installed")
def test_exchange_maps(tmp_path):
    exe = tmp_path / "swizzle_check"
    src = os.path.join(HERE, "native", "swizzle_check.hip")
    res = subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-host-only", "-o", str(exe), src],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    lines = [l for l in out.stdout.splitlines() if "NS=" in l]
    assert len(lines) >= 30 and all("linear bijection" in l and l.endswith(": 0") for l in lines), out.stdout
