"""The Milton renderer plugin renders on the device, and its film is
bit-identical to the same job through the C ABI; a scene description with a
missing array is rejected at init instead of being read through."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_plugin_render_matches_c_abi():
    drv = os.path.join(ROOT, "milton_amd", "mpt_plugin_driver")
    so = os.path.join(ROOT, "milton_amd", "libmilton_pt_plugin.so")
    out = subprocess.run([drv, so], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "identical to C ABI: yes" in out.stdout
    assert "missing array rejected: yes" in out.stdout
