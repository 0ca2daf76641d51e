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
    assert "MLT identical to C ABI: yes" in out.stdout
    assert "missing array rejected: yes" in out.stdout


def test_milton_adapter_renders_like_the_c_abi(tmp_path):
    """The Milton-side adapter, driven as Milton's loader and application
    drive a dynamic renderer (setOutput, setScene, setCamera, keys, init,
    render), writes the same film as mpt_render of the room, bit for bit."""
    from test_adapter import build_adapter_host
    exe = build_adapter_host(tmp_path)
    print("adapter host: --gpu", flush=True)
    out = subprocess.run([exe, "--gpu"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "values identical to the C ABI" in out.stdout
