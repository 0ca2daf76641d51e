"""The Milton-side adapter (milton_amd/adapter/milton_amd_renderer.cpp)
compiles against the subset of Milton's interface it uses
(tests/adapter/milton.h), and its converters turn the benchmark room, built
as Milton objects (Meshes, an oriented-emitter Plane, a PinholeCamera), into
the room's own scene description: the same triangles, materials, light and
camera, and the same closest hits through the host build of the device tree.
Without a device, init() fails cleanly at mpt_open.  The render itself is
checked on the device in tests/test_gpu_plugin.py."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_adapter_host(tmp_path):
    exe = str(tmp_path / "adapter_host")
    lib = os.path.join(ROOT, "milton_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "tests", "adapter"),
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "adapter", "host.cpp"),
                    "-o", exe, "-L", lib, "-lmilton_pt_plugin", "-lmilton_pt", f"-Wl,-rpath,{lib}"], check=True)
    return exe


def test_adapter_flattens_the_room(tmp_path):
    exe = build_adapter_host(tmp_path)
    out = subprocess.run([exe, "--cpu"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "20000 host rays agree" in out.stdout
    assert "init failed cleanly" in out.stdout
