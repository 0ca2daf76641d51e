"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point declared in include/*.h, the struct layouts the Python
binding uses match the header, and the seeded scene generators are
deterministic and have the specified shapes."""
import ctypes
import os
import re
import subprocess

import numpy as np

import milton_amd as m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(mpt_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = m.lib()
    names = _declared("milton_pt.h") + _declared("milton_pt_scenes.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.mpt_abi_version() == m.MPT_ABI_VERSION


_C_STRUCTS = {"Material": "mpt_material", "SceneDesc": "mpt_scene_desc", "KdParams": "mpt_kd_params",
              "Camera": "mpt_camera", "Job": "mpt_job", "Ray": "mpt_ray", "RenderStats": "mpt_render_stats",
              "ServerConfig": "mpt_server_config", "JobStatus": "mpt_job_status", "TuneResult": "mpt_tune_result",
              "EnvMapDesc": "mpt_env_map"}


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler
    gives the header's struct."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "milton_pt.h"', "int main(void) {"]
    for py, c in _C_STRUCTS.items():
        lines.append(f'  printf("{py} size %zu\\n", sizeof({c}));')
        for f, _ in getattr(m, py)._fields_:
            lines.append(f'  printf("{py} {f} %zu\\n", offsetof({c}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, got):
        py, field, val = line.split()
        cls = getattr(m, py)
        want = ctypes.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert int(val) == want, (py, field, int(val), want)


def test_struct_layouts():
    assert ctypes.sizeof(m.Material) == 64
    assert ctypes.sizeof(m.Camera) == 44 + 24
    assert ctypes.sizeof(m.Job) == 44 + 24 + 28 + 16 + 8 + 4 + 36   # + renderer, Metropolis fields
    assert ctypes.sizeof(m.Ray) == 24
    assert m.HIT_DTYPE.itemsize == 8
    assert ctypes.sizeof(m.KdParams) == 32
    # field order of mpt_scene_desc
    offs = [getattr(m.SceneDesc, f).offset for f, _ in m.SceneDesc._fields_]
    assert offs == sorted(offs)


def test_open_without_device_fails_cleanly():
    """mpt_open on a host with no visible device returns an error code
    instead of crashing (skipped when a device is present)."""
    L = m.lib()
    h = ctypes.c_void_p()
    rc = L.mpt_open(999, ctypes.byref(h))
    assert rc == -2 and not h.value


def test_room_generator():
    a = m.generate(m.MPT_SCENE_ROOM)
    assert a.indices.shape == (34, 3) and a.quads.shape == (1, 9)
    assert len(a.materials) == 4 and a.materials["Le"][3].tolist() == [15, 15, 15]
    # wall normals point into the room, box normals out of the boxes
    v = a.vertices[a.indices]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    c = v.mean(axis=1)
    walls = slice(0, 10)
    assert np.all(np.sum(n[walls] * (np.array([0.5, 0.5, 0.5]) - c[walls]), axis=1) > 0)
    boxes = [(slice(10, 22), [0.3, 0.3, 0.6]), (slice(22, 34), [0.7, 0.15, 0.35])]
    for sl, centre in boxes:
        assert np.all(np.sum(n[sl] * (c[sl] - np.array(centre)), axis=1) > 0)


def test_soup_generator_deterministic_and_lognormal():
    a = m.generate(m.MPT_SCENE_SOUP, 20000, 1)
    b = m.generate(m.MPT_SCENE_SOUP, 20000, 1)
    c = m.generate(m.MPT_SCENE_SOUP, 20000, 2)
    assert np.array_equal(a.vertices, b.vertices)
    assert not np.array_equal(a.vertices, c.vertices)
    v = a.vertices.reshape(-1, 3, 3).astype(np.float64)
    centroid = v.mean(axis=1)
    assert centroid.min() > -0.05 and centroid.max() < 1.05
    # offsets from the generating centre have lognormal length: the vertex
    # spread r satisfies ln r ~ N(ln 0.005 + const, ...); check the median edge
    e = np.linalg.norm(v[:, 1] - v[:, 0], axis=1)
    assert 0.004 < np.median(e) < 0.009
    assert len(a.indices) == 20000 and a.env.tolist() == [1, 1, 1]


def test_terrain_generator_shape():
    a = m.generate(m.MPT_SCENE_TERRAIN, 64, 3)
    assert len(a.indices) == 2 * 63 * 63
    assert a.spheres.shape == (500, 4)
    r = a.spheres[:, 3]
    assert r.min() >= 0.01 and r.max() <= 0.05
    assert set(np.unique(a.tri_material)) <= {0, 1}
    frac_phong = (a.tri_material == 1).mean()
    assert 0.15 < frac_phong < 0.3
    q = a.quads[0]
    n = np.cross(q[6:9], q[3:6])
    assert n[1] < 0   # the distant light faces down
    side = np.linalg.norm(q[3:6])
    assert abs(2 * np.degrees(np.arctan(side / 2 / np.linalg.norm(q[:3] - [2.0, 0, 2.0]))) - 0.5) < 0.01
    # spheres cover about 10% of the 4 x 4 terrain
    cover = np.pi * (r.astype(np.float64) ** 2).sum() / 16.0
    assert 0.05 < cover < 0.15


def test_generated_rays():
    r1 = m.generate_rays(m.MPT_SCENE_SOUP, 1000, 1000, seed=7)
    r2 = m.generate_rays(m.MPT_SCENE_SOUP, 1000, 1000, seed=7)
    assert np.array_equal(r1, r2)
    np.testing.assert_allclose(np.linalg.norm(r1["d"], axis=1), 1, rtol=1e-6)
    assert np.all((r1["o"][500:] >= 0) & (r1["o"][500:] <= 1))


def test_plugin_loads_like_milton():
    """The renderer plugin exports plugin_type / plugin_new / plugin_delete
    and reports MILTON_PLUGIN_TYPE_RENDERER (dynamic/DynamicPlugin.h:39-58),
    loaded as the scenefile loader does (MiltonJSONSceneLoader.cpp:1770-1813)."""
    import subprocess
    drv = os.path.join(ROOT, "milton_amd", "mpt_plugin_driver")
    so = os.path.join(ROOT, "milton_amd", "libmilton_pt_plugin.so")
    out = subprocess.run([drv, so, "--load-only"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "type 0" in out.stdout
    lib = ctypes.CDLL(so)
    for sym in ("plugin_type", "plugin_new", "plugin_delete"):
        assert hasattr(lib, sym)
    bad = subprocess.run([drv, os.path.join(ROOT, "milton_amd", "missing.so"), "--load-only"],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "failed to load" in bad.stderr


def test_batch_capacity_keeps_shadow_rays_in_budget():
    """A batch reserves one 32-byte shadow ray per light sample of each path;
    the engine lowers the batch so that paths x shadow rays per vertex stays
    within 2^28 (32-bit indices and 8 GiB per lane), never below 1024
    paths, and never above the context's batch size."""
    f = m.lib().mpt_batch_paths_for
    assert f(1 << 25, 0, 0, 1) == 1 << 25          # no lights
    assert f(1 << 25, 1, 0, 1) == 1 << 25          # one shadow ray per vertex
    assert f(1 << 25, 0, 1, 64) == 1 << 25         # a point light draws one sample
    assert f(1 << 25, 1, 0, 0) == 1 << 25          # 0 light samples means 1
    for lights, points, direct in [(1, 0, 4), (1, 0, 64), (3, 1, 64), (2, 2, 64), (3, 2, 84)]:
        spv = lights * direct + points
        for bp in (1 << 20, 1 << 25, 1 << 28):
            cap = f(bp, lights, points, direct)
            assert 1024 <= cap <= bp
            assert cap * spv <= 1 << 28 or cap == 1024
            assert cap * spv < 1 << 32
            if bp * spv <= 1 << 28:
                assert cap == bp
    # the largest shadow-ray count per vertex the engine accepts (254)
    assert f(1 << 28, 254, 0, 1) * 254 <= 1 << 28
