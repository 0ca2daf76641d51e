"""Milton scene files (mpt_scenefile_*): the tolerant JSON grammar, the scene
graph flattening (materials, transforms, shapes), the mapping of each element
to the scene description, one-sided triangles against the oracle, and clear
errors for what this path does not render.  CPU only."""
import math
import os

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from conftest import classify_hits, rays64

REFERENCE_SCENES = "/root/reference/mlt/src/exec/scenes"

# A small Cornell-style room in the Milton scenefile format: one-sided
# "plane" walls given by their vertices, a transformed unit plane as the
# light, a sphere, comments, trailing commas and repeated keys.
ROOM = """
{
   /* a test room */
   "scenefile" : {
      "version" : 0.1,
      "renderer" : { "type" : "pathTracer", "noSuperSamples" : 16, "maxDepth" : 5, "seed" : 3, },
      "camera" : {
         "type" : "pinhole", "eye" : [ 0.5, 0.5, -1.4 ], "look" : [ 0, 0, 1 ], "up" : [ 0, 1, 0 ],
         "heightAngle" : 40,
      },
      "output" : { "width" : 64, "height" : 48 },
      "scene" : {
         "material" : {
            "emitter" : { "type" : "oriented", "power" : [ 3, 3, 2.5 ] },
            "bsdf" : { "type" : "diffuse", "kd" : .75 },
            "transform" : {
               "translate" : [ 0.5, 0.999, 0.5 ],
               "rotate" : [ 0, 0, 0, 1, 0, 0, 180 ],
               "scale" : [ 0.25, 1, 0.25 ],
               "plane" : { },
            },
         },
         "material" : {
            "bsdf" : { "type" : "diffuse", "kd" : 0.75 },
            // floor, ceiling, back wall (front faces towards the inside)
            "plane" : { "vertices" : [ 0,0,0,  0,0,1,  1,0,1,  1,0,0 ] },
            "plane" : { "vertices" : [ 0,1,0,  1,1,0,  1,1,1,  0,1,1 ] },
            "plane" : { "vertices" : [ 0,0,1,  0,1,1,  1,1,1,  1,0,1 ] },
         },
         "material" : {
            "bsdf" : { "type" : "diffuse", "kd" : [ 0.6, 0.1, 0.1 ] },
            "plane" : { "vertices" : [ 0,0,0,  0,1,0,  0,1,1,  0,0,1 ] },
         },
         "material" : {
            "bsdf" : { "type" : "diffuse", "kd" : [ 0.1, 0.6, 0.1 ] },
            "plane" : { "vertices" : [ 1,0,0,  1,0,1,  1,1,1,  1,1,0 ] },
         },
         "material" : {
            "bsdf" : { "type" : "dielectric", "ks" : 1, "transparency" : 1 },
            "medium" : { "ior" : 1.5 },
            "sphere" : { "position" : [ 0.35, 0.2, 0.6 ], "radius" : 0.2 },
         },
         "material" : {
            "bsdf" : { "type" : "modifiedPhong", "kd" : 0.2, "ks" : 0.6, "n" : 50 },
            "transform" : {
               "translate" : [ 0.7, 0.15, 0.35 ],
               "scale" : [ 0.2, 0.3, 0.2 ],
               "triangle" : { "vertices" : [ -0.5,0,-0.5,  -0.5,0,0.5,  0.5,0,0.5 ] },
            },
         },
      },
   },
}
"""


def _room():
    return m.load_scenefile(text=ROOM)


def test_grammar_and_settings():
    a, job = _room()
    assert (job.width, job.height, job.spp, job.max_depth, job.seed) == (64, 48, 16, 5, 3)
    assert list(job.camera.eye) == pytest.approx([0.5, 0.5, -1.4])
    assert job.camera.height_angle == pytest.approx(40)
    assert job.camera.aspect == pytest.approx(1.0)   # the camera default, not W/H
    # 5 walls x 2 one-sided triangles + 1 standalone triangle; 1 quad light; 1 sphere
    assert len(a.indices) == 11 and len(a.quads) == 1 and len(a.spheres) == 1
    assert a.tri_flags is not None and np.all(a.tri_flags == m.MPT_TRI_ONE_SIDED)
    assert len(a.vertices) == 33   # standalone triangles keep their own vertices


def test_light_quad_and_emission():
    a, _ = _room()
    q = a.quads[0]
    c, ex, ez = q[:3], q[3:6], q[6:9]
    np.testing.assert_allclose(c, [0.5, 0.999, 0.5], atol=1e-6)
    # rotate 180 degrees about x: x stays, z flips; scale 0.25 in x and z
    np.testing.assert_allclose(ex, [0.25, 0, 0], atol=1e-6)
    np.testing.assert_allclose(ez, [0, 0, -0.25], atol=1e-6)
    mat = a.materials[a.quad_material[0]]
    area = 0.25 * 0.25
    np.testing.assert_allclose(mat["Le"], np.array([3, 3, 2.5]) / (math.pi * area), rtol=1e-6)


def test_materials_and_sphere():
    a, _ = _room()
    s = a.spheres[0]
    np.testing.assert_allclose(s, [0.35, 0.2, 0.6, 0.2], atol=1e-6)
    sm = a.materials[a.sphere_material[0]]
    assert sm["type"] == m.MPT_BSDF_DIELECTRIC and sm["ior"] == pytest.approx(1.5)
    assert sm["ior_g"] == 0 and sm["ior_b"] == 0


def test_dispersive_ior():
    a, _ = m.load_scenefile(text=ROOM.replace('"medium" : { "ior" : 1.5 }', '"medium" : { "ior" : [1.4, 1.55, 1.8] }'))
    sm = a.materials[a.sphere_material[0]]
    assert sm["ior"] == pytest.approx(1.4) and sm["ior_g"] == pytest.approx(1.55) and sm["ior_b"] == pytest.approx(1.8)
    assert sm["transparency"] == pytest.approx(1) and np.allclose(sm["ks"], 1)
    red = a.materials[a.tri_material[6]]
    np.testing.assert_allclose(red["kd"], [0.6, 0.1, 0.1], rtol=1e-6)
    tri = a.materials[a.tri_material[10]]
    assert tri["type"] == m.MPT_BSDF_PHONG and tri["exponent"] == pytest.approx(50)
    np.testing.assert_allclose(tri["kd"], 0.2, rtol=1e-6)
    np.testing.assert_allclose(tri["ks"], 0.6, rtol=1e-6)


def test_transform_order_and_rotation():
    """Node transform = translate * rotate * scale; children compose under
    their parents; rotate is right-handed about the given axis and point."""
    text = """{ "scenefile" : { "scene" : {
        "transform" : { "translate" : [1, 2, 3],
          "transform" : { "rotate" : [0, 0, 0,  0, 1, 0,  90], "scale" : [2, 1, 1],
             "plane" : { } } },
        "transform" : { "rotate" : [1, 0, 0,  0, 0, 1,  90],
             "triangle" : { "vertices" : [2,0,0, 2,1,0, 3,0,0] } } } } }"""
    a, _ = m.load_scenefile(text=text)
    c, ex, ez = a.quads[0][:3], a.quads[0][3:6], a.quads[0][6:9]
    np.testing.assert_allclose(c, [1, 2, 3], atol=1e-6)
    # scale x by 2, then rotate +90 about y: x -> -z, z -> x
    np.testing.assert_allclose(ex, [0, 0, -2], atol=1e-6)
    np.testing.assert_allclose(ez, [1, 0, 0], atol=1e-6)
    # +90 about z through (1, 0, 0): (2,0,0) -> (1,1,0), (2,1,0) -> (0,1,0), (3,0,0) -> (1,2,0)
    np.testing.assert_allclose(a.vertices[:3], [[1, 1, 0], [0, 1, 0], [1, 2, 0]], atol=1e-6)


def test_defaults():
    a, job = m.load_scenefile(text='{ "scenefile" : { "scene" : { "sphere" : { } } } }')
    assert (job.width, job.height, job.spp) == (480, 480, 1)
    assert list(job.camera.eye) == pytest.approx([2, 2, 2])
    np.testing.assert_allclose(np.array(job.camera.look) / np.linalg.norm(job.camera.look),
                               -np.ones(3) / math.sqrt(3), atol=1e-6)
    assert job.camera.height_angle == pytest.approx(45)
    np.testing.assert_allclose(a.spheres[0], [0, 0, 0, 0.5], atol=1e-7)   # unit sphere, diameter 1
    assert a.materials[0]["type"] == m.MPT_BSDF_DIFFUSE and np.allclose(a.materials[0]["kd"], 0.5)
    assert a.tonemap == m.MPT_TONEMAP_REINHARD


@pytest.mark.parametrize("kind, want", [("bidirectionalPathTracer", m.MPT_RENDERER_BIDIR),
                                        ("bidirPathTracer", m.MPT_RENDERER_BIDIR),
                                        ("pathTracer", m.MPT_RENDERER_PATH), (None, m.MPT_RENDERER_PATH)])
def test_renderer_type(kind, want):
    """MiltonJSONSceneLoader.cpp:720-721: both names select the bidirectional
    renderer; other types render with the path tracer."""
    r = '"renderer" : { "type" : "%s" }, ' % kind if kind else ""
    _, job = m.load_scenefile(text='{ "scenefile" : { %s"scene" : { "sphere" : { } } } }' % r)
    assert job.renderer == want


def test_mlt_renderer_keys():
    """MiltonJSONSceneLoader.cpp:731-738 and MLTRenderer::init / render
    (renderers/mlt/MLTRenderer.cpp:52-91): mlt selects Metropolis light
    transport with the reference's defaults, one chain per render thread,
    and mutations per chain = noSuperSamples x pixels / chains."""
    _, job = m.load_scenefile(text='{ "scenefile" : { "renderer" : { "type" : "mlt", "noRenderThreads" : 8, '
                                   '"noSuperSamples" : 16 }, "output" : { "width" : 32, "height" : 16 }, '
                                   '"scene" : { "sphere" : { } } } }')
    assert job.renderer == m.MPT_RENDERER_MLT
    assert (job.mlt_chains, job.mlt_mutations, job.mlt_seed_paths, job.mlt_max_depth, job.mlt_max_rejections,
            job.mlt_flags) == (8, 16 * 32 * 16 // 8, 10000, 10, 500, 0)
    assert list(job.mlt_probs) == [1.0, 1.0, 1.0]
    _, job = m.load_scenefile(text='{ "scenefile" : { "renderer" : { "type" : "MLT", "mltMaxDepth" : 7, '
                                   '"mltNoInitialPaths" : 100, "mltMaxConsequtiveRejections" : 0, '
                                   '"mltBidirPathMutationProb" : 0.5, "mltLensSubpathMutationProb" : 0, '
                                   '"mltPerturbationPathMutationProb" : 2, "mltFilterProposed" : false, '
                                   '"mltMutations" : 99 }, "scene" : { "sphere" : { } } } }')
    assert (job.mlt_chains, job.mlt_mutations, job.mlt_seed_paths, job.mlt_max_depth, job.mlt_max_rejections,
            job.mlt_flags) == (1, 99, 100, 7, 0, 1)
    assert list(job.mlt_probs) == [0.5, 0.0, 2.0]


@pytest.mark.parametrize("name, want", [("reinhard", m.MPT_TONEMAP_REINHARD), ("default", m.MPT_TONEMAP_REINHARD),
                                        ("null", m.MPT_TONEMAP_GAMMA)])
def test_output_tonemap(name, want):
    a, _ = m.load_scenefile(text='{ "scenefile" : { "output" : { "width" : 8, "height" : 8, "tonemap" : "%s" },'
                                 ' "scene" : { "sphere" : { } } } }' % name)
    assert a.tonemap == want


@pytest.mark.parametrize("snippet, word", [
    ('"cone" : { }', "cone"),
    ('"material" : { "emitter" : { "type" : "oriented", "power" : 1 }, "cube" : { } }', "cubes"),
    ('"transform" : { "arbitraryTransform" : [1,0.5,0,0, 0,1,0,0, 0,0,1,0, 0,0,0,1], "cube" : { } }', "sheared"),
    ('"material" : { "emitter" : { "type" : "oriented", "power" : 1 }, "triangle" : { } }', "triangles"),
    ('"material" : { "emitter" : { "type" : "omni", "power" : 1 }, "plane" : { } }', "omni"),
    ('"transform" : { "scale" : [1, 2, 1], "sphere" : { } }', "ellipsoid"),
    ('"material" : { "bsdf" : { "type" : "velvet" }, "plane" : { } }', "velvet"),
    ('"mesh" : { "path" : "no/such/file.obj" }', "cannot be opened"),
])
def test_unsupported_elements_fail_clearly(snippet, word):
    with pytest.raises(m.MptError) as e:
        m.load_scenefile(text='{ "scenefile" : { "scene" : { %s } } }' % snippet)
    assert word in str(e.value)


def test_syntax_error_names_the_line():
    with pytest.raises(m.MptError) as e:
        m.load_scenefile(text='{ "scenefile" : {\n "scene" : { "plane" : { } \n } ')
    assert "line" in str(e.value)


def test_obj_and_ply_meshes(tmp_path):
    obj = tmp_path / "quad.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1 4//1\n")
    ply = tmp_path / "tri.ply"
    ply.write_text("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                   "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
                   "0 0 0\n2 0 0\n0 4 0\n3 0 1 2\n")
    scene = tmp_path / "s.js"
    scene.write_text('{ "scenefile" : { "scene" : { "mesh" : { "path" : "quad.obj" }, '
                     '"mesh" : { "path" : "tri.ply", "normalize" : true } } } }')
    a, _ = m.load_scenefile(path=str(scene))
    assert len(a.indices) == 3                      # quad as a fan of 2, plus 1
    assert a.tri_flags is not None and np.all(a.tri_flags == 0)   # mesh triangles are two-sided
    assert a.vertex_normal_mask is not None
    np.testing.assert_array_equal(a.vertex_normal_mask, [1, 1, 1, 1, 0, 0, 0])
    np.testing.assert_allclose(a.vertex_normals[:4], [[0, 0, 1]] * 4)
    # normalize: into [-0.5, 0.5] per axis (a flat axis keeps its coordinate centred)
    np.testing.assert_allclose(a.vertices[4:], [[-0.5, -0.5, 0], [0.5, -0.5, 0], [-0.5, 0.5, 0]], atol=1e-6)


def test_one_sided_triangles_against_oracle():
    """The device traversal (host build of it) and the oracle agree on the
    loaded room, whose walls cull their back faces; rays from outside the
    room pass through the walls' backs."""
    a, job = _room()
    rng = np.random.default_rng(5)
    n = 20000
    inside = rng.uniform(0.02, 0.98, (n // 2, 3))
    outside = rng.uniform(-1.5, 2.5, (n // 2, 3))
    o = np.concatenate([inside, outside]).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, m.RAY_DTYPE)
    rays["o"] = o
    rays["d"] = d.astype(np.float32)
    tree = m.HostKdTree(a)
    hits, _ = tree.trace_closest(rays)
    osc = po.OracleScene(a)
    ot, oi = osc.trace_closest(rays64(rays))
    match, non_tie, rel = classify_hits(osc, rays, hits["t"], hits["prim"], ot, oi)
    assert non_tie <= 1e-4, (match, non_tie)
    assert (rel <= 1e-5).mean() >= 0.9999
    # a ray that meets the floor from below misses it; from above, hits it
    up = np.zeros(2, m.RAY_DTYPE)
    up["o"] = [[0.3, -0.5, 0.3], [0.3, 0.5, 0.3]]
    up["d"] = [[0, 1, 0], [0, -1, 0]]
    h, _ = tree.trace_closest(up)
    ot2, oi2 = osc.trace_closest(rays64(up))
    assert h["prim"][1] in (0, 1) and oi2[1] in (0, 1)           # floor triangles
    assert h["prim"][0] not in (0, 1) and oi2[0] not in (0, 1)   # culled from below


def test_oracle_one_sided_closed_form():
    """The oracle's culling test follows the reference's (e1 . (d x e2) must
    reach EPSILON): a triangle is hit from its front at the analytic t and
    missed from its back; the same triangle two-sided is hit from both."""
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)   # front normal +z
    mats = np.zeros(1, m.MATERIAL_DTYPE)
    mats["kd"] = 0.5
    rays = np.array([[0.2, 0.2, 2, 0, 0, -1], [0.2, 0.2, -2, 0, 0, 1]], np.float64)
    # e1 = (1,0,0), e2 = (0,1,0): det = e1 . (d x e2) = -d_z, so the ray going -z hits
    for flags, expect in (([1], [2.0, np.inf]), ([0], [2.0, 2.0])):
        a = m.SceneArrays(verts, [[0, 1, 2]], [0], np.zeros((0, 4)), [], np.zeros((0, 9)), [], mats, [0, 0, 0],
                          tri_flags=flags)
        t, _ = po.OracleScene(a).trace_closest(rays)
        np.testing.assert_allclose(t, expect)


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SCENES), reason="reference scene files not present")
def test_reference_scene_files():
    """Every scene file Milton ships either loads or is refused with a
    message naming what is not supported; the Cornell box loads."""
    loaded, refused = [], []
    for name in sorted(os.listdir(REFERENCE_SCENES)):
        if not name.endswith(".js"):
            continue
        try:
            a, job = m.load_scenefile(path=os.path.join(REFERENCE_SCENES, name))
            loaded.append((name, a.num_prims))
        except m.MptError as e:
            msg = str(e)
            assert any(w in msg for w in ("not supported", "cannot be opened", "cannot be read", "syntax", "needs",
                                          "no supported")), msg
            refused.append((name, msg))
    print("loaded:", loaded)
    print("refused:", refused)
    names = [n for n, _ in loaded]
    assert "cornellBox.js" in names
    a, job = m.load_scenefile(path=os.path.join(REFERENCE_SCENES, "cornellBox.js"))
    assert len(a.quads) == 1 and len(a.indices) == 2 * 15   # light; 15 vertex planes of two triangles
    assert job.direct_samples == 4 and job.spp == 1024
    le = a.materials[a.quad_material[0]]["Le"]
    np.testing.assert_allclose(le, np.array([2e6, 2e6, 1.6e6]) / (math.pi * 140 * 125), rtol=1e-5)
    assert list(job.camera.eye) == pytest.approx([278, 273, -650])
    # plates.js: its lights are emitting spheres (area pi, Sphere.cpp:138-141) and planes
    a, _ = m.load_scenefile(path=os.path.join(REFERENCE_SCENES, "plates.js"))
    emitting = [a.materials[s]["Le"].max() > 0 for s in a.sphere_material]
    assert sum(emitting) == 2 and a.sphere_frames is not None


def _quad_normals(a):
    q = a.quads.reshape(-1, 9).astype(np.float64)
    n = np.cross(q[:, 6:9], q[:, 3:6])
    return q, n / np.linalg.norm(n, axis=1, keepdims=True)


@pytest.mark.parametrize("scale", [[2, 1, 0.5], [-1, 1, 1], [1, -0.5, 3]])
def test_cube_faces(scale):
    """Cube::getIntersection (shapes/Cube.cpp:17-52): six faces of the unit
    cube under the node's transform, each a quad whose normal points out of
    the cube (also under a mirroring scale), with face areas the products of
    the two scales along it."""
    text = ('{ "scenefile" : { "scene" : { "transform" : { "translate" : [1, 2, 3], "rotate" : [0, 0, 0, 0, 1, 0, 30],'
            ' "scale" : %s, "cube" : { } } } } }' % scale)
    a, _ = m.load_scenefile(text=text)
    assert a.quads.reshape(-1, 9).shape == (6, 9) and len(a.indices) == 0 and len(a.spheres) == 0
    q, n = _quad_normals(a)
    centre = np.array([1, 2, 3], float)
    assert (np.einsum("ij,ij->i", n, q[:, :3] - centre) > 0).all()
    sx, sy, sz = np.abs(scale)
    areas = np.linalg.norm(np.cross(q[:, 6:9], q[:, 3:6]), axis=1)
    assert sorted(np.round(areas, 6)) == sorted(np.round([sy * sz] * 2 + [sx * sz] * 2 + [sx * sy] * 2, 6))


def test_mirrored_plane_normal():
    """A plane under a scale that flips y keeps the normal the inverse
    transpose gives it: (0, -1, 0)."""
    a, _ = m.load_scenefile(text='{ "scenefile" : { "scene" : { "transform" : { "scale" : [1, -1, 1],'
                                 ' "plane" : { } } } } }')
    _, n = _quad_normals(a)
    assert np.allclose(n[0], [0, -1, 0])


def test_cube_hits_closed_form():
    """Axis-aligned cube [0, 2] x [1, 2] x [-1, 1]: rays from outside along
    each axis hit the near face at the distance to it, a ray from inside hits
    the far face, a ray passing beside the cube misses."""
    a, _ = m.load_scenefile(text='{ "scenefile" : { "scene" : { "transform" : { "translate" : [1, 1.5, 0],'
                                 ' "scale" : [2, 1, 2], "cube" : { } } } } }')
    osc = po.OracleScene(a)
    rays = np.array([[-3, 1.5, 0, 1, 0, 0],     # x = 0 face at t = 3
                     [1, 5, 0.5, 0, -1, 0],     # y = 2 face at t = 3
                     [0.5, 1.2, 4, 0, 0, -1],   # z = 1 face at t = 3
                     [1, 1.5, 0, 0, 0, 1],      # inside: z = 1 face at t = 1
                     [1, 1.5, 0, 0.8, 0, 0.6],  # inside: x = 2 face at t = 1.25 (z = 1 at 5/3)
                     [-3, 3, 0, 1, 0, 0]],      # above the cube: miss
                    float)
    t, prim = osc.trace_closest(rays)
    assert np.allclose(t[:5], [3, 3, 3, 1, 1.25], rtol=1e-12)
    assert (prim[:5] != 0xFFFFFFFF).all() and prim[5] == 0xFFFFFFFF and np.isinf(t[5])


def test_emitting_spheres_and_meshes(tmp_path):
    """An emitting sphere gets Le = P / (pi * pi) (Sphere::_getSurfaceArea is
    pi whatever the scale, shapes/Sphere.cpp:138-141) and its transform's
    linear part as its frame; an emitting mesh gets Le = P / (pi A) with A
    its world-space area (Mesh::_getSurfaceArea, shapes/Mesh.cpp:452-466),
    and its triangles share one shape id."""
    obj = tmp_path / "quad.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    scene = tmp_path / "s.js"
    scene.write_text('{ "scenefile" : { "scene" : { '
                     '"material" : { "emitter" : { "type" : "oriented", "power" : [3, 2, 1] }, '
                     '"transform" : { "translate" : [0, 2, 0], "rotate" : [0, 0, 0, 1, 0, 0, 90], '
                     '"scale" : [0.4, 0.4, 0.4], "sphere" : { } } }, '
                     '"material" : { "emitter" : { "type" : "oriented", "power" : 5 }, '
                     '"transform" : { "scale" : [2, 3, 1], "mesh" : { "path" : "quad.obj" } } }, '
                     '"plane" : { } } } }')
    a, _ = m.load_scenefile(path=str(scene))
    assert len(a.spheres) == 1 and len(a.indices) == 2
    le = a.materials[a.sphere_material[0]]["Le"]
    np.testing.assert_allclose(le, np.array([3, 2, 1]) / math.pi ** 2, rtol=1e-6)
    np.testing.assert_allclose(a.spheres[0], [0, 2, 0, 0.2], atol=1e-6)
    # the object axes after a 90 degree turn about x, scaled by 0.4
    np.testing.assert_allclose(a.sphere_frames[0].reshape(3, 3), [[0.4, 0, 0], [0, 0, 0.4], [0, -0.4, 0]],
                               atol=1e-6)
    assert a.tri_shape[0] == a.tri_shape[1]
    le = a.materials[a.tri_material[0]]["Le"]
    np.testing.assert_allclose(le, 5 / (math.pi * 6.0), rtol=1e-6)   # the quad scaled to 2 x 3
