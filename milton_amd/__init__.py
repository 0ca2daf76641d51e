"""Python binding of the MI355X path-tracing core (ctypes over include/milton_pt.h).

The native library milton_amd/libmilton_pt.so holds the HIP kernels, the kd-tree
builder and the C ABI.  This module loads it and mirrors the reference's
objects on the path: a Scene built from primitives and materials
(core/Scene.cpp), kd-tree BuildParams (accel/kdTreeAccel.h:68-116) and a
PathTracer renderer with init()/render() (renderers/renderers/PathTracer.cpp,
renderers/PointSampleRenderer.cpp).  Importing never falls back to another
implementation: without the built library every entry point raises.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MPT_LIBRARY names another build of the same library (for example one
# compiled with different kernel tunables); the in-tree build by default.
LIB_PATH = os.environ.get("MPT_LIBRARY") or os.path.join(_HERE, "libmilton_pt.so")

MPT_ABI_VERSION = 8
MPT_TRI_ONE_SIDED = 1
MPT_MISS = 0xFFFFFFFF
MPT_BSDF_DIFFUSE, MPT_BSDF_PHONG, MPT_BSDF_DIELECTRIC = 0, 1, 2
MPT_SCENE_ROOM, MPT_SCENE_SOUP, MPT_SCENE_TERRAIN = 1, 2, 3
MPT_OPT_BATCH_PATHS, MPT_OPT_TIMING, MPT_OPT_COUNTERS, MPT_OPT_GRAPHS, MPT_OPT_STACK, MPT_OPT_PERSISTENT = 1, 2, 3, 4, 5, 6
MPT_OPT_LANES = 7
MPT_OPT_FAIL_BATCH = 8
MPT_OPT_TREE_TOP = 9
MPT_OPT_TRACE_BLOCK, MPT_OPT_SHADE_BLOCK = 10, 11
MPT_TONEMAP_REINHARD, MPT_TONEMAP_GAMMA = 0, 1
MPT_RENDERER_PATH, MPT_RENDERER_BIDIR, MPT_RENDERER_MLT = 0, 1, 2
MPT_FILTER_BOX, MPT_FILTER_TRIANGLE, MPT_FILTER_GAUSSIAN, MPT_FILTER_MITCHELL, MPT_FILTER_LANCZOS = 0, 1, 2, 3, 4
MPT_SAVE_RADIANCE = 1
MPT_E_IO = -7
ERRORS = {-1: "invalid", -2: "device", -3: "nomem", -4: "budget", -5: "notfound", -6: "cancelled"}

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("kd", ctypes.c_float * 3), ("ks", ctypes.c_float * 3),
                ("exponent", ctypes.c_float), ("ior", ctypes.c_float), ("transparency", ctypes.c_float),
                ("Le", ctypes.c_float * 3), ("ior_g", ctypes.c_float), ("ior_b", ctypes.c_float),
                ("pad", ctypes.c_uint32)]


class EnvMapDesc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("rgb", _f32p),
                ("repeat_u", ctypes.c_float), ("repeat_v", ctypes.c_float), ("filter", ctypes.c_uint32),
                ("support", ctypes.c_uint32), ("params", ctypes.c_float * 2)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint32),
                ("num_vertices", ctypes.c_uint32), ("vertices", _f32p),
                ("num_triangles", ctypes.c_uint32), ("indices", _u32p), ("tri_material", _u32p),
                ("num_spheres", ctypes.c_uint32), ("spheres", _f32p), ("sphere_material", _u32p),
                ("num_quads", ctypes.c_uint32), ("quads", _f32p), ("quad_material", _u32p),
                ("num_materials", ctypes.c_uint32), ("materials", ctypes.POINTER(Material)),
                ("env_radiance", ctypes.c_float * 3), ("tri_flags", _u32p), ("vertex_normals", _f32p),
                ("vertex_normal_mask", _u8p), ("env_map", ctypes.POINTER(EnvMapDesc)),
                ("num_points", ctypes.c_uint32), ("points", _f32p), ("tri_shape", _u32p),
                ("sphere_frames", _f32p)]


class KdParams(ctypes.Structure):
    """kdTreeAccel::BuildParams (accel/kdTreeAccel.h:68-116)."""
    _fields_ = [("leaf_size", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("cost_traversal", ctypes.c_float), ("cost_intersect", ctypes.c_float),
                ("empty_bias", ctypes.c_float), ("num_threads", ctypes.c_uint32), ("bins", ctypes.c_uint32),
                ("no_clip", ctypes.c_uint32)]


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("look", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("height_angle", ctypes.c_float), ("aspect", ctypes.c_float), ("aperture", ctypes.c_float),
                ("focal_length", ctypes.c_float), ("focal_distance", ctypes.c_float), ("autofocus", ctypes.c_uint32),
                ("focal_point", ctypes.c_float * 2)]


class Job(ctypes.Structure):
    _fields_ = [("camera", Camera), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("spp", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("job_id", ctypes.c_uint32), ("direct_samples", ctypes.c_uint32),
                ("filter", ctypes.c_uint32), ("filter_support", ctypes.c_uint32), ("filter_params", ctypes.c_float * 2),
                ("direct_off", ctypes.c_uint32), ("light_select", ctypes.c_uint32), ("renderer", ctypes.c_uint32),
                ("mlt_chains", ctypes.c_uint32), ("mlt_mutations", ctypes.c_uint32),
                ("mlt_seed_paths", ctypes.c_uint32), ("mlt_max_depth", ctypes.c_uint32),
                ("mlt_max_rejections", ctypes.c_uint32), ("mlt_flags", ctypes.c_uint32),
                ("mlt_probs", ctypes.c_float * 3)]


EMITTER_POINT_DTYPE = np.dtype([("p", np.float32, 3), ("light", np.uint32), ("n", np.float32, 3),
                                ("pdf_area", np.float32), ("pdf_select", np.float32), ("pad", np.uint32)])


class Ray(ctypes.Structure):
    _fields_ = [("o", ctypes.c_float * 3), ("d", ctypes.c_float * 3)]


class RenderStats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("camera_rays", ctypes.c_uint64),
                ("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("render_ms", ctypes.c_double), ("extend_ms", ctypes.c_double),
                ("shade_ms", ctypes.c_double), ("shadow_ms", ctypes.c_double),
                ("extend_launches", ctypes.c_uint64), ("nodes_visited", ctypes.c_uint64),
                ("prims_tested", ctypes.c_uint64), ("extend_wall_ms", ctypes.c_double),
                ("shade_wall_ms", ctypes.c_double), ("shadow_wall_ms", ctypes.c_double),
                ("mlt_accepted", ctypes.c_uint64), ("mlt_samples", ctypes.c_uint64), ("mlt_b", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_active", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32)]


class TuneResult(ctypes.Structure):
    _fields_ = [("batch_paths", ctypes.c_uint64), ("stack", ctypes.c_uint32), ("persistent", ctypes.c_uint32),
                ("leaf_size", ctypes.c_uint32), ("trials", ctypes.c_uint32), ("paths_per_s", ctypes.c_double),
                ("baseline_paths_per_s", ctypes.c_double), ("trace_block", ctypes.c_uint32),
                ("shade_block", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32), ("error", ctypes.c_uint32), ("pixels_done", ctypes.c_uint64),
                ("pixels_total", ctypes.c_uint64), ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_FAILED = 0, 1, 2, 3, 4

MATERIAL_DTYPE = np.dtype([("type", "<u4"), ("kd", "<f4", 3), ("ks", "<f4", 3), ("exponent", "<f4"),
                           ("ior", "<f4"), ("transparency", "<f4"), ("Le", "<f4", 3), ("ior_g", "<f4"),
                           ("ior_b", "<f4"), ("pad", "<u4")])
assert MATERIAL_DTYPE.itemsize == ctypes.sizeof(Material) == 64
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3)])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4")])

_lib = None


class MptError(RuntimeError):
    pass


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MptError(f"{LIB_PATH} is missing: build it with `python -m milton_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    if L.mpt_abi_version() != MPT_ABI_VERSION:
        raise MptError(f"{LIB_PATH} has ABI {L.mpt_abi_version()}, this binding {MPT_ABI_VERSION}: rebuild it")
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "mpt_abi_version": ([], i32),
        "mpt_batch_paths_for": ([u64, u32, u32, u32], u64),
        "mpt_open": ([i32, ctypes.POINTER(vp)], i32),
        "mpt_close": ([vp], i32),
        "mpt_last_error": ([vp], ctypes.c_char_p),
        "mpt_sync": ([vp], i32),
        "mpt_set_option": ([vp, u32, u64], i32),
        "mpt_get_option": ([vp, u32, ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_scene_create": ([vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(KdParams), ctypes.POINTER(vp)], i32),
        "mpt_scene_destroy": ([vp], i32),
        "mpt_scene_stats": ([vp, ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_trace_closest": ([vp, vp, u64, vp], i32),
        "mpt_trace_any": ([vp, vp, vp, u64, vp], i32),
        "mpt_emitter_sample": ([vp, u64, u32, vp], i32),
        "mpt_trace_closest_stats": ([vp, vp, u64, vp, ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_render": ([vp, ctypes.POINTER(Job), vp, ctypes.POINTER(RenderStats)], i32),
        "mpt_film_create": ([vp, u32, u32, ctypes.POINTER(vp)], i32),
        "mpt_film_destroy": ([vp], i32),
        "mpt_render_film": ([vp, ctypes.POINTER(Job), vp, ctypes.POINTER(RenderStats)], i32),
        "mpt_render_film_range": ([vp, ctypes.POINTER(Job), u32, u32, vp, ctypes.POINTER(RenderStats)], i32),
        "mpt_film_download": ([vp, vp], i32),
        "mpt_film_download_raw": ([vp, vp], i32),
        "mpt_film_upload_raw": ([vp, vp], i32),
        "mpt_tonemap_from_name": ([ctypes.c_char_p], u32),
        "mpt_film_tonemap": ([vp, u32, vp], i32),
        "mpt_film_tonemap_noalloc": ([vp, u32, vp], i32),
        "mpt_film_save": ([vp, ctypes.c_char_p, u32, u32], i32),
        "mpt_image_write_ldr": ([ctypes.c_char_p, u32, u32, vp], i32),
        "mpt_image_write_hdr": ([ctypes.c_char_p, u32, u32, vp], i32),
        "mpt_image_read": ([ctypes.c_char_p, ctypes.POINTER(u32), ctypes.POINTER(u32), vp, ctypes.c_size_t], i32),
        "mpt_render_paths": ([vp, ctypes.POINTER(Job), u32, u32, vp], i32),
        "mpt_mlt_trace": ([vp, ctypes.POINTER(Job), u32, vp, ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_server_create": ([vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)], i32),
        "mpt_server_destroy": ([vp], i32),
        "mpt_submit": ([vp, vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_status": ([vp, u64, ctypes.POINTER(JobStatus)], i32),
        "mpt_cancel": ([vp, u64], i32),
        "mpt_result": ([vp, u64, vp], i32),
        "mpt_wait": ([vp, u64, u32], i32),
        "mpt_server_step": ([vp, ctypes.POINTER(ctypes.c_uint32)], i32),
        "mpt_server_drain": ([vp], i32),
        "mpt_server_start": ([vp], i32),
        "mpt_server_stop": ([vp], i32),
        "mpt_server_stats": ([vp, ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_server_memory": ([vp, ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_tune": ([vp, ctypes.POINTER(Job), u64, ctypes.POINTER(TuneResult)], i32),
        "mpt_tune_scene": ([vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(KdParams), ctypes.POINTER(Job), u64,
                            ctypes.POINTER(vp), ctypes.POINTER(TuneResult)], i32),
        "mpt_scenefile_load": ([ctypes.c_char_p, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t], i32),
        "mpt_scenefile_parse": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.c_char_p,
                                 ctypes.c_size_t], i32),
        "mpt_scenefile_desc": ([vp], ctypes.POINTER(SceneDesc)),
        "mpt_scenefile_job": ([vp, ctypes.POINTER(Job)], i32),
        "mpt_scenefile_tonemap": ([vp], u32),
        "mpt_scenefile_destroy": ([vp], i32),
        "mpt_kd_build": ([ctypes.POINTER(SceneDesc), ctypes.POINTER(KdParams), ctypes.POINTER(vp)], i32),
        "mpt_kd_stats": ([vp, ctypes.POINTER(ctypes.c_uint64)], i32),
        "mpt_kd_trace": ([vp, vp, u64, vp, ctypes.POINTER(ctypes.c_uint64), u32], i32),
        "mpt_kd_destroy": ([vp], i32),
        "mpt_gen_create": ([u32, u32, u32, ctypes.POINTER(vp)], i32),
        "mpt_gen_describe": ([vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(Camera)], i32),
        "mpt_gen_destroy": ([vp], i32),
        "mpt_gen_rays": ([vp, u64, u32, u32, u32, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------------------
# Scenes as numpy arrays (shared by the device path and the oracle).
# ---------------------------------------------------------------------------
class EnvMap:
    """A lat-long environment image for the background (mpt_env_map): image
    (H, W, 3) float32 from the top row, repeats, and the kernel filter its
    lookups go through (MPT_FILTER_*, support, params; zeros = defaults)."""

    def __init__(self, image, repeat_u=1.0, repeat_v=1.0, filter=MPT_FILTER_TRIANGLE, support=2, params=(0.0, 0.0)):
        self.image = np.ascontiguousarray(image, np.float32).reshape(np.shape(image)[0], -1, 3)
        self.repeat_u, self.repeat_v = float(repeat_u), float(repeat_v)
        self.filter, self.support = int(filter), int(support)
        self.params = (float(params[0]), float(params[1]))

    def desc(self):
        e = EnvMapDesc()
        e.height, e.width = self.image.shape[:2]
        e.rgb = self.image.ctypes.data_as(_f32p)
        e.repeat_u, e.repeat_v = self.repeat_u, self.repeat_v
        e.filter, e.support = self.filter, self.support
        e.params[0], e.params[1] = self.params
        return e


class SceneArrays:
    """Primitive and material arrays in the layout of mpt_scene_desc."""

    def __init__(self, vertices, indices, tri_material, spheres, sphere_material, quads, quad_material,
                 materials, env, camera=None, tri_flags=None, vertex_normals=None, vertex_normal_mask=None,
                 env_map=None, points=None, tri_shape=None, sphere_frames=None):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        self.tri_material = np.ascontiguousarray(tri_material, np.uint32).reshape(-1)
        self.spheres = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4)
        self.sphere_material = np.ascontiguousarray(sphere_material, np.uint32).reshape(-1)
        self.quads = np.ascontiguousarray(quads, np.float32).reshape(-1, 9)
        self.quad_material = np.ascontiguousarray(quad_material, np.uint32).reshape(-1)
        self.materials = np.ascontiguousarray(materials, MATERIAL_DTYPE).reshape(-1)
        self.env = np.ascontiguousarray(env, np.float32).reshape(3)
        self.camera = camera
        # optional: MPT_TRI_ONE_SIDED per triangle; a mesh file's vertex normals
        self.tri_flags = None if tri_flags is None else np.ascontiguousarray(tri_flags, np.uint32).reshape(-1)
        self.vertex_normals = (None if vertex_normals is None
                               else np.ascontiguousarray(vertex_normals, np.float32).reshape(-1, 3))
        self.vertex_normal_mask = (None if vertex_normal_mask is None
                                   else np.ascontiguousarray(vertex_normal_mask, np.uint8).reshape(-1))
        self.env_map = env_map   # optional EnvMap in place of the constant env
        # optional point lights: (x, y, z, power r, g, b) per light
        self.points = (np.zeros((0, 6), np.float32) if points is None
                       else np.ascontiguousarray(points, np.float32).reshape(-1, 6))
        # optional: the emitting shape each triangle belongs to; per sphere the
        # linear part of its transform (columns: images of the object axes)
        self.tri_shape = None if tri_shape is None else np.ascontiguousarray(tri_shape, np.uint32).reshape(-1)
        self.sphere_frames = (None if sphere_frames is None
                              else np.ascontiguousarray(sphere_frames, np.float32).reshape(-1, 9))

    @property
    def num_prims(self):
        return len(self.indices) + len(self.spheres) + len(self.quads)

    def desc(self):
        d = SceneDesc()
        d.version = MPT_ABI_VERSION
        d.num_vertices = len(self.vertices)
        d.vertices = self.vertices.ctypes.data_as(_f32p)
        d.num_triangles = len(self.indices)
        d.indices = self.indices.ctypes.data_as(_u32p)
        d.tri_material = self.tri_material.ctypes.data_as(_u32p)
        d.num_spheres = len(self.spheres)
        d.spheres = self.spheres.ctypes.data_as(_f32p)
        d.sphere_material = self.sphere_material.ctypes.data_as(_u32p)
        d.num_quads = len(self.quads)
        d.quads = self.quads.ctypes.data_as(_f32p)
        d.quad_material = self.quad_material.ctypes.data_as(_u32p)
        d.num_materials = len(self.materials)
        d.materials = self.materials.ctypes.data_as(ctypes.POINTER(Material))
        for k in range(3):
            d.env_radiance[k] = float(self.env[k])
        if self.tri_flags is not None:
            d.tri_flags = self.tri_flags.ctypes.data_as(_u32p)
        if self.vertex_normals is not None:
            d.vertex_normals = self.vertex_normals.ctypes.data_as(_f32p)
            d.vertex_normal_mask = self.vertex_normal_mask.ctypes.data_as(_u8p)
        if self.env_map is not None:
            d._env = self.env_map.desc()   # kept alive with the descriptor
            d.env_map = ctypes.pointer(d._env)
        d.num_points = len(self.points)
        d.points = self.points.ctypes.data_as(_f32p)
        if getattr(self, "tri_shape", None) is not None:
            d.tri_shape = self.tri_shape.ctypes.data_as(_u32p)
        if getattr(self, "sphere_frames", None) is not None:
            d.sphere_frames = self.sphere_frames.ctypes.data_as(_f32p)
        return d

    def materials_f32(self):
        """Materials as the oracle's float[16] rows."""
        m = self.materials
        out = np.zeros((len(m), 16), np.float32)
        out[:, 0] = m["type"]
        out[:, 1:4] = m["kd"]
        out[:, 4:7] = m["ks"]
        out[:, 7] = m["exponent"]
        out[:, 8] = m["ior"]
        out[:, 9] = m["transparency"]
        out[:, 10:13] = m["Le"]
        out[:, 13] = m["ior_g"]
        out[:, 14] = m["ior_b"]
        return out


def generate(kind, count=0, seed=1):
    """Seeded benchmark scene (MPT_SCENE_*) as SceneArrays with its camera."""
    L = lib()
    g = ctypes.c_void_p()
    rc = L.mpt_gen_create(kind, count, seed, ctypes.byref(g))
    if rc:
        raise MptError(f"mpt_gen_create failed: {ERRORS.get(rc, rc)}")
    try:
        d, cam = SceneDesc(), Camera()
        L.mpt_gen_describe(g, ctypes.byref(d), ctypes.byref(cam))

        def arr(p, n, dt, w):
            if n == 0:
                return np.zeros((0, w), dt)
            return np.ctypeslib.as_array(p, shape=(n * w,)).astype(dt).reshape(n, w).copy()

        mats = np.frombuffer(ctypes.string_at(d.materials, 64 * d.num_materials), MATERIAL_DTYPE).copy()
        sa = SceneArrays(arr(d.vertices, d.num_vertices, np.float32, 3),
                         arr(d.indices, d.num_triangles, np.uint32, 3),
                         arr(d.tri_material, d.num_triangles, np.uint32, 1),
                         arr(d.spheres, d.num_spheres, np.float32, 4),
                         arr(d.sphere_material, d.num_spheres, np.uint32, 1),
                         arr(d.quads, d.num_quads, np.float32, 9),
                         arr(d.quad_material, d.num_quads, np.uint32, 1),
                         mats, np.array(d.env_radiance[:], np.float32),
                         camera=camera_dict(cam))
        sa._gen = None
        return sa
    finally:
        L.mpt_gen_destroy(g)


def generate_rays(kind, count, n, seed=7, film=(1920, 1080), scene_seed=1):
    """Seeded parity rays for a generated scene (half camera rays, half random)."""
    L = lib()
    g = ctypes.c_void_p()
    rc = L.mpt_gen_create(kind, count, scene_seed, ctypes.byref(g))
    if rc:
        raise MptError(f"mpt_gen_create failed: {ERRORS.get(rc, rc)}")
    try:
        rays = np.zeros(n, RAY_DTYPE)
        rc = L.mpt_gen_rays(g, n, seed, film[0], film[1], _ptr(rays))
        if rc:
            raise MptError("mpt_gen_rays failed")
        return rays
    finally:
        L.mpt_gen_destroy(g)


def camera_dict(cam):
    return {"eye": list(cam.eye), "look": list(cam.look), "up": list(cam.up),
            "height_angle": cam.height_angle, "aspect": cam.aspect}


def make_job(camera, width, height, spp, max_depth, seed=1, job_id=0, direct_samples=1, film_filter=None,
             renderer=0, mlt=None):
    """film_filter: None (pixel bin) or (MPT_FILTER_*, support, (p0, p1)); renderer:
    MPT_RENDERER_PATH, MPT_RENDERER_BIDIR or MPT_RENDERER_MLT.  mlt: the
    Metropolis parameters over the reference's defaults: chains (1),
    mutations (spp * pixels / chains), seed_paths (10000), max_depth (10),
    max_rejections (500), filter_proposed (True), probs ((1, 1, 1))."""
    j = Job()
    for k in range(3):
        j.camera.eye[k] = camera["eye"][k]
        j.camera.look[k] = camera["look"][k]
        j.camera.up[k] = camera["up"][k]
    j.camera.height_angle = camera["height_angle"]
    j.camera.aspect = camera["aspect"]
    # thin lens (optional): aperture and focal length in metres, focal distance,
    # autofocus through focal_point
    j.camera.aperture = camera.get("aperture", 0.0)
    j.camera.focal_length = camera.get("focal_length", 0.0)
    j.camera.focal_distance = camera.get("focal_distance", 1.0)
    j.camera.autofocus = 1 if camera.get("autofocus", False) else 0
    fp = camera.get("focal_point", (0.5, 0.5))
    j.camera.focal_point[0], j.camera.focal_point[1] = fp[0], fp[1]
    j.width, j.height, j.spp, j.max_depth, j.seed, j.job_id = width, height, spp, max_depth, seed, job_id
    j.direct_samples = direct_samples
    j.renderer = renderer
    p = dict(mlt or {})
    j.mlt_chains = p.get("chains", 1)
    j.mlt_mutations = p.get("mutations", max(1, spp * width * height // max(1, j.mlt_chains)))
    j.mlt_seed_paths = p.get("seed_paths", 10000)
    j.mlt_max_depth = p.get("max_depth", 10)
    j.mlt_max_rejections = p.get("max_rejections", 500)
    j.mlt_flags = 0 if p.get("filter_proposed", True) else 1
    probs = p.get("probs", (1.0, 1.0, 1.0))
    for k in range(3):
        j.mlt_probs[k] = probs[k]
    if film_filter is not None:
        ff = list(film_filter)
        kind = ff[0]
        support = ff[1] if len(ff) > 1 else 2
        params = ff[2] if len(ff) > 2 else (0.0, 0.0)
        j.filter, j.filter_support = 1 + kind, support
        j.filter_params[0], j.filter_params[1] = params
    return j


# ---------------------------------------------------------------------------
# Device objects.
# ---------------------------------------------------------------------------
class Context:
    """A HIP device context (mpt_open)."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = ctypes.c_void_p()
        self._children = weakref.WeakSet()   # scenes and films, closed before the context
        rc = self.L.mpt_open(device, ctypes.byref(self.h))
        if rc:
            raise MptError(f"mpt_open({device}) failed: {ERRORS.get(rc, rc)}")

    def check(self, rc, what):
        if rc:
            msg = self.L.mpt_last_error(self.h)
            raise MptError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg.decode() if msg else ''}")

    def get_option(self, option):
        v = ctypes.c_uint64()
        self.check(self.L.mpt_get_option(self.h, option, ctypes.byref(v)), "mpt_get_option")
        return v.value

    def set_option(self, option, value):
        self.check(self.L.mpt_set_option(self.h, option, value), "mpt_set_option")

    def sync(self):
        self.check(self.L.mpt_sync(self.h), "mpt_sync")

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            self.L.mpt_close(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A device-resident scene with its kd-tree (Scene::init + kdTreeAccel::init)."""

    def __init__(self, ctx, arrays, kd=None, _handle=None):
        self.ctx, self.arrays = ctx, arrays
        self.h = ctypes.c_void_p()
        if _handle is not None:
            self.h = _handle
        else:
            d = arrays.desc()
            kp = kd if kd is not None else KdParams()
            ctx.check(ctx.L.mpt_scene_create(ctx.h, ctypes.byref(d), ctypes.byref(kp), ctypes.byref(self.h)),
                      "mpt_scene_create")
        ctx._children.add(self)

    @classmethod
    def tuned(cls, ctx, arrays, job, sample_paths=1 << 26, kd=None):
        """Builds the scene at the fastest kd leaf size for `job` and leaves the
        fastest engine settings on ctx (mpt_tune_scene); returns (scene, result)."""
        d = arrays.desc()
        kp = kd if kd is not None else KdParams()
        h = ctypes.c_void_p()
        res = TuneResult()
        ctx.check(ctx.L.mpt_tune_scene(ctx.h, ctypes.byref(d), ctypes.byref(kp), ctypes.byref(job), sample_paths,
                                       ctypes.byref(h), ctypes.byref(res)), "mpt_tune_scene")
        return cls(ctx, arrays, _handle=h), res.as_dict()

    def tune(self, job, sample_paths=1 << 26):
        """Engine settings for this scene and job (mpt_tune), left on the context."""
        res = TuneResult()
        self.ctx.check(self.ctx.L.mpt_tune(self.h, ctypes.byref(job), sample_paths, ctypes.byref(res)), "mpt_tune")
        return res.as_dict()

    def stats(self):
        out = (ctypes.c_uint64 * 8)()
        self.ctx.L.mpt_scene_stats(self.h, out)
        keys = ["nodes", "leaves", "refs", "max_depth", "device_bytes", "build_ms", "max_leaf", "triangles"]
        return dict(zip(keys, list(out)))

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        self.ctx.check(self.ctx.L.mpt_trace_closest(self.h, _ptr(rays), len(rays), _ptr(hits)),
                       "mpt_trace_closest")
        return hits

    def trace_closest_stats(self, rays):
        """Closest hits through the render's closest-hit kernel with its work
        counters: (hits, {"nodes": kd nodes fetched, "refs": leaf references
        tested})."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        ctr = (ctypes.c_uint64 * 2)()
        self.ctx.check(self.ctx.L.mpt_trace_closest_stats(self.h, _ptr(rays), len(rays), _ptr(hits), ctr),
                       "mpt_trace_closest_stats")
        return hits, {"nodes": ctr[0], "refs": ctr[1]}

    def trace_any(self, rays, tmax):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        tmax = np.ascontiguousarray(tmax, np.float32)
        occ = np.zeros(len(rays), np.uint8)
        self.ctx.check(self.ctx.L.mpt_trace_any(self.h, _ptr(rays), _ptr(tmax), len(rays), _ptr(occ)),
                       "mpt_trace_any")
        return occ

    def emitter_sample(self, n, seed):
        """EmitterSampler draws on the device (mpt_emitter_sample)."""
        out = np.zeros(n, EMITTER_POINT_DTYPE)
        self.ctx.check(self.ctx.L.mpt_emitter_sample(self.h, n, seed, _ptr(out)), "mpt_emitter_sample")
        return out

    def render(self, job):
        film = np.zeros((job.height, job.width, 3), np.float32)
        st = RenderStats()
        self.ctx.check(self.ctx.L.mpt_render(self.h, ctypes.byref(job), _ptr(film), ctypes.byref(st)), "mpt_render")
        return film, st.as_dict()

    def mlt_trace(self, job, steps):
        """The first `steps` mutations of every chain of a Metropolis job
        (mpt_mlt_trace): (chains, steps, 6) records (mutation, alpha,
        accepted, mutated path length, tentative film bin, path length
        after), b and the seed entries."""
        rec = np.zeros((job.mlt_chains, steps, 6), np.float64)
        b = ctypes.c_double(0)
        n = ctypes.c_uint64(0)
        self.ctx.check(self.ctx.L.mpt_mlt_trace(self.h, ctypes.byref(job), steps, _ptr(rec), ctypes.byref(b),
                                                ctypes.byref(n)), "mpt_mlt_trace")
        return rec, b.value, n.value

    def render_paths(self, job, pix_begin, pix_end):
        L = np.zeros(((pix_end - pix_begin) * job.spp, 3), np.float32)
        self.ctx.check(self.ctx.L.mpt_render_paths(self.h, ctypes.byref(job), pix_begin, pix_end, _ptr(L)),
                       "mpt_render_paths")
        return L

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.mpt_scene_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Film:
    """Device-resident film (numerator/denominator per pixel)."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, width, height
        self.h = ctypes.c_void_p()
        ctx.check(ctx.L.mpt_film_create(ctx.h, width, height, ctypes.byref(self.h)), "mpt_film_create")
        ctx._children.add(self)

    def render(self, scene, job):
        st = RenderStats()
        self.ctx.check(self.ctx.L.mpt_render_film(scene.h, ctypes.byref(job), self.h, ctypes.byref(st)),
                       "mpt_render_film")
        return st.as_dict()

    def render_range(self, scene, job, pix_begin, pix_end):
        """Only pixels [pix_begin, pix_end) of the job (a tile)."""
        st = RenderStats()
        self.ctx.check(self.ctx.L.mpt_render_film_range(scene.h, ctypes.byref(job), pix_begin, pix_end, self.h,
                                                        ctypes.byref(st)), "mpt_render_film_range")
        return st.as_dict()

    def download(self):
        out = np.zeros((self.height, self.width, 3), np.float32)
        self.ctx.check(self.ctx.L.mpt_film_download(self.h, _ptr(out)), "mpt_film_download")
        return out

    def download_raw(self):
        """(H, W, 4) float32: numerator RGB and denominator per pixel."""
        out = np.zeros((self.height, self.width, 4), np.float32)
        self.ctx.check(self.ctx.L.mpt_film_download_raw(self.h, _ptr(out)), "mpt_film_download_raw")
        return out

    def upload_raw(self, rgbw):
        a = np.ascontiguousarray(rgbw, np.float32)
        if a.shape != (self.height, self.width, 4):
            raise ValueError(f"expected shape {(self.height, self.width, 4)}, got {a.shape}")
        self.ctx.check(self.ctx.L.mpt_film_upload_raw(self.h, _ptr(a)), "mpt_film_upload_raw")

    def tonemap(self, tonemap=MPT_TONEMAP_REINHARD):
        """(H, W, 4) uint8: the finalized image (r, g, b, 255); tonemap is
        MPT_TONEMAP_* or a reference tone-map name ("reinhard", "default",
        anything else = gamma only)."""
        if isinstance(tonemap, str):
            tonemap = tonemap_from_name(tonemap)
        out = np.zeros((self.height, self.width, 4), np.uint8)
        self.ctx.check(self.ctx.L.mpt_film_tonemap(self.h, tonemap, _ptr(out)), "mpt_film_tonemap")
        return out

    def save(self, path, tonemap=MPT_TONEMAP_REINHARD, radiance=False):
        """Writes .png / .ppm (tone-mapped) or .hdr / .rgbe / .pfm (tone-mapped
        bytes / 255, or the linear film with radiance=True)."""
        if isinstance(tonemap, str):
            tonemap = tonemap_from_name(tonemap)
        self.ctx.check(self.ctx.L.mpt_film_save(self.h, os.fsencode(path), tonemap,
                                                MPT_SAVE_RADIANCE if radiance else 0), "mpt_film_save")

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.mpt_film_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tonemap_from_name(name):
    """ToneMap::create's choice: "reinhard" and "default" -> Reinhard, else gamma."""
    return lib().mpt_tonemap_from_name(name.encode())


def write_image(path, image):
    """Writes a host image: (H, W, 4|3) uint8 to .png/.ppm/.hdr/.rgbe/.pfm, or
    (H, W, 3) float32 radiance to .hdr/.rgbe/.pfm."""
    a = np.asarray(image)
    h, w = a.shape[:2]
    if a.dtype == np.uint8:
        if a.shape[2] == 3:
            a = np.concatenate([a, np.full((h, w, 1), 255, np.uint8)], axis=2)
        a = np.ascontiguousarray(a)
        rc = lib().mpt_image_write_ldr(os.fsencode(path), w, h, _ptr(a))
    else:
        a = np.ascontiguousarray(a[..., :3], np.float32)
        rc = lib().mpt_image_write_hdr(os.fsencode(path), w, h, _ptr(a))
    if rc:
        raise MptError(f"cannot write {path} (error {rc})")


def read_image(path):
    """(H, W, 3) float32 linear RGB of a .hdr / .rgbe / .pfm / .ppm file."""
    L = lib()
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    rc = L.mpt_image_read(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), None, 0)
    if rc != -1:
        raise MptError(f"cannot read {path} (error {rc})")
    out = np.zeros((h.value, w.value, 3), np.float32)
    rc = L.mpt_image_read(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), _ptr(out), out.size)
    if rc:
        raise MptError(f"cannot read {path} (error {rc})")
    return out


class Server:
    """Continuous-batching render server over resident scenes (mpt_server_*)."""

    def __init__(self, ctx, hbm_budget=0, max_active=0, max_inflight=0):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        cfg = ServerConfig(hbm_budget, max_active, max_inflight)
        ctx.check(ctx.L.mpt_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(self.h)), "mpt_server_create")
        ctx._children.add(self)

    def submit(self, scene, job):
        """Returns a handle; raises MptError (budget) if the job can never be admitted."""
        h = ctypes.c_uint64()
        self.ctx.check(self.ctx.L.mpt_submit(self.h, scene.h, ctypes.byref(job), ctypes.byref(h)), "mpt_submit")
        return h.value

    def status(self, handle):
        st = JobStatus()
        self.ctx.check(self.ctx.L.mpt_status(self.h, handle, ctypes.byref(st)), "mpt_status")
        return st.as_dict()

    def cancel(self, handle):
        self.ctx.check(self.ctx.L.mpt_cancel(self.h, handle), "mpt_cancel")

    def result(self, handle, width, height):
        out = np.zeros((height, width, 3), np.float32)
        self.ctx.check(self.ctx.L.mpt_result(self.h, handle, _ptr(out)), "mpt_result")
        return out

    def wait(self, handle, timeout_ms=600000):
        rc = self.ctx.L.mpt_wait(self.h, handle, timeout_ms)
        if rc < 0:
            self.ctx.check(rc, "mpt_wait")
        return rc == 0

    def step(self):
        n = ctypes.c_uint32()
        self.ctx.check(self.ctx.L.mpt_server_step(self.h, ctypes.byref(n)), "mpt_server_step")
        return n.value

    def drain(self):
        self.ctx.check(self.ctx.L.mpt_server_drain(self.h), "mpt_server_drain")

    def start(self):
        self.ctx.check(self.ctx.L.mpt_server_start(self.h), "mpt_server_start")

    def stop(self):
        self.ctx.check(self.ctx.L.mpt_server_stop(self.h), "mpt_server_stop")

    def stats(self):
        out = (ctypes.c_uint64 * 6)()
        self.ctx.check(self.ctx.L.mpt_server_stats(self.h, out), "mpt_server_stats")
        return dict(zip(["batches", "paths", "extension_rays", "shadow_rays", "jobs_done", "peak_film_bytes"],
                        list(out)))

    def memory(self):
        out = (ctypes.c_uint64 * 6)()
        self.ctx.check(self.ctx.L.mpt_server_memory(self.h, out), "mpt_server_memory")
        return dict(zip(["budget", "queue_bytes", "film_bytes", "peak_film_bytes", "batch_paths", "lane_bytes"],
                        list(out)))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.mpt_server_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostKdTree:
    """The device kd-tree built on the host (mpt_kd_build), with host-thread
    closest hits through the same traversal code as the kernels."""

    def __init__(self, arrays, kd=None):
        self.L = lib()
        self.h = ctypes.c_void_p()
        d = arrays.desc()
        kp = kd if kd is not None else KdParams()
        rc = self.L.mpt_kd_build(ctypes.byref(d), ctypes.byref(kp), ctypes.byref(self.h))
        if rc:
            raise MptError(f"mpt_kd_build failed: {ERRORS.get(rc, rc)}")

    def stats(self):
        out = (ctypes.c_uint64 * 6)()
        self.L.mpt_kd_stats(self.h, out)
        return dict(zip(["nodes", "leaves", "refs", "max_depth", "build_ms", "max_leaf"], list(out)))

    def trace_closest(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        ctr = (ctypes.c_uint64 * 2)()
        rc = self.L.mpt_kd_trace(self.h, _ptr(rays), len(rays), _ptr(hits), ctr, threads)
        if rc:
            raise MptError(f"mpt_kd_trace failed: {ERRORS.get(rc, rc)}")
        return hits, {"nodes": ctr[0], "refs": ctr[1]}

    def close(self):
        if self.h:
            self.L.mpt_kd_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy(ptr, count, dtype, cols=None):
    if count == 0 or not ptr:
        return np.zeros((0,) if cols is None else (0, cols), dtype)
    n = count * (cols or 1)
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))), (n,))
    arr = np.array(arr, dtype)
    return arr if cols is None else arr.reshape(-1, cols)


def load_scenefile(path=None, text=None, base_dir=None):
    """A Milton scene file (mpt_scenefile_load / _parse) -> (SceneArrays, Job).
    The arrays are copies; the job holds the file's camera, film size,
    samples per pixel (noSuperSamples), depth cap and seed; arrays.tonemap
    the output's tone map (MPT_TONEMAP_*)."""
    L = lib()
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(1024)
    if text is not None:
        rc = L.mpt_scenefile_parse(text.encode(), (base_dir or "").encode(), ctypes.byref(h), err, len(err))
    else:
        rc = L.mpt_scenefile_load(str(path).encode(), ctypes.byref(h), err, len(err))
    if rc:
        raise MptError(f"scene file: {err.value.decode(errors='replace')}")
    try:
        d = L.mpt_scenefile_desc(h).contents
        nt, nv = d.num_triangles, d.num_vertices
        mats = np.zeros(d.num_materials, MATERIAL_DTYPE)
        ctypes.memmove(mats.ctypes.data, d.materials, mats.nbytes)
        arrays = SceneArrays(
            _copy(d.vertices, nv, np.float32, 3), _copy(d.indices, nt, np.uint32, 3),
            _copy(d.tri_material, nt, np.uint32), _copy(d.spheres, d.num_spheres, np.float32, 4),
            _copy(d.sphere_material, d.num_spheres, np.uint32), _copy(d.quads, d.num_quads, np.float32, 9),
            _copy(d.quad_material, d.num_quads, np.uint32), mats, list(d.env_radiance),
            tri_flags=_copy(d.tri_flags, nt, np.uint32) if d.tri_flags else None,
            vertex_normals=_copy(d.vertex_normals, nv, np.float32, 3) if d.vertex_normals else None,
            vertex_normal_mask=_copy(d.vertex_normal_mask, nv, np.uint8) if d.vertex_normal_mask else None)
        if d.env_map:
            e = d.env_map.contents
            img = _copy(e.rgb, e.width * e.height, np.float32, 3).reshape(e.height, e.width, 3)
            arrays.env_map = EnvMap(img, e.repeat_u or 1.0, e.repeat_v or 1.0, e.filter, e.support or 2,
                                    (e.params[0], e.params[1]))
        if d.num_points:
            arrays.points = _copy(d.points, d.num_points, np.float32, 6)
        if d.tri_shape:
            arrays.tri_shape = _copy(d.tri_shape, nt, np.uint32)
        if d.sphere_frames:
            arrays.sphere_frames = _copy(d.sphere_frames, d.num_spheres, np.float32, 9)
        job = Job()
        L.mpt_scenefile_job(h, ctypes.byref(job))
        c = job.camera
        arrays.camera = {"eye": list(c.eye), "look": list(c.look), "up": list(c.up), "height_angle": c.height_angle,
                         "aspect": c.aspect, "aperture": c.aperture, "focal_length": c.focal_length,
                         "focal_distance": c.focal_distance, "autofocus": bool(c.autofocus),
                         "focal_point": list(c.focal_point)}
        arrays.tonemap = L.mpt_scenefile_tonemap(h)   # the output's MPT_TONEMAP_*
        return arrays, job
    finally:
        L.mpt_scenefile_destroy(h)
