"""The bidirectional path tracer on the CPU: the oracle's restatement of
BidirectionalPathTracer::sample (renderers/renderers/BidirectionalPathTracer.
cpp:34-174) over Path (renderers/utils/Path.cpp: append/prepend, _initL,
_initE, the join of the subpaths, getPd, getContribution) and PathVertex.

Pinned by closed forms in which every split but one has density zero, so the
power heuristic gives that split weight 1 and a path's value is computed here
from the geometry alone:
  * a camera looking at an emitter: the eye subpath ends on it (k = 2, s = 0)
    and the pixel is Le, clamped to 1 (RADIANCE_HACK);
  * a tiny light over a diffuse floor: the light's sampled point joined to the
    floor vertex (s = 1, t = 2) gives Le A cos cos / r^2 kd / pi, next-event
    estimation's value, for the point EmitterSampler drew;
  * a point light over the floor: P / (4 pi) kd / pi cos / r^2, with
    OmniEmitter's pdf 1 and Le1 = 1 / (4 pi) (UniformOnSphereSampler.cpp:32-34);
  * a mirror between the camera and a light: the eye subpath reflects into the
    light with the roulette's 0.95 at the mirror, Le / 0.95, or gives 0 when
    the roulette ends it (specular vertices join nothing).
"""
import math

import numpy as np

import milton_amd as m
from oracle import pyoracle as po

KEY = 0x4D505431
CAMERA_DEPTH, LIGHT_DEPTH = 255, 1024


def _u(seed, job_id, pixel, sample, depth, stream):
    """The four 24-bit uniforms of one Philox block of the path streams."""
    r = po.philox([pixel, sample, (depth << 8) | stream, job_id], [seed, KEY])
    return (r >> 8).astype(np.float64) / 16777216.0


def _quad_hit(o, d, c, ex, ez):
    n = np.cross(ez, ex)
    dn = n @ d
    if dn == 0:
        return np.inf
    t = n @ (c - o) / dn
    p = o + t * d - c
    if t > 1e-7 and abs(p @ ex / (ex @ ex)) <= 0.5 and abs(p @ ez / (ez @ ez)) <= 0.5:
        return t
    return np.inf


def _paths(w, h, spp):
    pix = np.repeat(np.arange(w * h, dtype=np.uint32), spp)
    smp = np.tile(np.arange(spp, dtype=np.uint32), w * h)
    return pix, smp


def _bidir(scene, w, h, spp, depth, seed):
    job = m.make_job(scene.camera, w, h, spp, depth, seed=seed, renderer=m.MPT_RENDERER_BIDIR)
    pix, smp = _paths(w, h, spp)
    return job, pix, smp, po.OracleScene(scene).render_paths(job, pix, smp)


def _scene(quads, quad_mat, mats, points=None):
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [],
                      np.asarray(quads, np.float32), quad_mat, mats, [0, 0, 0], points=points)
    return a


def test_direct_view_of_an_emitter_is_its_radiance_clamped():
    for le, want in ((0.6, np.float32(0.6)), (1.5, 1.0)):
        mats = np.zeros(1, m.MATERIAL_DTYPE)
        mats[0]["type"] = m.MPT_BSDF_DIFFUSE
        mats[0]["kd"] = 0.5
        mats[0]["Le"] = le
        a = _scene([[0, 0, 0, 1, 0, 0, 0, 0, 1]], [0], mats)   # normal +y
        a.camera = {"eye": [0.1, 2.0, 0.05], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 40.0,
                    "aspect": 1.0}
        job, pix, smp, L = _bidir(a, 12, 12, 2, 4, seed=3)
        q = a.quads.astype(np.float64)[0]
        hits = 0
        for i in range(len(pix)):
            r = po.camera_ray(job, int(pix[i]), int(smp[i]))
            if np.isfinite(_quad_hit(r[:3], r[3:], q[:3], q[3:6], q[6:])):
                hits += 1
                assert np.allclose(L[i], float(want), rtol=1e-12), (L[i], want)
            else:
                assert (L[i] == 0).all()
        assert 0 < hits < len(pix)


KD = 0.6
H = 1.5


def _floor_mats(light_le):
    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = KD
    mats[1]["type"] = m.MPT_BSDF_DIFFUSE   # the light: black, emitting
    mats[1]["Le"] = light_le
    return mats


def test_tiny_light_over_a_floor_is_next_event_estimation():
    side, le = 1e-4, 1e7
    mats = _floor_mats(le)
    light = [0.2, H, -0.1, side, 0, 0, 0, 0, -side]   # normal -y
    a = _scene([[0, 0, 0, 30, 0, 0, 0, 0, 30], light], [0, 1], mats)
    a.camera = {"eye": [0.3, 3.0, -0.2], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 50.0, "aspect": 1.0}
    seed = 7
    job, pix, smp, L = _bidir(a, 12, 12, 4, 4, seed)
    q = a.quads.astype(np.float64)[1]
    c, ex, ez = q[:3], q[3:6], q[6:]
    nl = np.cross(ez, ex)
    area = float(np.linalg.norm(nl))
    nl /= area
    kd = float(np.float32(KD))
    le32 = float(np.float32(le))
    lit = 0
    for i in range(len(pix)):
        r = po.camera_ray(job, int(pix[i]), int(smp[i]))
        t = -r[1] / r[4]
        x = r[:3] + t * r[3:]
        # the light subpath's start: EmitterSampler's light 0, Plane::getPosition at (u, v)
        u = _u(seed, 0, int(pix[i]), int(smp[i]), LIGHT_DEPTH, 0)
        lp = c + ex * (u[1] - 0.5) + ez * (u[2] - 0.5)
        w = lp - x
        d = np.linalg.norm(w)
        w /= d
        G = abs(w[1]) * abs(nl @ w) / d ** 2
        want = le32 * area * G * kd / math.pi
        assert abs(L[i] - want).max() <= 1e-6 * want, (i, L[i], want)
        lit += 1
    assert lit == len(pix)


def test_point_light_over_a_floor():
    P = np.array([4.0, 2.0, 1.0])
    lpos = np.array([0.2, H, -0.1])
    mats = _floor_mats(0.0)[:1]
    a = _scene([[0, 0, 0, 20, 0, 0, 0, 0, 20]], [0], mats, points=[[*lpos, *P]])
    a.camera = {"eye": [0.3, 3.0, -0.2], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 50.0, "aspect": 1.0}
    job, pix, smp, L = _bidir(a, 12, 12, 4, 3, seed=5)
    lp = lpos.astype(np.float32).astype(np.float64)
    pw = P.astype(np.float32).astype(np.float64)
    kd = float(np.float32(KD))
    for i in range(len(pix)):
        r = po.camera_ray(job, int(pix[i]), int(smp[i]))
        t = -r[1] / r[4]
        x = r[:3] + t * r[3:]
        w = lp - x
        d = np.linalg.norm(w)
        want = kd / math.pi * pw / (4 * math.pi) * (w[1] / d) / d ** 2
        assert np.allclose(L[i], want, rtol=1e-9, atol=0), (L[i], want)


def test_mirror_reflects_the_eye_subpath_into_the_light():
    le = 0.5
    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIELECTRIC   # mirror: never refracts
    mats[0]["ks"] = 1.0
    mats[0]["ior"] = 1.5
    mats[0]["transparency"] = 0.0
    mats[1]["type"] = m.MPT_BSDF_DIFFUSE      # the light: black, emitting
    mats[1]["Le"] = le
    a = _scene([[0, 0, 0, 4, 0, 0, 0, 0, 4], [0, 1, 1, 1, 0, 0, 0, 0, -1]], [0, 1], mats)
    a.camera = {"eye": [0.0, 1.0, -1.0], "look": [0, -1, 1], "up": [0, 1, 1], "height_angle": 20.0, "aspect": 1.0}
    seed = 2
    job, pix, smp, L = _bidir(a, 20, 20, 4, 6, seed)
    q = a.quads.astype(np.float64)
    mirror, light = (q[0, :3], q[0, 3:6], q[0, 6:]), (q[1, :3], q[1, 3:6], q[1, 6:])
    le32 = float(np.float32(le))
    direct = via = ended = 0
    for i in range(len(pix)):
        r = po.camera_ray(job, int(pix[i]), int(smp[i]))
        o, d = r[:3], r[3:]
        tm, tl = _quad_hit(o, d, *mirror), _quad_hit(o, d, *light)
        if tl < tm:
            direct += 1
            assert np.allclose(L[i], le32, rtol=1e-12)
            continue
        p = o + tm * d
        t2 = _quad_hit(p, d * np.array([1, -1, 1]), *light)
        # the mirror is eye vertex 1: its roulette uniform, q = min(0.95, ks / 1)
        rr = _u(seed, 0, int(pix[i]), int(smp[i]), 1, 2)[3]
        if np.isfinite(t2) and rr < 0.95:
            via += 1
            assert np.allclose(L[i], le32 / 0.95, rtol=1e-12), (L[i], le32 / 0.95)
        else:
            ended += np.isfinite(t2)
            assert (L[i] == 0).all(), L[i]
    assert via > len(pix) // 5 and ended > 0
