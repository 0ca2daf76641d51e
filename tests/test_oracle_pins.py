"""Independent pins of five oracle functions that the device path is checked
against, each by a closed form or an invariant computed here from first
principles (not by calling the oracle's own formula):

  * pinhole camera rays (ThinLensCamera::getWorldRay with aperture 0,
    cameras/ThinLensCamera.cpp:131-135, 181-256): every ray projected back
    through the reference's worldToFilm matrix chain A B C D E
    (_computeMatrix, :181-216) lands at its film position;
  * film strata (SuperSampleGenerator / JitteredSampleGenerator,
    renderers/generators/SuperSampleGenerator.cpp:48-91,
    JitteredSampleGenerator.cpp:30-31; core/Viewport.h:29-36): exactly one
    sample in each of the floor(sqrt(spp))^2 cells of every pixel, row-major,
    cycling for the samples beyond the square;
  * emission after a reflective specular bounce, Le * area / t^2
    (renderers/renderers/PathTracer.cpp:50-69, 97-98): a mirror facing a
    small light;
  * the depth-0 channel rule (PathTracer.cpp:107-114): a glass sphere under a
    white environment is a furnace -- every path through the glass is
    non-zero in exactly one channel with value 3 / 0.95^k, and the image mean
    is the environment in every channel;
  * barycentric interpolation of averaged vertex normals
    (shapes/Mesh.cpp:335-377, 419-433; MeshTriangle.cpp:344-361): a bent
    two-triangle quad under a point light.
"""
import math

import numpy as np

import milton_amd as m
from oracle import pyoracle as po


# ---------------------------------------------------------------------------
# Camera and strata.
def world_to_film(cam, near=0.01, far=100.0):
    """The reference's worldToFilm = A B C D E (ThinLensCamera::_computeMatrix
    with the basis of setOrientation), built here as 4x4 matrices."""
    eye = np.array(cam["eye"], np.float64)
    look = np.array(cam["look"], np.float64)
    up = np.array(cam["up"], np.float64)
    n = -look
    u = np.cross(up, n)
    n /= np.linalg.norm(n)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    k = near / far
    A = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1 / (k - 1), k / (k - 1)], [0, 0, -1, 0]], np.float64)
    B = np.diag([1 / far, 1 / far, 1 / far, 1.0])
    cot_h = 1.0 / math.tan(math.radians(cam["height_angle"]) / 2)
    C = np.diag([cot_h / cam["aspect"], cot_h, 1.0, 1.0])
    D = np.eye(4)
    D[0, :3], D[1, :3], D[2, :3] = u, v, n
    E = np.eye(4)
    E[:3, 3] = -eye
    return A @ B @ C @ D @ E


def film_xy(M, ray):
    """Film position (x, y) in [0, 1]^2 of the point one unit along a ray."""
    p = M @ np.append(ray[:3] + ray[3:], 1.0)
    X, Y = p[0] / p[3], p[1] / p[3]     # NDC (2x - 1, 1 - 2y)
    return (X + 1) / 2, (1 - Y) / 2


CAMERAS = [
    {"eye": [0.5, 0.5, -1.4], "look": [0, 0, 1], "up": [0, 1, 0], "height_angle": 40.0, "aspect": 1.0},
    {"eye": [0.1, 0.3, -1.0], "look": [0.3, -0.2, 1.0], "up": [0, 1, 0], "height_angle": 55.0, "aspect": 16 / 9},
    {"eye": [2.0, 1.0, 0.5], "look": [-1.0, -0.3, 0.2], "up": [0.1, 1.0, 0.3], "height_angle": 30.0, "aspect": 0.5},
]


def test_pinhole_rays_project_back_to_their_film_positions():
    for cam in CAMERAS:
        W, H, spp = 7, 5, 9
        job = m.make_job(cam, W, H, spp, 1, seed=3)
        M = world_to_film(cam)
        look = np.array(cam["look"]) / np.linalg.norm(cam["look"])
        for pixel in range(W * H):
            px, py = pixel % W, pixel // W
            for s in range(spp):
                r = po.camera_ray(job, pixel, s)
                assert np.allclose(r[:3], cam["eye"], atol=1e-6)
                assert abs(np.linalg.norm(r[3:]) - 1) < 1e-12
                assert r[3:] @ look > 0                       # rays leave forward
                x, y = film_xy(M, r)
                assert px <= x * W < px + 1 and py <= y * H < py + 1, (cam, pixel, s, x * W, y * H)


def test_film_corners_and_axes():
    """The film's centre maps to the look direction, its top edge to half the
    height angle above it, its right edge to atan(aspect tan(h / 2)) to the
    right (look x up), for both aspect orientations."""
    for cam in CAMERAS:
        M = world_to_film(cam)
        look = np.array(cam["look"], np.float64)
        look /= np.linalg.norm(look)
        up = np.array(cam["up"], np.float64)
        right = np.cross(look, up)
        right /= np.linalg.norm(right)
        top = np.cross(right, look)
        h = math.radians(cam["height_angle"]) / 2
        eye = np.array(cam["eye"], np.float64)
        # directions from the angles, projected to the film
        for d, want in ((look, (0.5, 0.5)),
                        (math.cos(h) * look + math.sin(h) * top, (0.5, 0.0)),
                        (look + cam["aspect"] * math.tan(h) * right, (1.0, 0.5)),
                        (look - cam["aspect"] * math.tan(h) * right - math.tan(h) * top, (0.0, 1.0))):
            x, y = film_xy(M, np.concatenate([eye, d / np.linalg.norm(d)]))
            assert abs(x - want[0]) < 1e-12 and abs(y - want[1]) < 1e-12, (cam, want, x, y)


def test_strata_one_sample_per_cell():
    cam = CAMERAS[1]
    for spp, side in ((16, 4), (9, 3), (12, 3), (1, 1), (64, 8)):
        W, H = 4, 3
        job = m.make_job(cam, W, H, spp, 1, seed=11)
        M = world_to_film(cam)
        for pixel in range(W * H):
            px, py = pixel % W, pixel // W
            cells = []
            for s in range(spp):
                x, y = film_xy(M, po.camera_ray(job, pixel, s))
                cx, cy = int((x * W - px) * side), int((y * H - py) * side)
                assert 0 <= cx < side and 0 <= cy < side
                cells.append(cy * side + cx)
            sq = side * side
            # row-major cells in sample order, cycling past the square
            assert cells == [s % sq for s in range(spp)], (spp, pixel, cells)
            assert sorted(cells[:sq]) == list(range(sq))


# ---------------------------------------------------------------------------
# Specular emission: Le * area / t^2 after a mirror bounce.
def _mirror_scene(L0=5.0, a=1.0):
    mats = np.zeros(3, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIELECTRIC   # mirror: never refracts
    mats[0]["ks"] = 1.0
    mats[0]["ior"] = 1.5
    mats[0]["transparency"] = 0.0
    mats[1]["type"] = m.MPT_BSDF_DIFFUSE      # the light: black, emitting L0
    mats[1]["Le"] = L0
    quads = np.array([[0, 0, 0, 4, 0, 0, 0, 0, 4],                # mirror, y = 0, normal +y
                      [0, 1, 1, a, 0, 0, 0, 0, -a]], np.float32)  # light at (0, 1, 1), normal -y
    a_ = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], quads, [0, 1], mats, [0, 0, 0])
    a_.camera = {"eye": [0.0, 1.0, -1.0], "look": [0, -1, 1], "up": [0, 1, 1], "height_angle": 20.0, "aspect": 1.0}
    return a_


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


def test_specular_bounce_emission_closed_form():
    L0, a = 5.0, 1.0
    scene = _mirror_scene(L0, a)
    job = m.make_job(scene.camera, 24, 24, 4, 6, seed=2)
    n = 24 * 24 * 4
    pix = np.repeat(np.arange(24 * 24, dtype=np.uint32), 4)
    smp = np.tile(np.arange(4, dtype=np.uint32), 24 * 24)
    L = po.OracleScene(scene).render_paths(job, pix, smp)
    q = scene.quads.astype(np.float64)
    mirror, light = (q[0, :3], q[0, 3:6], q[0, 6:]), (q[1, :3], q[1, 3:6], q[1, 6:])
    area = float(np.linalg.norm(np.cross(light[1], light[2])))
    Le = float(np.float32(L0))
    via_mirror = channels = 0
    for i in range(n):
        r = po.camera_ray(job, int(pix[i]), int(smp[i]))
        o, d = r[:3], r[3:]
        tm, tl = _quad_hit(o, d, *mirror), _quad_hit(o, d, *light)
        if tl < tm:                          # the light seen directly: depth-0 emission
            assert np.allclose(L[i], Le, rtol=1e-12)
            continue
        assert np.isfinite(tm)
        p = o + tm * d
        d2 = d * np.array([1, -1, 1])        # reflect about +y
        t2 = _quad_hit(p, d2, *light)
        if not np.isfinite(t2):
            assert (L[i] == 0).all()
            continue
        via_mirror += 1
        want = 3 * Le * area / t2 ** 2       # PathTracer.cpp:59-64, then the x3 rule of :107-114
        nz = np.nonzero(L[i])[0]
        assert len(nz) == 1, L[i]
        channels |= 1 << int(nz[0])
        assert abs(L[i, nz[0]] - want) <= 1e-9 * want, (L[i], want)
    assert via_mirror > n // 5 and channels == 7


# ---------------------------------------------------------------------------
# Furnace: a glass sphere under a white environment.
def test_glass_sphere_furnace_and_channel_rule():
    mats = np.zeros(1, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIELECTRIC
    mats[0]["ks"] = 1.0
    mats[0]["ior"] = 1.5
    mats[0]["transparency"] = 1.0
    scene = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.array([[0, 0, 0, 1.0]]), [0],
                          np.zeros((0, 9)), [], mats, [1, 1, 1])
    scene.camera = {"eye": [0, 0, -3.0], "look": [0, 0, 1], "up": [0, 1, 0], "height_angle": 45.0, "aspect": 1.0}
    W, spp = 48, 16
    job = m.make_job(scene.camera, W, W, spp, 64, seed=7)
    pix = np.repeat(np.arange(W * W, dtype=np.uint32), spp)
    smp = np.tile(np.arange(spp, dtype=np.uint32), W * W)
    L = po.OracleScene(scene).render_paths(job, pix, smp)
    miss = (L == 1.0).all(axis=1)
    glass = ~miss
    assert 0.3 < glass.mean() < 0.9
    nz = (L[glass] != 0).sum(axis=1)
    # the x3 rule keeps one channel; a path Russian roulette ends inside the
    # glass (probability 0.05 per vertex from depth 2) returns 0
    assert (nz <= 1).all() and (nz == 0).mean() < 0.02
    glass = np.nonzero(glass)[0][nz == 1]
    v = L[glass].max(axis=1)
    k = np.log(v / 3) / np.log(1 / 0.95)     # Russian roulette survivals at depth >= 2
    assert np.allclose(k, np.round(k), atol=1e-9) and (np.round(k) >= 0).all()
    ch = np.argmax(L[glass], axis=1)
    frac = np.bincount(ch, minlength=3) / len(glass)
    assert np.all(np.abs(frac - 1 / 3) < 4 * math.sqrt(2 / 9 / len(glass)))
    mean, sd = L.mean(axis=0), L.std(axis=0) / math.sqrt(len(L))
    print(f"furnace mean {mean} +- {sd}", flush=True)
    assert np.all(np.abs(mean - 1.0) < 4 * sd)


# ---------------------------------------------------------------------------
# Interpolated vertex normals on a bent quad.
KD, P_LIGHT, LIGHT = 0.7, np.array([3.0, 2.0, 1.0]), np.array([0.4, 2.0, 0.6])


def _bent_quad():
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0.5, 1]], np.float64)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)   # face normals +y and (0.5, 1, -0.5)/|.|
    mats = np.zeros(1, m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = KD
    a = m.SceneArrays(verts, idx, [0, 0], np.zeros((0, 4)), [], np.zeros((0, 9)), [], mats, [0, 0, 0],
                      points=[[*LIGHT, *P_LIGHT]])
    a.camera = {"eye": [0.5, 2.5, 0.5], "look": [0, -1, 0], "up": [0, 0, 1], "height_angle": 30.0, "aspect": 1.0}
    return a


def test_interpolated_normals_closed_form():
    a = _bent_quad()
    V = a.vertices.astype(np.float64)
    faces = a.indices.astype(int)
    fn = []
    for f in faces:
        nrm = np.cross(V[f[1]] - V[f[0]], V[f[2]] - V[f[0]])
        fn.append(nrm / np.linalg.norm(nrm))
    vn = np.zeros_like(V)
    for f, nrm in zip(faces, fn):
        vn[f] += nrm
    vn = (vn.T / np.linalg.norm(vn, axis=1)).T          # Gouraud: normalised sum of unit face normals
    vn = vn.astype(np.float32).astype(np.float64)
    job = m.make_job(a.camera, 16, 16, 4, 1, seed=4)
    pix = np.repeat(np.arange(256, dtype=np.uint32), 4)
    smp = np.tile(np.arange(4, dtype=np.uint32), 256)
    L = po.OracleScene(a).render_paths(job, pix, smp)
    lp = LIGHT.astype(np.float32).astype(np.float64)
    pw = P_LIGHT.astype(np.float32).astype(np.float64)
    kd = float(np.float32(KD))
    lit = tested = 0
    for i in range(len(pix)):
        r = po.camera_ray(job, int(pix[i]), int(smp[i]))
        o, d = r[:3], r[3:]
        best, hit = np.inf, None
        for fi, f in enumerate(faces):
            A, B, C = V[f[0]], V[f[1]], V[f[2]]
            nrm = np.cross(B - A, C - A)
            t = nrm @ (A - o) / (nrm @ d)
            p = o + t * d
            # barycentrics from sub-triangle areas (MeshTriangle.cpp:344-361)
            tot = np.linalg.norm(nrm)
            wa = np.linalg.norm(np.cross(B - p, C - p)) / tot
            wb = np.linalg.norm(np.cross(p - A, C - A)) / tot
            wc = 1 - wa - wb
            if t > 0 and wc >= -1e-12 and abs(wa + wb + wc - 1) < 1e-9 and t < best and \
                    abs(np.linalg.norm(np.cross(A - p, B - p)) / tot - wc) < 1e-9:
                best, hit = t, (p, wa * vn[f[0]] + wb * vn[f[1]] + wc * vn[f[2]])
        if hit is None:
            assert (L[i] == 0).all()
            continue
        tested += 1
        p, ns = hit
        ns /= np.linalg.norm(ns)
        w = lp - p
        dist = np.linalg.norm(w)
        w /= dist
        want = np.zeros(3)
        if ns @ d < 0 < ns @ w:              # one-sided diffuse (DiffuseBSDF.cpp:22, 41)
            want = kd / math.pi * pw / (4 * math.pi) * abs(ns @ w) / dist ** 2
            lit += 1
        assert np.allclose(L[i], want, rtol=1e-6, atol=1e-12), (i, L[i], want)
    assert tested > 500 and lit > 500
