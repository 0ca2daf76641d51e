"""Pins the CPU oracle before it is trusted as the checker.

The reference ships no golden vectors or known-answer tests for this path
(its tests/ programs are visual or stale), so each oracle function is pinned
against published known answers, closed forms, brute force, or invariants the
mathematics fixes, chosen so that a dropped term, a wrong sign or a swapped
operand fails.
"""
import math

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from conftest import rays64


# ---------------------------------------------------------------------------
# RNG
# ---------------------------------------------------------------------------
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors of Random123 (Salmon et al. 2011)."""
    assert list(po.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(po.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(po.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ---------------------------------------------------------------------------
# Scenes and traversal
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def room():
    arrays = m.generate(m.MPT_SCENE_ROOM)
    return arrays, po.OracleScene(arrays)


@pytest.fixture(scope="module")
def soup():
    arrays = m.generate(m.MPT_SCENE_SOUP, 20_000, 1)
    return arrays, po.OracleScene(arrays), po.OracleScene(arrays, naive=True)


def _ray(o, d):
    d = np.asarray(d, np.float64)
    return np.concatenate([o, d / np.linalg.norm(d)])[None]


def test_room_axis_aligned_closed_form(room):
    arrays, osc = room
    nt = len(arrays.indices)
    o = [0.5, 0.8, 0.5]   # above both boxes
    cases = [([1, 0, 0], 0.5, 2), ([-1, 0, 0], 0.5, 1), ([0, -1, 0], 0.8, 0), ([0, 1, 0], 0.199, None),
             ([0, 0, 1], 0.5, 0)]
    for d, t_exp, mat in cases:
        t, i = osc.trace_closest(_ray(o, d))
        assert abs(t[0] - t_exp) < 1e-6, (d, t)
        if mat is None:
            assert i[0] == nt          # the light quad
        else:
            assert arrays.tri_material[i[0]] == mat
    t, i = osc.trace_closest(_ray(o, [0, 0, -1]))
    assert i[0] == po.MISS and math.isinf(t[0])


def test_room_oblique_closed_form(room):
    """A ray from the room centre to a known floor point hits at |p - o|."""
    _, osc = room
    rng = np.random.default_rng(0)
    for _ in range(50):
        target = np.array([rng.uniform(0.9, 0.99), 0.0, rng.uniform(0.55, 0.95)])
        o = np.array([0.95, 0.9, 0.95])
        d = target - o
        t, i = osc.trace_closest(_ray(o, d))
        assert abs(t[0] - np.linalg.norm(d)) < 1e-9


def test_kd_equals_naive(soup):
    """kd traversal == linear scan (accel/kdTreeAccel.cpp:45-59 recipe)."""
    arrays, kd, naive = soup
    rays = rays64(m.generate_rays(m.MPT_SCENE_SOUP, 20_000, 20_000, seed=9))
    t1, i1 = kd.trace_closest(rays)
    t2, i2 = naive.trace_closest(rays)
    assert (i1 == i2).mean() >= 0.9999
    same = (i1 == i2) & (i1 != po.MISS)
    assert np.all(t1[same] == t2[same])
    assert (i1 != po.MISS).sum() > 500


def test_kd_invariants(soup, room):
    _, kd, _ = soup
    st = kd.kd_stats(check=True)
    assert st["violations"] == 0
    assert st["max_depth"] <= 24
    assert st["refs"] >= 20_000
    st2 = room[1].kd_stats(check=True)
    assert st2["violations"] == 0


def test_any_hit_consistent_with_closest(soup):
    arrays, kd, naive = soup
    rays = rays64(m.generate_rays(m.MPT_SCENE_SOUP, 20_000, 10_000, seed=13))
    t, _ = kd.trace_closest(rays)
    rng = np.random.default_rng(1)
    tmax = rng.uniform(0.001, 0.5, len(rays))
    occ = kd.trace_any(rays, tmax)
    occ_n = naive.trace_any(rays, tmax)
    clear = np.abs(t - tmax) > 1e-6
    assert np.all(occ[clear] == (t[clear] < tmax[clear]))
    assert np.all(occ == occ_n)


def test_sphere_and_quad_closed_form():
    """Sphere (shapes/Sphere.cpp:16-45) and plane (shapes/Plane.cpp:25-38)."""
    mats = np.zeros(1, m.MATERIAL_DTYPE)
    mats["kd"] = 0.5
    arrays = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], [[0.0, 0.0, 0.0, 0.25]], [0],
                           [[0.0, 0.0, 2.0, 1.0, 0, 0, 0, 1.0, 0]], [0], mats, [0, 0, 0])
    osc = po.OracleScene(arrays)
    rng = np.random.default_rng(2)
    for _ in range(100):
        o = rng.uniform(-2, 2, 3)
        o = o / np.linalg.norm(o) * rng.uniform(0.5, 1.5)
        d = -o + rng.uniform(-0.1, 0.1, 3)
        d /= np.linalg.norm(d)
        b = np.dot(o, d)
        c = np.dot(o, o) - 0.25 ** 2
        disc = b * b - c
        t, i = osc.trace_closest(np.concatenate([o, d])[None])
        if disc > 0 and -b - math.sqrt(disc) > 1e-7:
            assert i[0] == 0
            assert abs(t[0] - (-b - math.sqrt(disc))) < 1e-12
    # inside the sphere: the far root
    t, i = osc.trace_closest(_ray([0, 0, 0], [0, 1, 0]))
    assert i[0] == 0 and abs(t[0] - 0.25) < 1e-12
    # quad in the plane z = 2 (normal ez x ex = -z), |x|,|y| <= 0.5
    t, i = osc.trace_closest(_ray([0.3, -0.4, 0.5], [0, 0, 1]))
    assert i[0] == 1 and abs(t[0] - 1.5) < 1e-12
    t, i = osc.trace_closest(_ray([0.6, 0.0, 0.5], [0, 0, 1]))
    assert i[0] == po.MISS


# ---------------------------------------------------------------------------
# BSDFs
# ---------------------------------------------------------------------------
def _mat(type_, kd=0.5, ks=0.0, n=1.0, ior=1.5, transparency=1.0):
    row = np.zeros(16, np.float32)
    row[0] = type_
    row[1:4] = kd
    row[4:7] = ks
    row[7] = n
    row[8] = ior
    row[9] = transparency
    return row


def test_diffuse_cosine_sampling():
    """DiffuseBSDF: f/pdf = kd, E[cos] = 2/3 under cosine sampling, back-side
    hits are absorbed (one-sided, DiffuseBSDF.cpp:21-48)."""
    mat = _mat(0, kd=[0.3, 0.5, 0.7])
    n = np.array([0.3, 0.8, -0.52])
    n /= np.linalg.norm(n)
    wi = -n + np.array([0.2, 0.0, 0.1])
    wi /= np.linalg.norm(wi)
    rng = np.random.default_rng(4)
    cos = []
    for _ in range(4000):
        u = rng.random(4)
        wo, fs, pdf, spec = po.bsdf_sample(mat, n, wi, u)
        assert not spec
        assert abs(np.linalg.norm(wo) - 1) < 1e-12
        np.testing.assert_allclose(fs, np.float32([0.3, 0.5, 0.7]).astype(np.float64), rtol=1e-12)
        assert abs(pdf - 1 / math.pi) < 1e-15
        cos.append(np.dot(wo, n))
    cos = np.array(cos)
    assert cos.min() >= 0
    assert abs(cos.mean() - 2 / 3) < 3 * cos.std() / math.sqrt(len(cos))
    wo, fs, pdf, _ = po.bsdf_sample(mat, n, -wi, [0.3, 0.4, 0.5, 0.6])
    assert np.all(wo == 0)


def test_phong_sampling_and_consistency():
    """Modified Phong (ModifiedPhongBSDF.cpp:84-181): lobe choice by average
    kd / ks, fs = f/pdf of the sampled lobe, the specular lobe follows
    cos(alpha) = u^(1/(n+1)) about the mirror direction (E = (n+1)/(n+2)),
    the diffuse lobe is cosine-weighted (E[cos] = 2/3), the BRDF is
    reciprocal and the specular lobe pdf is normalised."""
    kd, ks, n_exp = 0.2, 0.6, 50.0
    mat = _mat(1, kd=kd, ks=ks, n=n_exp)
    n = np.array([0.0, 1.0, 0.0])
    rng = np.random.default_rng(5)
    wi = np.array([0.0, -1.0, 0.0])      # normal incidence: mirror direction = n
    spec_cos, diff_cos, absorbed = [], [], 0
    for _ in range(20000):
        u = rng.random(4)
        wo, fs, pdf, _ = po.bsdf_sample(mat, n, wi, u)
        if np.all(wo == 0):
            absorbed += 1
            continue
        f, _ = po.bsdf_eval(mat, n, wi, wo)
        np.testing.assert_allclose(fs, f / pdf, rtol=1e-9)
        (diff_cos if u[3] - 1e-7 <= np.float32(kd) else spec_cos).append(wo[1])
    total = 20000
    for frac, p in [(len(diff_cos) / total, kd), (len(spec_cos) / total, ks), (absorbed / total, 1 - kd - ks)]:
        assert abs(frac - p) < 4 * math.sqrt(p * (1 - p) / total)
    sc, dc = np.array(spec_cos), np.array(diff_cos)
    assert abs(sc.mean() - (n_exp + 1) / (n_exp + 2)) < 4 * sc.std() / math.sqrt(len(sc))
    assert abs(dc.mean() - 2 / 3) < 4 * dc.std() / math.sqrt(len(dc))
    # reciprocity f(wi -> wo) = f(-wo -> -wi)
    for _ in range(200):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[1], b[1] = -abs(a[1]), abs(b[1])
        wi, wo = a / np.linalg.norm(a), b / np.linalg.norm(b)
        f1, _ = po.bsdf_eval(mat, n, wi, wo)
        f2, _ = po.bsdf_eval(mat, n, -wo, -wi)
        np.testing.assert_allclose(f1, f2, rtol=1e-12, atol=1e-15)
    # the specular lobe pdf (n+1)/(2 pi) cos^n integrates to 1 over the
    # hemisphere around the mirror direction
    th = np.linspace(0, math.pi / 2, 20001)
    integrand = (n_exp + 1) / (2 * math.pi) * np.cos(th) ** n_exp * np.sin(th) * 2 * math.pi
    assert abs(np.trapezoid(integrand, th) - 1) < 1e-6


def test_fresnel_textbook():
    """Fresnel::getFr (Fresnel.h:45-90) against the normal-incidence closed
    form, total internal reflection, and Brewster's angle (Rp = 0)."""
    n = np.array([0.0, 0.0, 1.0])
    wi = np.array([0.0, 0.0, -1.0])
    assert abs(po.fresnel(wi, n, 1.0, 1.5) - ((1.5 - 1) / 2.5) ** 2) < 1e-12
    assert abs(po.fresnel(-wi, n, 1.0, 1.5) - ((1.5 - 1) / 2.5) ** 2) < 1e-12
    th = math.radians(60)
    assert po.fresnel(np.array([math.sin(th), 0, math.cos(th)]), n, 1.0, 1.5) == 1.0   # inside, TIR
    tb = math.atan(1.5)
    wi_b = np.array([math.sin(tb), 0, -math.cos(tb)])
    ct = math.cos(math.asin(math.sin(tb) / 1.5))
    rs = (math.cos(tb) - 1.5 * ct) / (math.cos(tb) + 1.5 * ct)
    assert abs(po.fresnel(wi_b, n, 1.0, 1.5) - rs * rs / 2) < 1e-12


def test_dielectric_snell_and_channel():
    """DielectricBSDF: refraction obeys Snell's law with the air IOR 1.0003,
    keeps only the path's IOR channel; reflection mirrors about the facing
    normal (DielectricBSDF.cpp:55-158)."""
    mat = _mat(2, kd=0.0, ks=1.0, ior=1.5)
    n = np.array([0.0, 1.0, 0.0])
    th = 0.6
    wi = np.array([math.sin(th), -math.cos(th), 0.0])
    wo, fs, pdf, spec = po.bsdf_sample(mat, n, wi, [0.1, 0.2, 0.3, 0.0], ior_index=1)
    assert spec and pdf == 1
    sin_t = math.hypot(wo[0], wo[2])
    assert abs(1.0003 * math.sin(th) - 1.5 * sin_t) < 1e-12 and wo[1] < 0
    np.testing.assert_array_equal(fs, [0, 1, 0])
    wo, fs, _, _ = po.bsdf_sample(mat, n, wi, [0.1, 0.2, 0.3, 0.999999], ior_index=1)
    np.testing.assert_allclose(wo, [math.sin(th), math.cos(th), 0], atol=1e-15)
    np.testing.assert_array_equal(fs, [1, 1, 1])


def test_dielectric_dispersive_ior():
    """A dispersive IOR spectrum: each channel refracts with its own index
    (m_ior[index] at the path's channel, DielectricBSDF.cpp:65-87)."""
    mat = _mat(2, kd=0.0, ks=1.0, ior=1.4)
    mat[13], mat[14] = 1.55, 1.8   # ior_g, ior_b
    n = np.array([0.0, 1.0, 0.0])
    th = 0.6
    wi = np.array([math.sin(th), -math.cos(th), 0.0])
    for ch, eta in ((0, 1.4), (1, 1.55), (2, 1.8)):
        wo, fs, _, _ = po.bsdf_sample(mat, n, wi, [0.1, 0.2, 0.3, 0.0], ior_index=ch)
        sin_t = math.hypot(wo[0], wo[2])
        assert abs(1.0003 * math.sin(th) - float(np.float32(eta)) * sin_t) < 1e-12, ch
        assert fs[ch] == 1 and fs.sum() == 1


# ---------------------------------------------------------------------------
# Estimator closed forms
# ---------------------------------------------------------------------------
def _floor_scene(env, light_size=0.0, light_h=1.0, Le=0.0, kd=0.5):
    """A 20 x 20 diffuse floor at y = 0 facing +y, optional small light quad
    above the origin facing down, constant environment."""
    v = np.array([[-10, 0, -10], [-10, 0, 10], [10, 0, 10], [10, 0, -10]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    mats = np.zeros(2, m.MATERIAL_DTYPE)
    mats["kd"][0] = kd
    mats["Le"][1] = Le
    quads = np.zeros((0, 9))
    qm = []
    if light_size > 0:
        s = light_size
        quads = np.array([[0, light_h, 0, s, 0, 0, 0, 0, -s]], np.float32)
        qm = [1]
    cam = {"eye": [0.0, 2.0, -2.0], "look": [0.0, -1.0, 1.0], "up": [0.0, 1.0, 0.0], "height_angle": 30.0,
           "aspect": 1.0}
    return m.SceneArrays(v, idx, [0, 0], np.zeros((0, 4)), [], quads, qm, mats, env, camera=cam)


def test_estimator_floor_under_environment():
    """PathTracer with a diffuse floor under a constant environment E: a path
    hitting the floor from above returns kd * E exactly (f/pdf = kd, the
    bounce escapes, PathTracer.cpp:36-41, 88-105); a miss returns E; with a
    depth cap of 1 the floor path returns 0."""
    arrays = _floor_scene([2.0, 1.0, 0.5])
    osc = po.OracleScene(arrays)
    job = m.make_job(arrays.camera, 16, 16, 4, 4, seed=3)
    pix = np.repeat(np.arange(256, dtype=np.uint32), 4)
    smp = np.tile(np.arange(4, dtype=np.uint32), 256)
    L = osc.render_paths(job, pix, smp)
    hit = np.isclose(L, [1.0, 0.5, 0.25]).all(axis=1)
    miss = np.isclose(L, [2.0, 1.0, 0.5]).all(axis=1)
    assert np.all(hit | miss) and hit.mean() > 0.5
    job1 = m.make_job(arrays.camera, 16, 16, 4, 1, seed=3)
    L1 = osc.render_paths(job1, pix, smp)
    assert np.all(np.isclose(L1, 0).all(axis=1) | np.isclose(L1, [2.0, 1.0, 0.5]).all(axis=1))


def test_estimator_direct_light_closed_form():
    """Next-event estimation (DirectIllumination.cpp:46-136) from a tiny
    light of side s at height h onto a floor point p: L = kd/pi * Le * s^2 *
    cos_p * cos_l / r^2 (the point-light limit, exact to O(s^2 / r^2))."""
    s, h, Le, kd = 1e-5, 1.0, 1e10, 0.5
    arrays = _floor_scene([0, 0, 0], light_size=s, light_h=h, Le=Le, kd=kd)
    osc = po.OracleScene(arrays)
    job = m.make_job(arrays.camera, 16, 16, 1, 1, seed=5)
    pix = np.arange(256, dtype=np.uint32)
    L = osc.render_paths(job, pix, np.zeros(256, np.uint32))
    for k in range(0, 256, 17):
        ray = po.camera_ray(job, int(pix[k]), 0)
        o, d = ray[:3], ray[3:]
        t = -o[1] / d[1]
        p = o + t * d
        r = np.array([0.0, h, 0.0]) - p
        dist = np.linalg.norm(r)
        cos_p = r[1] / dist
        expected = kd / math.pi * Le * s * s * cos_p * cos_p / dist ** 2
        assert abs(L[k, 0] - expected) <= 1e-4 * expected


def _form_factor_rect(p, h, x0, x1, z0, z1):
    """Configuration factor from a floor point p (facing +y) to the rectangle
    [x0, x1] x [z0, z1] at height h above it, facing down: the corner formula
    for a rectangle with a corner above the point (Howell's catalogue, B-1
    parallel case) combined over the four corners with signs."""
    def corner(x, z):
        X, Z = abs(x) / h, abs(z) / h
        f = (X / math.sqrt(1 + X * X) * math.atan(Z / math.sqrt(1 + X * X))
             + Z / math.sqrt(1 + Z * Z) * math.atan(X / math.sqrt(1 + Z * Z))) / (2 * math.pi)
        return math.copysign(1, x) * math.copysign(1, z) * f
    px, pz = p[0], p[2]
    return (corner(x1 - px, z1 - pz) - corner(x0 - px, z1 - pz) - corner(x1 - px, z0 - pz)
            + corner(x0 - px, z0 - pz))


def test_estimator_direct_samples_area_light_closed_form():
    """NEE from a large square light with 1 and 4 light samples per vertex
    (DirectIllumination.cpp:98-133, noDirectSamples): the mean radiance of a
    floor point is kd * Le * F, F the point's configuration factor to the
    light (closed form), and 4 stratified samples cut the variance well
    below a quarter of one sample's."""
    side, h, Le, kd = 1.0, 0.5, 3.0, 0.5
    arrays = _floor_scene([0, 0, 0], light_size=side, light_h=h, Le=Le, kd=kd)
    # a low camera (below the light) looking at the floor under the light
    arrays.camera = {"eye": [0.0, 0.3, -2.0], "look": [0.0, -0.3, 2.0], "up": [0.0, 1.0, 0.0],
                     "height_angle": 20.0, "aspect": 1.0}
    osc = po.OracleScene(arrays)
    var = {}
    for n in (1, 4):
        job = m.make_job(arrays.camera, 8, 8, 4096, 1, seed=11, direct_samples=n)
        for pixel in (27, 36):
            pix = np.full(4096, pixel, np.uint32)
            smp = np.arange(4096, dtype=np.uint32)
            L = osc.render_paths(job, pix, smp)[:, 0]
            expect = []
            for k in range(4096):
                ray = po.camera_ray(job, pixel, k)
                o, d = ray[:3], ray[3:]
                pt = o + (-o[1] / d[1]) * d
                expect.append(kd * Le * _form_factor_rect(pt, h, -side / 2, side / 2, -side / 2, side / 2))
            expect = np.array(expect)
            sigma = L.std() / math.sqrt(len(L))
            assert abs(L.mean() - expect.mean()) <= 4 * sigma + 1e-3 * expect.mean(), (n, pixel)
            # the variance the light sampling adds on top of the pixel footprint's
            var[(n, pixel)] = np.var(L - expect)
    for pixel in (27, 36):
        assert var[(4, pixel)] < var[(1, pixel)] / 4


def test_three_sigma_oracle_self_consistency(room):
    """Two independent seeds of the oracle agree within 3 sigma per pixel on
    >= 98% of pixels (the statistical test used for the device is sound)."""
    arrays, osc = room
    job1 = m.make_job(arrays.camera, 24, 24, 64, 5, seed=1)
    job2 = m.make_job(arrays.camera, 24, 24, 64, 5, seed=2)
    s1, q1, _ = osc.render(job1)
    s2, q2, _ = osc.render(job2)
    n = 64
    m1, m2 = s1 / n, s2 / n
    v1 = (q1 / n - m1 ** 2) * n / (n - 1) / n
    v2 = (q2 / n - m2 ** 2) * n / (n - 1) / n
    ok = (np.abs(m1 - m2) <= 3 * np.sqrt(v1 + v2)) | (np.abs(m1 - m2) < 1e-9)
    assert ok.all(axis=1).mean() >= 0.98
