"""The EmitterSampler in the oracle (stats/samplers/EmitterSampler.cpp:19-96)
and next-event estimation with lights drawn by power (mpt_job.light_select),
pinned by what the sampler's definition fixes:

  * a light is drawn with probability power / total power, where the test
    computes each power from the scene arrays itself (Le pi A for an area
    light, A from the quad's edges or the mesh's triangles, pi for a sphere;
    the given P for a point light): frequencies within 5 sigma;
  * pdf_area is 1 / (sum of the lights' areas) for every sample (getPd);
  * a quad's points lie on the quad, uniformly (Plane::getPosition at two
    uniforms); a mesh's points pick each triangle with frequency 1 / n
    whatever its area (Mesh::getPosition's triangle index) and lie on it;
    a sphere's points are NaN exactly when the uniform point's |s_y| > 1/2
    (asin's domain in Sphere::_getUV), half of them by Archimedes' theorem,
    and lie at object-space distance 1 from the centre otherwise;
  * with one light, drawing by power gives the same paths as sampling every
    light; with two lights and one sample per vertex at depth 1, every path
    carries exactly one light's contribution (as rendered with that light
    alone) over its probability, the brighter light for a share p_A of the
    paths; with 16 samples per vertex the estimate agrees per pixel with
    sampling every light (room with a second light and a point light).
"""
import math

import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from test_area_lights import CAMERA, FLOOR, KD, _square_mesh


def _mats(les):
    mats = np.zeros(1 + len(les), m.MATERIAL_DTYPE)
    mats[0]["type"] = m.MPT_BSDF_DIFFUSE
    mats[0]["kd"] = KD
    for i, le in enumerate(les):
        mats[1 + i]["type"] = m.MPT_BSDF_DIFFUSE
        mats[1 + i]["Le"] = le
    return mats


def _down_quad(c, a, b):
    """A quad facing -y with edges a (along z) and b (along x): normal ez x ex."""
    return [*c, 0, 0, a, b, 0, 0]


def _lights_scene():
    """Two quads, a two-triangle mesh, a turned sphere and a point light."""
    les = [[4.0, 4.0, 4.0], [1.0, 2.0, 6.0], [3.0, 3.0, 0.0], [2.0, 1.0, 1.0]]
    quads = [FLOOR[0], _down_quad([0.2, 1.0, 0.1], 0.3, 0.2), _down_quad([-0.4, 1.2, 0.3], 0.1, 0.5)]
    e = 0.2
    c0, c1 = np.array([0.5, 0.9, -0.2]), np.array([-0.3, 0.8, -0.4])
    v = np.array([c0, c0 + [e, 0, 0], c0 + [0, 0, e], c1, c1 + [2 * e, 0, 0], c1 + [0, 0, 2 * e]], np.float64)
    idx = [[0, 1, 2], [3, 4, 5]]
    frame = 0.3 * np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], np.float64)
    a = m.SceneArrays(v, idx, [3, 3], [[0.0, 1.5, 0.0, 0.15]], [4], quads, [0, 1, 2], _mats(les), [0, 0, 0],
                      tri_shape=[5, 5], sphere_frames=frame.reshape(1, 9),
                      points=[[0.3, 0.6, 0.2, 2.0, 3.0, 4.0]])
    a.camera = CAMERA
    # powers from the arrays: emitting quads (ez x ex areas), the sphere
    # (area pi whatever its size), the mesh (its triangles), the point
    q = np.asarray(quads, np.float64)
    area_q = [np.linalg.norm(np.cross(q[i, 6:9], q[i, 3:6])) for i in (1, 2)]
    area_m = 0.5 * e * e + 0.5 * 4 * e * e
    areas = area_q + [math.pi, area_m]
    le_mean = [np.mean(les[0]), np.mean(les[1]), np.mean(les[3]), np.mean(les[2])]   # quads, sphere, mesh
    power = [lm * math.pi * ar for lm, ar in zip(le_mean, areas)] + [3.0]
    return a, np.array(power) / sum(power), sum(areas)


def test_selection_frequencies_and_pdfs():
    a, share, area = _lights_scene()
    n = 400_000
    s = po.OracleScene(a).emitter_sample(n, seed=5)
    counts = np.bincount(s["light"], minlength=len(share))
    assert len(counts) == len(share)
    sigma = np.sqrt(n * share * (1 - share))
    print("counts", counts, "expected", n * share)
    assert np.all(np.abs(counts - n * share) < 5 * sigma)
    np.testing.assert_allclose(s["pdf_select"], share[s["light"]], rtol=1e-6)
    np.testing.assert_allclose(s["pdf_area"], 1.0 / area, rtol=1e-6)


def test_points_on_the_lights():
    a, share, _ = _lights_scene()
    n = 200_000
    s = po.OracleScene(a).emitter_sample(n, seed=9)
    p, nrm, li = s["p"], s["n"], s["light"]
    q = a.quads.astype(np.float64)
    # quads (lights 0, 1 = quads 1, 2): on the plane, inside, uniform
    for k, qi in ((0, 1), (1, 2)):
        sel = li == k
        c, ex, ez = q[qi, :3], q[qi, 3:6], q[qi, 6:9]
        d = p[sel] - c
        uu = d @ ex / (ex @ ex)
        vv = d @ ez / (ez @ ez)
        np.testing.assert_allclose(d, np.outer(uu, ex) + np.outer(vv, ez), atol=1e-12)
        assert np.all(np.abs(uu) <= 0.5) and np.all(np.abs(vv) <= 0.5)
        for x in (uu, vv):   # uniform on [-1/2, 1/2]: mean 0, variance 1/12
            assert abs(x.mean()) < 5 * math.sqrt(1 / 12 / sel.sum())
            assert abs(x.var() - 1 / 12) < 0.01
        np.testing.assert_allclose(nrm[sel], np.tile(np.cross(ez, ex) / np.linalg.norm(np.cross(ez, ex)),
                                                     (sel.sum(), 1)), atol=1e-12)
    # the mesh (light 3): each triangle half the time, on its plane y = const
    sel = li == 3
    on_first = np.abs(p[sel, 1] - 0.9) < 1e-6
    on_second = np.abs(p[sel, 1] - 0.8) < 1e-6
    assert np.all(on_first | on_second)
    f = on_first.mean()
    assert abs(f - 0.5) < 5 * math.sqrt(0.25 / sel.sum())
    # the sphere (light 2): NaN for half the draws, else at object-space
    # distance 1 (frame columns are the images of the object axes)
    sel = li == 2
    bad = np.isnan(p[sel]).any(axis=1)
    assert abs(bad.mean() - 0.5) < 5 * math.sqrt(0.25 / sel.sum())
    frame = a.sphere_frames.reshape(3, 3).astype(np.float64)          # rows: images of x, y, z
    obj = np.linalg.solve(frame.T, (p[sel][~bad] - a.spheres[0, :3]).T).T
    np.testing.assert_allclose(np.linalg.norm(obj, axis=1), 1.0, rtol=1e-9)
    assert np.all(obj[:, 2] >= math.sqrt(0.5) - 1e-9)   # |x|, |y| <= 1/2 puts the point near +z
    # the point light (light 4): its position, no normal
    sel = li == 4
    np.testing.assert_allclose(p[sel], np.tile(a.points[0, :3], (sel.sum(), 1)), rtol=1e-7)
    assert np.all(nrm[sel] == 0)


def test_no_lights_is_an_error():
    a = m.SceneArrays(np.zeros((0, 3)), np.zeros((0, 3)), [], np.zeros((0, 4)), [], FLOOR, [0], _mats([]),
                      [1, 1, 1])
    with pytest.raises(ValueError):
        po.OracleScene(a).emitter_sample(10, 1)


def _paths(a, select, direct, side=16, spp=2, depth=1):
    job = m.make_job(a.camera, side, side, spp, depth, seed=3)
    job.direct_samples = direct
    job.light_select = select
    pix = np.repeat(np.arange(side * side, dtype=np.uint32), spp)
    smp = np.tile(np.arange(spp, dtype=np.uint32), side * side)
    return job, pix, smp, po.OracleScene(a).render_paths(job, pix, smp)


def test_one_light_same_paths():
    a = m.generate(m.MPT_SCENE_ROOM)
    for direct in (1, 4):
        L0 = _paths(a, 0, direct, depth=5)[3]
        L1 = _paths(a, 1, direct, depth=5)[3]
        assert np.array_equal(L0, L1)
        assert L0.mean() > 0.05


def _two_lights(which=(0, 1)):
    s = 0.05
    centres = [np.array([0.1, 1.0, 0.2]), np.array([-0.3, 0.8, -0.1])]
    les = [[5.0, 5.0, 5.0], [1.0, 1.0, 1.0]]
    v, idx, tm, ts, mats = [], [], [], [], []
    for k in which:
        vk, ik = _square_mesh(centres[k], s)
        idx.append(ik + 4 * len(v))
        v.append(vk)
        tm += [1 + len(mats)] * 2
        ts += [len(mats)] * 2
        mats.append(les[k])
    a = m.SceneArrays(np.concatenate(v), np.concatenate(idx), tm, np.zeros((0, 4)), [], FLOOR, [0], _mats(mats),
                      [0, 0, 0], tri_shape=ts)
    a.camera = CAMERA
    return a


def test_two_lights_drawn_by_power():
    """One light sample per vertex at depth 1: each path carries exactly the
    contribution of light A alone over p_A, or of light B alone over p_B
    (the same random stream samples the light in both scenes), p_A = 5/6
    from the powers 5 : 1 of two equal squares, and A is drawn for a share
    p_A of the paths."""
    both = _two_lights()
    _, _, _, L = _paths(both, 1, 1, side=32, spp=4)
    L_a = _paths(_two_lights((0,)), 0, 1, side=32, spp=4)[3]
    L_b = _paths(_two_lights((1,)), 0, 1, side=32, spp=4)[3]
    p_a = 5.0 / 6.0   # up to the float32 vertices of the two squares
    is_a = np.all(np.abs(L - L_a / p_a) <= 1e-6 * np.abs(L_a / p_a) + 1e-15, axis=1)
    is_b = np.all(np.abs(L - L_b / (1 - p_a)) <= 1e-6 * np.abs(L_b / (1 - p_a)) + 1e-15, axis=1)
    lit = (L_a[:, 0] > 0) & (L_b[:, 0] > 0) & (np.abs(L_a[:, 0] / p_a - L_b[:, 0] / (1 - p_a)) > 1e-6)
    assert lit.mean() > 0.9
    assert np.all(is_a | is_b)
    frac = is_a[lit].mean()
    print(f"light A drawn for {frac:.4f} of the paths (p_A {p_a:.4f})")
    assert abs(frac - p_a) < 5 * math.sqrt(p_a * (1 - p_a) / lit.sum())


def test_drawn_by_power_matches_every_light_on_average():
    """16 samples per vertex: per pixel, the mean over 16 paths of the
    estimate drawing lights by power agrees with sampling every light
    within 5 standard errors, on the room with a second light and a point
    light."""
    a = m.generate(m.MPT_SCENE_ROOM)
    a.points = np.array([[0.3, 0.7, 0.4, 0.6, 0.5, 0.4]], np.float32)
    q = a.quads[0].copy()
    q[:3] = [0.75, 0.999, 0.75]
    a.quads = np.concatenate([a.quads, q[None]])
    a.quad_material = np.concatenate([a.quad_material, a.quad_material[:1]])
    res = []
    for sel in (0, 1):
        job = m.make_job(a.camera, 24, 24, 64, 3, seed=8)
        job.direct_samples = 4
        job.light_select = sel
        s, s2, ctr = po.OracleScene(a).render(job)
        res.append((s / 64, np.sqrt(np.maximum(s2 / 64 - (s / 64) ** 2, 0) / 64), ctr))
    (m0, e0, c0), (m1, e1, c1) = res
    z = np.abs(m0 - m1) / np.sqrt(e0 ** 2 + e1 ** 2 + 1e-12)
    print(f"mean {m0.mean():.4f} vs {m1.mean():.4f}; shadow rays {c0['shadow']} vs {c1['shadow']}")
    assert np.mean(z < 3) > 0.97 and np.all(z < 6)
    assert abs(m0.mean() - m1.mean()) < 0.02 * m0.mean()
    assert c1["shadow"] < 0.5 * c0["shadow"]   # 4 rays per vertex instead of 2 x 4 + 1
