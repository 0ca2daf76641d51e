"""The device kd-tree and traversal, run on host threads (mpt_kd_build /
mpt_kd_trace), against the oracle's kd-tree and its brute-force loop.

The host entry points compile the same builder, record layout, traversal
(short stack with restart) and watertight intersection the kernels use, so
these checks pin the tree and the traversal logic without a device; the GPU
suites then only have to show the device runs them the same way.  Criteria
as for the device: every disagreement with the oracle's closest primitive is
an exact tie, and t agrees to 1e-5 relative.
"""
import numpy as np
import pytest

import milton_amd as m
from oracle import pyoracle as po
from conftest import classify_hits, rays64


def _check(osc, rays, hits, max_non_tie=1e-4):
    ot, oi = osc.trace_closest(rays64(rays))
    match, non_tie, rel = classify_hits(osc, rays, hits["t"], hits["prim"], ot, oi)
    assert non_tie <= max_non_tie, (match, non_tie)
    assert len(rel) == 0 or (rel <= 1e-5).mean() >= 0.99999, rel.max()
    return match


@pytest.fixture(scope="module")
def soup50k():
    return m.generate(m.MPT_SCENE_SOUP, 50_000, 1)


def test_host_trace_soup_vs_oracle(soup50k):
    tree = m.HostKdTree(soup50k)
    rays = m.generate_rays(m.MPT_SCENE_SOUP, 50_000, 100_000, seed=3)
    hits, ctr = tree.trace_closest(rays)
    assert _check(po.OracleScene(soup50k), rays, hits) > 0.9999
    assert ctr["nodes"] > 0 and ctr["refs"] > 0


def test_host_trace_room_vs_oracle():
    a = m.generate(m.MPT_SCENE_ROOM, 0, 1)
    tree = m.HostKdTree(a)
    rays = m.generate_rays(m.MPT_SCENE_ROOM, 0, 50_000, seed=5)
    hits, _ = tree.trace_closest(rays)
    _check(po.OracleScene(a), rays, hits, max_non_tie=0.0)


def test_host_trace_terrain_vs_oracle():
    a = m.generate(m.MPT_SCENE_TERRAIN, 256, 1)
    tree = m.HostKdTree(a)
    rays = m.generate_rays(m.MPT_SCENE_TERRAIN, 256, 50_000, seed=9)
    hits, _ = tree.trace_closest(rays)
    _check(po.OracleScene(a), rays, hits)


def test_host_trace_brute_force():
    a = m.generate(m.MPT_SCENE_SOUP, 3_000, 4)
    tree = m.HostKdTree(a)
    rays = m.generate_rays(m.MPT_SCENE_SOUP, 3_000, 20_000, seed=4)
    hits, _ = tree.trace_closest(rays)
    _check(po.OracleScene(a, naive=True), rays, hits, max_non_tie=0.0)


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_leaf_size_and_clipping_keep_hits(soup50k, leaf):
    """The autotuner's leaf-size range and the clipped/unclipped builders give
    the same closest hits; clipping only removes references."""
    rays = m.generate_rays(m.MPT_SCENE_SOUP, 50_000, 30_000, seed=6)
    out = {}
    for no_clip in (0, 1):
        kp = m.KdParams()
        kp.leaf_size = leaf
        kp.no_clip = no_clip
        tree = m.HostKdTree(soup50k, kp)
        hits, ctr = tree.trace_closest(rays)
        out[no_clip] = (hits, ctr, tree.stats())
    (h0, c0, s0), (h1, c1, s1) = out[0], out[1]
    osc = po.OracleScene(soup50k)
    _check(osc, rays, h0)
    _check(osc, rays, h1)
    assert c0["refs"] <= c1["refs"]
    assert s0["max_leaf"] >= 1 and s0["max_depth"] <= 48


def test_max_depth_and_stats(soup50k):
    kp = m.KdParams()
    kp.max_depth = 6
    tree = m.HostKdTree(soup50k, kp)
    st = tree.stats()
    assert st["max_depth"] <= 6
    assert st["leaves"] <= 2 ** 6 and st["nodes"] == 2 * st["leaves"] - 1
    rays = m.generate_rays(m.MPT_SCENE_SOUP, 50_000, 5_000, seed=8)
    hits, _ = tree.trace_closest(rays)
    _check(po.OracleScene(soup50k), rays, hits)


def test_invalid_scene_rejected():
    a = m.generate(m.MPT_SCENE_ROOM, 0, 1)
    a.tri_material = np.full_like(a.tri_material, 999)
    with pytest.raises(m.MptError):
        m.HostKdTree(a)
