import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def ctx():
    import milton_amd as m
    c = m.Context(0)
    yield c
    c.close()


def rays64(rays):
    """Structured float32 rays -> float64 (n, 6) for the oracle."""
    return np.concatenate([rays["o"], rays["d"]], axis=1).astype(np.float64)


def classify_hits(oracle_scene, rays, gpu_t, gpu_id, orc_t, orc_id):
    """Compares closest hits; returns (match fraction, non-tie mismatch
    fraction, relative t errors of matching hits).

    A disagreement counts as an exact tie when the double-precision t of the
    primitive the device chose equals the oracle's closest t to 1e-6 relative
    (shared edges and coincident faces)."""
    r = rays64(rays)
    same = gpu_id == orc_id
    diff = np.nonzero(~same)[0]
    non_tie = 0
    if len(diff):
        both_hit = (gpu_id[diff] != 0xFFFFFFFF) & (orc_id[diff] != 0xFFFFFFFF)
        t_alt = np.full(len(diff), np.inf)
        idx = diff[both_hit]
        if len(idx):
            t_alt[both_hit] = oracle_scene.prim_t(r[idx], gpu_id[idx])
        with np.errstate(invalid="ignore"):
            tie = both_hit & (np.abs(t_alt - orc_t[diff]) <= 1e-6 * np.maximum(orc_t[diff], 1e-3))
        non_tie = int((~tie).sum())
    hit = same & (orc_id != 0xFFFFFFFF)
    rel = np.abs(gpu_t[hit].astype(np.float64) - orc_t[hit]) / orc_t[hit]
    n = len(gpu_id)
    return same.mean(), non_tie / n, rel


def _tri_fp64(arrays, prims, rays):
    """Moller-Trumbore quantities of triangles `prims` for `rays` in double
    precision (shapes/MeshTriangle.cpp:142-167 without its tests): the
    determinant, the barycentric u, v and the distance t."""
    V = arrays.vertices.astype(np.float64)
    idx = arrays.indices[prims].astype(np.int64)
    a, b, c = V[idx[:, 0]], V[idx[:, 1]], V[idx[:, 2]]
    o, d = rays[:, :3], rays[:, 3:]
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = o - a
        u = np.einsum("ij,ij->i", s, p) / det
        q = np.cross(s, e1)
        v = np.einsum("ij,ij->i", d, q) / det
        t = np.einsum("ij,ij->i", e2, q) / det
    return det, u, v, t


def _edge_distance(o, d, V):
    """Smallest 3D distance between the ray's line and the edges of the
    triangle V (3 x 3)."""
    d = d / np.linalg.norm(d)
    best = np.inf
    for k in range(3):
        P, e = V[k], V[(k + 1) % 3] - V[k]
        a = (P - o) - ((P - o) @ d) * d   # components across the ray
        b = e - (e @ d) * d
        w = 0.0 if b @ b == 0 else float(np.clip(-(a @ b) / (b @ b), 0.0, 1.0))
        best = min(best, float(np.linalg.norm(a + w * b)))
    return best


def _silhouette_distance(arrays, k, ns, ray):
    """Distance between a ray and the outline of shape k of the spheres
    (k < ns) and quads (k >= ns) that follow the triangles: for a sphere
    | distance of its centre from the ray's line - radius |, for a quad the
    distance from the ray's line to its edges."""
    o, d = ray[:3], ray[3:] / np.linalg.norm(ray[3:])
    if k < ns:
        c, rad = arrays.spheres[k, :3].astype(np.float64), float(arrays.spheres[k, 3])
        f = c - o
        return abs(float(np.linalg.norm(f - (f @ d) * d)) - rad)
    q = arrays.quads[k - ns].astype(np.float64)
    c, ex, ez = q[:3], q[3:6], q[6:]
    corners = [c - ex / 2 - ez / 2, c + ex / 2 - ez / 2, c + ex / 2 + ez / 2, c - ex / 2 + ez / 2]
    return min(_edge_distance(o, d, np.array([corners[j], corners[(j + 1) % 4], corners[j]])) for j in range(4))


def explain_mismatches(arrays, oracle_scene, rays, gpu_t, gpu_id, orc_t, orc_id):
    """Sorts every closest-hit disagreement between the device and the
    oracle into the cases the two precisions are allowed to resolve
    differently, and counts what is left unexplained.

    Disputed primitives are the device's and the oracle's choices.  Classes,
    checked in this order:
      tie     the device's primitive is hit at the oracle's t (1e-6 relative);
      det     a disputed triangle's determinant |d . (e1 x e2)| lies within a
              factor 4 of the reference's EQ(det, 0) threshold 1e-7
              (MeshTriangle.cpp:144): FP32 and FP64 may fall on either side;
      t_eps   a disputed hit lies at t <= 1e-6, next to the t > 1e-7 rule
              (accel/kdTreeAccel.cpp:988);
      edge    the ray passes within 1e-6 (about 8 FP32 ulps of unit
              coordinates) of a disputed triangle's boundary in 3D: inside in
              one precision, outside in the other (u, v >= 0 and u + v <= 1 are
              inclusive on both sides; for a grazing triangle a tiny distance
              to the edge is a large barycentric one);
      tangent a disputed sphere's distance from the ray's line equals its
              radius to 1e-6 (the quadratic's discriminant changes sign
              between precisions), or the ray passes within 1e-6 of a
              disputed quad's boundary;
      kd_eps  the two hits are within 1e-6 of each other in t: the oracle's
              kd walk accepts a hit up to EPSILON past a cell's exit and pops
              with tMax - 2 EPSILON (kdTreeAccel.cpp:912-913, 999-1014).
    Returns (counts per class, indices of unexplained rays)."""
    r = rays64(rays)
    miss = 0xFFFFFFFF
    diff = np.nonzero(gpu_id != orc_id)[0]
    nt = len(arrays.indices)
    counts = {"tie": 0, "det": 0, "t_eps": 0, "edge": 0, "tangent": 0, "kd_eps": 0}
    ns = len(arrays.spheres)
    left = []
    for i in diff:
        g, o = int(gpu_id[i]), int(orc_id[i])
        tg = np.inf
        if g != miss:
            tg = float(oracle_scene.prim_t(r[i:i + 1], np.array([g], np.uint32))[0])
        to = float(orc_t[i]) if o != miss else np.inf
        if g != miss and o != miss and abs(tg - to) <= 1e-6 * max(to, 1e-3):
            counts["tie"] += 1
            continue
        tris = [p for p in (g, o) if p != miss and p < nt]
        if tris:
            det, u, v, t = _tri_fp64(arrays, np.array(tris), np.repeat(r[i:i + 1], len(tris), axis=0))
            if np.any((np.abs(det) <= 4e-7) & (np.abs(det) >= 2.5e-8)):
                counts["det"] += 1
                continue
            if np.any(np.abs(t) <= 1e-6) or min(tg, to) <= 1e-6:
                counts["t_eps"] += 1
                continue
            scale = max(1.0, float(np.abs(r[i, :3]).max()))
            if min(_edge_distance(r[i, :3], r[i, 3:], arrays.vertices[arrays.indices[p]].astype(np.float64))
                   for p in tris) <= 1e-6 * scale:
                counts["edge"] += 1
                continue
        others = [p for p in (g, o) if p != miss and p >= nt]
        if others and min(_silhouette_distance(arrays, p - nt, ns, r[i]) for p in others) <= \
                1e-6 * max(1.0, float(np.abs(r[i, :3]).max())):
            counts["tangent"] += 1
            continue
        if np.isfinite(tg) and np.isfinite(to) and abs(tg - to) <= 1e-6:
            counts["kd_eps"] += 1
            continue
        left.append(int(i))
    return counts, left
