"""The fast traversal's SAH BVH builder (host code, no GPU): through mt_fast_bvh_build, on a
sphere, a terrain patch, a random soup, coincident triangles and a mesh below the leaf size.
Pinned by structure (every triangle in exactly one leaf of 1-4 slots, boxes tight over what
they hold, the root box = the mesh bounds) and by a float64 restatement of the fast traversal
(nearest entry first, culled at the closest hit so far) returning the brute-force closest hit."""
import ctypes as C

import numpy as np
import pytest

from mitrace_amd import scenegen
from mitrace_amd._lib import check, load

LEAF = 0x80000000
SLOT = 0x7FFFFFFF


def build(P):
    L = load()
    P = np.ascontiguousarray(P, np.float32).reshape(-1, 9)
    n = len(P)
    npair, nslot, root = C.c_uint32(), C.c_uint32(), C.c_uint32()
    fp = P.ctypes.data_as(C.POINTER(C.c_float))
    u32 = C.POINTER(C.c_uint32)
    check(L.mt_fast_bvh_build(fp, n, None, C.byref(npair), None, None, C.byref(nslot), None, None))
    pairs = np.zeros((npair.value, 16), np.float32)
    tris = np.zeros(nslot.value, np.uint32)
    last = np.zeros(nslot.value, np.uint32)
    box = np.zeros(6, np.float32)
    check(L.mt_fast_bvh_build(fp, n, pairs.ctypes.data_as(C.POINTER(C.c_float)), C.byref(npair),
                              tris.ctypes.data_as(u32), last.ctypes.data_as(u32), C.byref(nslot),
                              C.byref(root), box.ctypes.data_as(C.POINTER(C.c_float))))
    return pairs, tris, last, root.value, box


def words(pairs):
    w = pairs.view(np.uint32)
    return w[:, 3], w[:, 7]  # A.word, B.word


def walk(pairs, last, root, box):
    """(box, word, leaf slots) of every node, depth first"""
    wa, wb = words(pairs)
    out, stack = [], [(box[:3], box[3:], int(root))]
    while stack:
        mn, mx, w = stack.pop()
        if w & LEAF:
            s, slots = w & SLOT, []
            while True:
                slots.append(s)
                if last[s]:
                    break
                s += 1
            out.append((mn, mx, w, slots))
        else:
            r = pairs[w]
            out.append((mn, mx, w, None))
            stack.append((r[0:3], r[4:7], int(wa[w])))
            stack.append((r[8:11], r[12:15], int(wb[w])))
    return out


def meshes():
    rng = np.random.default_rng(7)
    ico = scenegen.icosphere(3).positions.reshape(-1, 9)
    ter = scenegen.terrain_mesh(40, 10.0, 3.0, rng)[0].positions.reshape(-1, 9)
    soup = (rng.random((1500, 3, 3)) * 10.0 + rng.normal(size=(1500, 1, 3)) * 0.3).reshape(-1, 9)
    same = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), (40, 1))
    small = rng.random((3, 9))
    return {"icosphere": ico, "terrain": ter, "soup": soup, "coincident": same, "three": small}


@pytest.mark.parametrize("name", ["icosphere", "terrain", "soup", "coincident", "three"])
def test_structure_and_tight_boxes(name):
    P = np.asarray(meshes()[name], np.float32).reshape(-1, 9)
    pairs, tris, last, root, box = build(P)
    T = P.reshape(-1, 3, 3)
    tmin, tmax = T.min(axis=1), T.max(axis=1)
    np.testing.assert_array_equal(box[:3], tmin.min(axis=0))
    np.testing.assert_array_equal(box[3:], tmax.max(axis=0))
    assert np.array_equal(np.sort(tris), np.arange(len(P))), "every triangle exactly once"
    nodes = walk(pairs, last, root, box)
    seen = np.zeros(len(tris), bool)
    for mn, mx, w, slots in nodes:
        if slots is None:
            continue
        assert 1 <= len(slots) <= 4
        assert not seen[slots].any()
        seen[slots] = True
        t = tris[slots]
        np.testing.assert_array_equal(mn, tmin[t].min(axis=0))  # tight: the union of its boxes
        np.testing.assert_array_equal(mx, tmax[t].max(axis=0))
    assert seen.all()
    # inner boxes are the union of their children's
    wa, wb = words(pairs)
    for mn, mx, w, slots in nodes:
        if slots is None:
            r = pairs[w]
            np.testing.assert_array_equal(mn, np.minimum(r[0:3], r[8:11]))
            np.testing.assert_array_equal(mx, np.maximum(r[4:7], r[12:15]))


def _tri_hit(o, d, v0, v1, v2):
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-12:
        return np.inf
    s = o - v0
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return t if (u >= 0 and v >= 0 and u + v <= 1 and t > 0) else np.inf


def _brute(o, d, P):
    """closest t over all triangles (vectorised form of _tri_hit)"""
    v0, e1, e2 = P[:, 0], P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    p = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, p)
    ok = np.abs(det) >= 1e-12
    det = np.where(ok, det, 1.0)
    s = o - v0
    u = np.einsum("ij,ij->i", s, p) / det
    q = np.cross(s, e1)
    v = q @ d / det
    t = np.einsum("ij,ij->i", e2, q) / det
    hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)
    return t[hit].min() if hit.any() else np.inf


def _enter(o, inv, mn, mx, tmax):
    with np.errstate(invalid="ignore"):
        t1, t2 = (mn - o) * inv, (mx - o) * inv
    tn = max(np.nanmax(np.minimum(t1, t2)), 0.0)
    tf = min(np.nanmin(np.maximum(t1, t2)), tmax)
    return tn if tn <= tf else np.inf


@pytest.mark.parametrize("name", ["icosphere", "terrain", "soup"])
def test_traversal_finds_the_brute_force_closest_hit(name):
    P = np.asarray(meshes()[name], np.float64).reshape(-1, 3, 3)
    pairs32, tris, last, root, box = build(P.astype(np.float32))
    wa, wb = words(pairs32)
    pairs = pairs32.astype(np.float64)
    rng = np.random.default_rng(3)
    lo, hi = P.reshape(-1, 3).min(0), P.reshape(-1, 3).max(0)
    c, ext = (lo + hi) / 2, hi - lo
    for _ in range(150):
        o = c + (rng.random(3) - 0.5) * ext * 1.5
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        with np.errstate(divide="ignore"):
            inv = 1.0 / d
        brute = _brute(o, d, P)
        best, visited = np.inf, 0
        stack = [(_enter(o, inv, box[:3], box[3:], np.inf), int(root))]
        while stack:
            te, w = stack.pop()
            if te >= best:
                continue
            visited += 1
            if w & LEAF:
                s = w & SLOT
                while True:
                    best = min(best, _tri_hit(o, d, *P[tris[s]]))
                    if last[s]:
                        break
                    s += 1
                continue
            r = pairs[w]
            da = _enter(o, inv, r[0:3], r[4:7], best)
            db = _enter(o, inv, r[8:11], r[12:15], best)
            kids = sorted([(da, int(wa[w])), (db, int(wb[w]))], reverse=True)  # nearer popped first
            stack.extend(k for k in kids if k[0] < best)
        if np.isinf(brute):
            assert np.isinf(best)
        else:
            assert best == pytest.approx(brute, rel=1e-9)
        assert visited <= 2 * len(pairs) + 2
