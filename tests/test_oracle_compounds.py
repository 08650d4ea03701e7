"""Compound and decorated colliders in the oracle: the restatement's placement of each leaf
(CollectTransformedShapes: bounds tests against the soft body's local box, moves into and
rotations of each child) against the reference's own CollectTransformedShapes
(CompoundShapeVisitors.h:226-272, StaticCompoundShape.cpp:543-570, MutableCompoundShape.cpp:443-470,
RotatedTranslatedShape.cpp:169-176, OffsetCenterOfMassShape.cpp:135-142), and the leaf
descriptions the Jolt-side converter makes of them.
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from joltphysics_amd.records import OP_CULL, OP_OFFSET, OP_ROTATE, OP_TRANSLATE, SHAPE_BOX, SHAPE_SPHERE, SHAPE_TAPERED_CAPSULE, \
    ShapeSpec


def _worlds(colliders):
    from oracle import JoltReference, JoltWorld, Restatement
    w = JoltWorld(JoltReference("nofma"), (0.0, 0.0, -9.81))
    ids = [w.add_collider(c) for c in colliders]
    bounds = [w.collider_bounds(i) for i in ids]
    for c, i in zip(colliders, ids):
        c.device_params, c.com = w.collider_device(i)
        c.parts = w.collider_parts(i)
    r = Restatement((0.0, 0.0, -9.81), 32)
    for c, b in zip(colliders, bounds):
        r.add_collider(c, b)
    return w, r, ids


@pytest.fixture(scope="module")
def compound_worlds():
    cols = scene_zoo._compound_colliders()
    w, r, ids = _worlds(cols)
    yield cols, w, r, ids
    w.close()
    r.close()


def _random_pose(rng, identity):
    pos = rng.uniform(-0.6, 0.6, 3).astype(np.float32)
    q = np.array([0, 0, 0, 1], np.float32)
    if not identity:
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
    centre, half = rng.uniform(-1.0, 1.0, 3), rng.uniform(0.02, 1.2, 3)
    return pos, q, np.concatenate([centre - half, centre + half]).astype(np.float32)


def test_leaf_transforms_and_culling_match_reference(compound_worlds):
    """Bit-equal leaf transforms, in the reference's order, and the same leaves culled, over
    random soft-body poses and local boxes (small boxes cull most leaves, large ones none)."""
    cols, w, r, ids = compound_worlds
    rng = np.random.default_rng(7)
    culled = kept = 0
    for trial in range(400):
        pos, q, box = _random_pose(rng, trial % 3 == 0)
        for c, i in zip(cols, ids):
            want = w.collect_leaves(i, pos, q, box)
            got = r.collect_leaves(i, pos, q, box)
            assert len(got) == len(want), f"collider {i} trial {trial}: {len(got)} leaves, reference {len(want)}"
            for k, (m, g) in enumerate(zip(want, got)):
                assert np.array_equal(m[0], g), f"collider {i} trial {trial} leaf {k}: transforms differ"
            culled += len(c.parts) - len(want)
            kept += len(want)
    assert culled > 100 and kept > 100, (culled, kept)


def test_leaf_scales_are_the_reference_ones(compound_worlds):
    """With a box that culls nothing, the reference collects every leaf; its scale is the one
    the converter hands the device (scaled compounds, rotated children of uniform scale,
    identity-rotation children of a non-uniform, mirroring scale)."""
    cols, w, r, ids = compound_worlds
    box = np.array([-50, -50, -50, 50, 50, 50], np.float32)
    for c, i in zip(cols, ids):
        leaves = w.collect_leaves(i, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), box)
        assert len(leaves) == len(c.parts)
        for (m, scale), part in zip(leaves, c.parts):
            assert np.array_equal(np.float32(scale), np.float32(part.scale)), (i, scale, part.scale)


def test_far_box_collects_nothing_and_the_body_is_left_out(compound_worlds):
    """A local box away from every child culls every leaf of a compound (the reference then
    leaves the body out); a decorated single shape has no bounds test and is always kept."""
    cols, w, r, ids = compound_worlds
    box = np.array([40, 40, 40, 41, 41, 41], np.float32)
    for c, i in zip(cols, ids):
        n = len(w.collect_leaves(i, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), box))
        has_cull = any((p.ops["kind"] == OP_CULL).any() for p in c.parts)
        assert n == (0 if has_cull else len(c.parts)), (i, n)
        assert len(r.collect_leaves(i, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), box)) == n


def test_static_compound_tree_gives_one_bounds_step_per_level():
    """A static compound of 25 children has a tree of 4-wide nodes at least two levels deep:
    every leaf carries one bounds step per node on its way down (its own entry in its node
    included), then the move and the rotation into the child; the mutable compound carries one
    bounds step per child, a rotated-translated shape only a rotation, an offset centre of mass
    only the offset."""
    cols = scene_zoo._compound_colliders(sensor=False)
    w, r, ids = _worlds(cols)
    try:
        kinds = [[tuple(int(k) for k in p.ops["kind"]) for p in c.parts] for c in cols]
        assert len(kinds[0]) == 25
        for k in kinds[0]:
            depth = len(k) - 2
            assert depth >= 2 and k == (OP_CULL,) * depth + (OP_TRANSLATE, OP_ROTATE), k
        assert all(k == (OP_CULL, OP_TRANSLATE, OP_ROTATE) for k in kinds[1]) and len(kinds[1]) == 5
        assert kinds[2] == [(OP_ROTATE,)]
        assert kinds[3] == [(OP_OFFSET,)]
    finally:
        w.close()
        r.close()


def test_static_compound_walk_visits_children_last_first():
    """Four spheres in a static compound (one tree node): the reference collects them last child
    first (StaticCompoundShape::WalkTree pops the node's children from a stack), and so does the
    converter's order (checked on the leaf positions)."""
    S = ShapeSpec
    from joltphysics_amd.records import SHAPE_COMPOUND, Collider
    from joltphysics_amd.scenes import collider_world_bounds
    kids = [(S.basic(SHAPE_SPHERE, 0.1), (float(x), 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)) for x in (-0.6, -0.2, 0.2, 0.6)]
    c = Collider(SHAPE_COMPOUND, (), shape=S.static_compound(kids))
    c.bounds = collider_world_bounds(c)
    w, r, ids = _worlds([c])
    try:
        leaves = w.collect_leaves(0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), np.array([-9, -9, -9, 9, 9, 9], np.float32))
        xs = [float(m[0][0, 3]) for m in leaves]
        assert xs == sorted(xs, reverse=True), xs
        got = r.collect_leaves(0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), np.array([-9, -9, -9, 9, 9, 9], np.float32))
        assert [float(m[0, 3]) for m in got] == xs
    finally:
        w.close()
        r.close()


def test_oriented_box_test_against_brute_force():
    """The bounds test is a separating-axis test: where it says "apart", a dense sampling of the
    soft body's box (seen from the compound) finds no point inside the child's box, and where
    the sampling finds one, it says "overlap".  A box child, random soft-body
    poses and boxes (rotated, so every one of the 15 axes can separate); a second child far
    away never passes."""
    S = ShapeSpec
    from joltphysics_amd.records import SHAPE_COMPOUND, Collider, _rotate
    from joltphysics_amd.scenes import collider_world_bounds
    kids = [(S.basic(SHAPE_BOX, 0.3, 0.2, 0.1), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)),
            (S.basic(SHAPE_SPHERE, 0.01), (5.0, 5.0, 5.0), (0.0, 0.0, 0.0, 1.0))]
    c = Collider(SHAPE_COMPOUND, (), shape=S.mutable_compound(kids))
    c.bounds = collider_world_bounds(c)
    w, r, ids = _worlds([c])
    try:
        bmin, bmax = c.parts[0].ops["v"][0][:3], c.parts[0].ops["v"][0][3:]
        rng = np.random.default_rng(11)
        g = np.linspace(0.0, 1.0, 9)
        grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        seen = {True: 0, False: 0}
        for _ in range(300):
            pos, q, box = _random_pose(rng, False)
            pos = pos * 0.8
            # the far sphere never passes, so one leaf means the box child passed
            kept = len(r.collect_leaves(0, pos, q, box)) == 1
            assert kept == (len(w.collect_leaves(0, pos, q, box)) == 1)
            # sample points of the soft body's box, into the compound's frame
            pts = box[:3] + grid * (box[3:] - box[:3])
            com = np.array(c.com, np.float64)
            world = np.array([_rotate(q, p) for p in pts]) + pos
            local = np.array([_rotate(np.array([-c.rot[0], -c.rot[1], -c.rot[2], c.rot[3]]), p - com) for p in world])
            inside = ((local >= bmin - 1e-6) & (local <= bmax + 1e-6)).all(axis=1).any()
            if inside:
                assert kept
            seen[kept] += 1
        assert seen[True] > 20 and seen[False] > 20, seen
    finally:
        w.close()
        r.close()


def test_tapered_capsule_parameters_are_the_shapes_own():
    """The converter reads a tapered capsule's centres, radii, sin and tan from the shape's saved
    state: the centres are the constructor's (TaperedCapsuleShape.cpp:117-131), half height
    + / - the centre offset, which GetHalfHeight() (rebuilt from the centres) can miss by an
    ulp; the cone's sin is (bottom - top) / (top centre - bottom centre)."""
    from joltphysics_amd.records import Collider
    f = np.float32
    for hh, top, bottom in ((0.06, 0.03, 0.06), (0.2, 0.12, 0.06), (0.37, 0.011, 0.29), (0.1, 0.2, 0.05)):
        c = Collider(SHAPE_TAPERED_CAPSULE, (hh, top, bottom))
        c.bounds = (-1, -1, -1, 1, 1, 1)
        w, r, ids = _worlds([c])
        try:
            p = np.float32(c.device_params)
            off = f(0.5) * (f(bottom) - f(top))
            assert p[0] == f(hh) + off and p[1] == -f(hh) + off, (p, hh, top, bottom)
            assert p[2] == f(top) and p[3] == f(bottom)
            assert p[4] == (f(bottom) - f(top)) / (p[0] - p[1])
            assert abs(float(p[5]) - np.tan(np.arcsin(float(p[4])))) < 1e-6
        finally:
            w.close()
            r.close()
