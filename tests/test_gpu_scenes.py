"""GPU parity on scenes covering every constraint type, collider shape (plane, sphere,
box, capsule, cylinder, tapered capsule, convex hull), the LDS-resident and
streamed storage modes, pressure, LRA, sleeping and batching.

The reference (compiled from its sources with FMA contraction off, loaded
through oracle/oracle.py) builds the settings with its own CreateConstraints
and Optimize and runs PhysicsSystem::Update; the HIP path runs the same
settings through the C ABI.  Bar: float-equal positions, velocities and body
positions at every checked frame.
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from helpers import float_equal, max_abs, reference_run, run_gpu
from joltphysics_amd.records import SHAPE_PLANE, _rotate

pytestmark = pytest.mark.gpu


def _same_plane(a, b):
    """Two plane colliders holding the same plane (translated within it)."""
    if a.type != SHAPE_PLANE or b.type != SHAPE_PLANE or tuple(a.params) != tuple(b.params) or a.rot != b.rot:
        return False
    n = _rotate(a.rot, a.params[:3])
    return abs(float(np.dot(n, np.subtract(a.pos, b.pos)))) == 0.0


def _check_contact_bodies(sc, got, want, what):
    """Equal collider per vertex.  Where two colliders hold the same plane a vertex touches
    both equally and the first in the list wins; the reference's list is in broadphase
    query order, which for such a pair may differ from creation order, so either one is
    accepted (positions and contact points are identical either way)."""
    assert ((got >= 0) == (want >= 0)).all(), f"{what}: contact flags differ"
    for k in np.nonzero(got != want)[0]:
        assert _same_plane(sc.colliders[got[k]], sc.colliders[want[k]]), \
            f"{what}: vertex {k} collider {got[k]} != {want[k]}"


@pytest.mark.parametrize("name", list(scene_zoo.ALL))
def test_scene_bit_exact(name):
    sc = scene_zoo.ALL[name]()
    frames = (1, 2, sc.frames)
    print(f"scene {name}: {len(sc.bodies)} bodies, frames {frames}", flush=True)
    settings, bounds, ref = reference_run(sc, frames, record_contacts=True)
    gpu = run_gpu(sc, settings, bounds=bounds, frames_to_keep=frames)
    for f in frames:
        for i in range(len(sc.bodies)):
            r, g = ref[f][i], gpu[f][i]
            dp, dv = max_abs(g["pos"], r["pos"]), max_abs(g["vel"], r["vel"])
            print(f"  frame {f} body {i}: max |dpos| {dp:.3e} max |dvel| {dv:.3e}", flush=True)
            assert float_equal(g["pos"], r["pos"]), f"{name}: positions differ at frame {f}, body {i}"
            assert float_equal(g["vel"], r["vel"]), f"{name}: velocities differ at frame {f}, body {i}"
            assert float_equal(g["body_pos"], r["body_pos"]), f"{name}: body position differs at frame {f}, body {i}"
            # the contact listener's manifold of the frame's last collision step
            _check_contact_bodies(sc, g["contact_collider"], r["contact_collider"], f"{name} frame {f} body {i}")
            assert (g["sensor_hits"] == r["sensor_hits"]).all(), f"{name}: sensor contacts differ at frame {f}, body {i}"
            hit = r["contact_collider"] >= 0
            assert float_equal(g["contact_point"][hit], r["contact_point"][hit]), \
                f"{name}: contact points differ at frame {f}, body {i}"


def test_convex_hull_tables_are_validated():
    """xpbd_set_colliders rejects hull tables that index outside themselves (a vertex
    index past the points, a face running past the index table) before any upload."""
    import numpy as np
    from joltphysics_amd.records import SHAPE_CONVEX_HULL, Collider, ConvexHull
    from joltphysics_amd.xpbd import Context, XpbdError
    cube = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    faces = [[0, 1, 3, 2], [4, 6, 7, 5], [0, 4, 5, 1], [2, 3, 7, 6], [0, 2, 6, 4], [1, 5, 7, 3]]
    bad = ConvexHull.from_polytope(cube, faces)
    bad.vertex_index = bad.vertex_index.copy()
    bad.vertex_index[3] = 8
    c = Collider(SHAPE_CONVEX_HULL, tuple(cube.ravel()), hull=bad, bounds=(-0.5, -0.5, -0.5, 0.5, 0.5, 0.5))
    ctx = Context(0)
    with pytest.raises(XpbdError, match="vertex index"):
        ctx.set_colliders([c])
    bad.vertex_index[3] = 3
    bad.face_count = bad.face_count.copy()
    bad.face_count[5] = 30
    with pytest.raises(XpbdError, match="face"):
        ctx.set_colliders([c])
    ctx.close()


def test_compound_parts_are_validated():
    """xpbd_set_colliders rejects compound leaves with an unknown placement step, a leaf shape
    type it does not collide with, or more steps than XPBD_MAX_PART_OPS."""
    from joltphysics_amd.records import PART_OP, SHAPE_COMPOUND, SHAPE_SPHERE, Collider, ColliderPart
    from joltphysics_amd.xpbd import Context, XpbdError
    ops = np.zeros(2, PART_OP)
    ops["kind"] = (2, 3)
    ops["v"][1][3] = 1.0
    part = ColliderPart(SHAPE_SPHERE, (0.2,), ops=ops)
    c = Collider(SHAPE_COMPOUND, (), parts=[part], bounds=(-1, -1, -1, 1, 1, 1))
    ctx = Context(0)
    try:
        ctx.set_colliders([c])
        part.ops = ops.copy()
        part.ops["kind"][0] = 9
        with pytest.raises(XpbdError, match="placement step"):
            ctx.set_colliders([c])
        part.ops = np.zeros(1025, PART_OP)
        part.ops["kind"] = 2
        with pytest.raises(XpbdError, match="XPBD_MAX_PART_OPS"):
            ctx.set_colliders([c])
        part.ops = ops
        part.type = 42
        with pytest.raises(XpbdError, match="not supported"):
            ctx.set_colliders([c])
    finally:
        ctx.close()
