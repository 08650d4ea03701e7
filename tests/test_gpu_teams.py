"""GPU parity of large bodies run by teams of workgroups.

* Exact mode: streamed scenes of the zoo with every streamed body on a team (one
  member per parallel-group window, the non-parallel group on member 0) are
  bit-exact with the reference (FMA off): positions, velocities, body positions and
  contact reports at every checked frame.
* Coloured mode (XPBD_TAIL_COLOURED): the non-parallel group re-coloured within each
  constraint type.  Its result is that of the reference run on settings whose
  non-parallel group lists its constraints colour by colour (a colour's constraints
  share no written vertex, so any order within a colour gives the same floats).  The
  colouring is restated here independently (greedy, lowest free colour, in the
  reference's order) and must give the library's colour count; the reference on the
  reordered settings must then equal the GPU bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from helpers import float_equal, max_abs, reference_run
from joltphysics_amd import scenes
from joltphysics_amd.xpbd import Context, build_settings

pytestmark = pytest.mark.gpu

TEAM_SCENES = ("cloth100_sphere", "cloth_hulls_streamed", "block_lra_boxes_streamed", "cloth_on_mesh_streamed",
               "cloth_on_height_field_streamed", "cloth_on_compounds_streamed")


def _run_team(sc, settings, bounds, frames, coloured=False, schedule="exact"):
    ctx = Context(0)
    try:
        ctx.set_gravity(sc.gravity)
        for c, b in zip(sc.colliders, bounds):
            c.bounds = tuple(float(x) for x in b)
        ctx.set_colliders(sc.colliders)
        ctx.set_team_min_windows(1)
        ids = [ctx.create_body(ctx.create_settings(s, schedule), b.position, b.rotation, b.params)
               for b, s in zip(sc.bodies, settings)]
        if coloured:
            for i in ids:
                ctx.set_tail_mode(i, True)
        out, done = {}, 0
        for f in sorted(frames):
            ctx.update(sc.dt, sc.collision_steps, f - done)
            done = f
            out[f] = [ctx.read(i) for i in ids]
            for i, r in zip(ids, out[f]):
                r["contact_collider"], r["contact_point"] = ctx.contacts(i)
                r["placement"] = ctx.placement(i)
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("name", TEAM_SCENES)
def test_team_exact_bit_exact(name):
    sc = scene_zoo.ALL[name]()
    frames = tuple(sorted({1, 2, sc.frames}))
    print(f"scene {name} on teams: frames {frames}", flush=True)
    settings, bounds, ref = reference_run(sc, frames, record_contacts=True)
    gpu = _run_team(sc, settings, bounds, frames)
    for f in frames:
        for i in range(len(sc.bodies)):
            r, g = ref[f][i], gpu[f][i]
            print(f"  frame {f} body {i} {g['placement']}: max |dpos| {max_abs(g['pos'], r['pos']):.3e} "
                  f"max |dvel| {max_abs(g['vel'], r['vel']):.3e}", flush=True)
            if len(settings[i].vertices) > 8192:
                assert g["placement"][0] == "team" and g["placement"][3] >= 2, g["placement"]
            assert float_equal(g["pos"], r["pos"]), f"{name}: positions differ at frame {f}, body {i}"
            assert float_equal(g["vel"], r["vel"]), f"{name}: velocities differ at frame {f}, body {i}"
            assert float_equal(g["body_pos"], r["body_pos"]), f"{name}: body position differs at frame {f}, body {i}"
            hit = r["contact_collider"] >= 0
            assert (g["contact_collider"] == r["contact_collider"]).all(), f"{name}: contact colliders differ, frame {f}"
            assert float_equal(g["contact_point"][hit], r["contact_point"][hit]), f"{name}: contact points differ, frame {f}"


def colour_order(items, nv):
    """Greedy colouring of one run of constraints of one type (in the reference's order):
    each takes the lowest colour that no earlier one holds on a shared vertex that either
    writes (an LRA reads its vertex 0).  Returns the item indices colour by colour and the
    number of colours."""
    writers = [set() for _ in range(nv)]
    readers = [set() for _ in range(nv)]
    colour_of = []
    for verts, writes in items:
        forbidden = set()
        for v, w in zip(verts, writes):
            forbidden |= writers[v]
            if w:
                forbidden |= readers[v]
        c = 0
        while c in forbidden:
            c += 1
        for v, w in zip(verts, writes):
            (writers if w else readers)[v].add(c)
        colour_of.append(c)
    n = max(colour_of) + 1 if colour_of else 0
    order = sorted(range(len(items)), key=lambda k: (colour_of[k], k))
    return order, n


def recoloured_settings(s):
    """Settings whose non-parallel group lists each type's constraints colour by colour."""
    r = s.copy()
    g = s.groups
    prev = g[-2] if len(g) > 1 else np.zeros(1, g.dtype)[0]
    nv = len(s.vertices)
    total = 0
    for arr, key, nvert, lra in ((r.volumes, "volume_end", 4, False), (r.bends, "bend_end", 4, False),
                                 (r.edges, "edge_end", 2, False), (r.lras, "lra_end", 2, True)):
        lo, hi = int(prev[key]), int(g[-1][key])
        seg = arr[lo:hi].copy()
        keep = [k for k in range(len(seg)) if not (lra and seg["v"][k][0] == seg["v"][k][1])]
        items = [(tuple(int(x) for x in seg["v"][k][:nvert]), (not lra, True) if lra else (True,) * nvert)
                 for k in keep]
        order, n = colour_order(items, nv)
        total += n
        kept = set(keep)
        skipped = [k for k in range(len(seg)) if k not in kept]
        arr[lo:hi] = seg[[keep[k] for k in order] + skipped]
    return r, total


def test_team_coloured_equals_reference_on_recoloured_settings():
    sc = scene_zoo.block_lra_boxes_streamed(frames=2)
    frames = (1, 2)
    settings, bounds, _ = reference_run(sc, (1,))
    recol, colours = zip(*(recoloured_settings(s) for s in settings))
    print(f"block_lra_boxes_streamed coloured: {colours} colours", flush=True)
    _, _, want = reference_run(sc, frames, settings=list(recol))
    gpu = _run_team(sc, settings, bounds, frames, coloured=True)
    ctx = Context(0)
    try:
        st = ctx.create_settings(settings[0]).stats()
    finally:
        ctx.close()
    assert st["tail_colourable"] and st["tail_colours"] == colours[0], (st, colours)
    for f in frames:
        g, w = gpu[f][0], want[f][0]
        print(f"  frame {f} {g['placement']}: max |dpos| {max_abs(g['pos'], w['pos']):.3e}", flush=True)
        assert g["placement"][0] == "team"
        assert float_equal(g["pos"], w["pos"]) and float_equal(g["vel"], w["vel"]), f"frame {f}"
        assert float_equal(g["body_pos"], w["body_pos"]), f"frame {f}"


def test_team_coloured_schedule_equals_reference_on_recoloured_settings():
    """Settings under XPBD_SCHEDULE_COLOURED on a team with the coloured non-parallel group
    (the fastest path for a large body): every group coloured, the last one as the team's colour
    passes; equal to the reference on the settings with every run of every group sorted by
    colour."""
    from helpers import recoloured_settings as recolour_all
    sc = scene_zoo.block_lra_boxes_streamed(frames=2)
    frames = (1, 2)
    settings, bounds, _ = reference_run(sc, (1,))
    _, _, want = reference_run(sc, frames, settings=[recolour_all(s)[0] for s in settings])
    gpu = _run_team(sc, settings, bounds, frames, coloured=True, schedule="coloured")
    for f in frames:
        g, w = gpu[f][0], want[f][0]
        print(f"  frame {f} {g['placement']}: max |dpos| {max_abs(g['pos'], w['pos']):.3e}", flush=True)
        assert g["placement"][0] == "team"
        assert float_equal(g["pos"], w["pos"]) and float_equal(g["vel"], w["vel"]), f"frame {f}"
        assert float_equal(g["body_pos"], w["body_pos"]), f"frame {f}"


def test_team_size_and_fallbacks():
    """A team per large body, one member per parallel-group window; team_min_windows 0 keeps
    the one-workgroup path; XPBD_TAIL_COLOURED is refused for LDS-resident bodies."""
    from joltphysics_amd.xpbd import XpbdError
    sc = scene_zoo.big_cloth_on_sphere(frames=1)
    s = build_settings(sc.bodies[0].mesh)
    small = build_settings(scenes.MeshSpec(*scenes.cloth(8, 0.05)[:2], attrs=scenes.attrs(1e-4, 1e-4, 1e-4)))
    ctx = Context(0)
    try:
        ctx.set_gravity(sc.gravity)
        ctx.set_colliders(sc.colliders)
        st = ctx.create_settings(s)
        a = ctx.create_body(st, sc.bodies[0].position, sc.bodies[0].rotation, sc.bodies[0].params)
        b = ctx.create_body(ctx.create_settings(small), (0, 0, 5), (0, 0, 0, 1), sc.bodies[0].params)
        with pytest.raises(XpbdError):
            ctx.set_tail_mode(b, True)
        ctx.set_team_min_windows(1)
        ctx.update(sc.dt, sc.collision_steps, 1)
        kind, _, _, size = ctx.placement(a)
        assert kind == "team" and size == st.stats()["parallel_windows"]
        assert ctx.placement(b)[0] == "bundle"
        ctx.set_team_min_windows(0)
        ctx.update(sc.dt, sc.collision_steps, 1)
        assert ctx.placement(a)[0] == "streamed"
    finally:
        ctx.close()
