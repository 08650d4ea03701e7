"""The level schedule equals the reference's sequential order (CPU, no device).

The reference applies its update groups one after the other and, inside a group, its
constraints one by one in ProcessGroup's order: volumes, bends, skinned, edges, rods,
LRAs (SoftBodyMotionProperties.cpp:1030-1055).  The device runs levels: the records of
a level concurrently, the levels in order.  The programs the device would get
(xpbd_schedule_build: the windowed program of streamed bodies, the identity program
bundles merge) are read back and checked against the reference's order:

* every constraint appears exactly once, with its own vertices and parameters;
* for every vertex, any two constraints that touch it, at least one of them writing it,
  sit in levels ordered as the constraints are in the sequence (a write after a write or
  a read, a read after a write: strictly later level).

Together these make every constraint read exactly the values it reads in the sequential
loop and write the same values, so the level program's results are those of the
sequential loop bit for bit, whatever order the records of a level run in.  The coloured
non-parallel group is checked for what its mode promises: the constraints of a colour
touch disjoint written vertices, and every constraint of the group appears once.
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from joltphysics_amd import scenes
from joltphysics_amd.xpbd import build_settings, host_schedule

TYPE_VOLUME, TYPE_BEND, TYPE_EDGE, TYPE_LRA = 0, 1, 2, 3
NOP = 0xFFFFFFFF
SECOND_HALF = 1 << 15


def sequence(s):
    """The reference's processing order: (type, index) per position, and per type the
    position of each constraint (-1: never projected)."""
    g = s.groups
    prev = {k: 0 for k in g.dtype.names}
    types, index = [], []
    for row in g:
        for t, key, arr in ((TYPE_VOLUME, "volume_end", s.volumes), (TYPE_BEND, "bend_end", s.bends),
                            (TYPE_EDGE, "edge_end", s.edges), (TYPE_LRA, "lra_end", s.lras)):
            k = np.arange(prev[key], int(row[key]))
            if t == TYPE_LRA:
                k = k[arr["v"][k, 0] != arr["v"][k, 1]]     # an LRA of a vertex to itself never moves it
            types.append(np.full(len(k), t))
            index.append(k)
        prev = {k: int(row[k]) for k in g.dtype.names}
    types, index = np.concatenate(types), np.concatenate(index)
    pos = {t: np.full(n, -1, np.int64) for t, n in ((TYPE_VOLUME, len(s.volumes)), (TYPE_BEND, len(s.bends)),
                                                      (TYPE_EDGE, len(s.edges)), (TYPE_LRA, len(s.lras)))}
    for t in pos:
        sel = types == t
        pos[t][index[sel]] = np.nonzero(sel)[0]
    return len(types), pos


def decode(recs, local_to_body):
    """Type and body vertices of packed records (window-local vertex numbers mapped)."""
    x, y = recs[:, 0], recs[:, 1]
    t = (x >> 14) & 3
    v = np.stack([x & 0x3FFF, x >> 16, y & 0x7FFF, y >> 16], axis=1).astype(np.int64)
    return t, local_to_body(v)


def levels_of(recs, rec_index, lvl_ranges, local_maps):
    """(type, constraint index, level, body vertices, p0, p1) of every real record; the
    second half of a two-lane bend is dropped (it carries the same constraint)."""
    out = []
    level = 0
    for (begin, offsets), local in zip(lvl_ranges, local_maps):
        for L in range(len(offsets) - 1):
            lo, hi = int(offsets[L]), int(offsets[L + 1])
            r = recs[lo:hi]
            keep = r[:, 0] != NOP
            r = r[keep]
            t, v = decode(r, local)
            second = (t == TYPE_BEND) & ((r[:, 1] & SECOND_HALF) != 0)
            keep2 = ~second
            out.append((t[keep2], rec_index[lo:hi][keep][keep2].astype(np.int64), np.full(keep2.sum(), level),
                        v[keep2], r[keep2, 2].view(np.float32), r[keep2, 3].view(np.float32)))
            level += 1
    return [np.concatenate([o[k] for o in out]) for k in range(6)]


def check_program(s, t, idx, level, verts, p0, p1):
    n_seq, pos = sequence(s)
    arrays = {TYPE_VOLUME: (s.volumes, "six_rest_volume", "compliance", 4),
              TYPE_BEND: (s.bends, "compliance", "initial_angle", 4),
              TYPE_EDGE: (s.edges, "rest", "compliance", 2), TYPE_LRA: (s.lras, "max_distance", None, 2)}
    seqpos = np.zeros(len(t), np.int64)
    for ty, (arr, a, b, nvert) in arrays.items():
        sel = np.nonzero(t == ty)[0]
        i = idx[sel]
        assert (i < len(arr)).all()
        seqpos[sel] = pos[ty][i]
        assert (pos[ty][i] >= 0).all(), f"type {ty}: a record for a constraint that is never projected"
        assert (verts[sel, :nvert] == arr["v"][i][:, :nvert].astype(np.int64)).all(), f"type {ty}: vertices differ"
        assert (p0[sel] == arr[a][i]).all(), f"type {ty}: {a} differs"
        if b is not None:
            assert (p1[sel] == arr[b][i]).all(), f"type {ty}: {b} differs"
    # every constraint exactly once
    assert np.array_equal(np.sort(seqpos), np.arange(n_seq)), "constraints missing or repeated"

    # vertex accesses: (vertex, sequence position, level, write)
    nvert = np.where(t == TYPE_EDGE, 2, np.where(t == TYPE_LRA, 2, 4))
    acc_v, acc_s, acc_l, acc_w = [], [], [], []
    for j in range(4):
        sel = nvert > j
        acc_v.append(verts[sel, j])
        acc_s.append(seqpos[sel])
        acc_l.append(level[sel])
        acc_w.append(~((t[sel] == TYPE_LRA) & (j == 0)))
    v, sq, lv, w = (np.concatenate(a) for a in (acc_v, acc_s, acc_l, acc_w))
    o = np.lexsort((sq, v))
    v, sq, lv, w = v[o], sq[o], lv[o], w[o]
    n = len(v)
    ar = np.arange(n)
    # last write at or before each access (same vertex)
    lw = np.maximum.accumulate(np.where(w, ar, -1))
    prev_w = np.concatenate([[-1], lw[:-1]])
    has = (prev_w >= 0) & (v[np.maximum(prev_w, 0)] == v)
    assert (lv[has] > lv[prev_w[has]]).all(), "an access does not come after the previous write of its vertex"
    # next write after each read (same vertex)
    nw = np.minimum.accumulate(np.where(w, ar, n)[::-1])[::-1]
    next_w = np.concatenate([nw[1:], [n]])
    r = ~w & (next_w < n)
    r[r] = v[next_w[r]] == v[r]
    assert (lv[next_w[r]] > lv[r]).all(), "a write does not come after an earlier read of its vertex"
    return n_seq, int(level.max()) + 1 if len(level) else 0


def windowed_program(h):
    ranges, maps = [], []
    for w in h["windows"]:
        vbegin, vcount, lbegin, nlev = (int(x) for x in w)
        ranges.append((lbegin, h["lvl"][lbegin:lbegin + nlev + 1]))
        if vcount:
            vl = h["vlist"][vbegin:vbegin + vcount].astype(np.int64)
            maps.append(lambda v, vl=vl: np.where(v < len(vl), vl[np.minimum(v, len(vl) - 1)], -1))
        else:
            maps.append(lambda v: v)
    return levels_of(h["recs"], h["rec_index"], ranges, maps)


MESHES = {
    "cloth32_resident": lambda: scenes.config1_cloth_on_box().bodies[0].mesh,
    "cloth100_streamed": lambda: scene_zoo.big_cloth_on_sphere().bodies[0].mesh,
    "tet_cube": lambda: scene_zoo.tet_cube_on_plane().bodies[0].mesh,
    "lra_cloth": lambda: scene_zoo.lra_cloth().bodies[0].mesh,
    "block21_lra_streamed": lambda: scene_zoo.block_lra_boxes_streamed().bodies[0].mesh,
    "icosphere4": lambda: scenes.mesh_for_key(("ico", 4)),
    "cloth127_streamed": lambda: scenes.mesh_for_key(("cloth", 127)),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_level_programs_follow_the_reference_order(name):
    s = build_settings(MESHES[name]())
    h = host_schedule(s)
    n, nlev = check_program(s, *windowed_program(h))
    print(f"{name}: {n} constraints in {nlev} levels over {len(h['windows'])} windows", flush=True)
    if len(h["identity_level_begin"]):
        ident = levels_of(h["identity_recs"], h["identity_rec_index"], [(0, h["identity_level_begin"])], [lambda v: v])
        check_program(s, *ident)


@pytest.mark.parametrize("name", ["block21_lra_streamed", "cloth127_streamed"])
def test_coloured_group_is_independent_within_colours(name):
    s = build_settings(MESHES[name]())
    h = host_schedule(s)
    cr, cb = h["colour_recs"], h["colour_begin"]
    assert len(cb) >= 2
    v = cr[:, :4].astype(np.int64)
    t = cr[:, 4]
    nvert = np.where((t == TYPE_EDGE) | (t == TYPE_LRA), 2, 4)
    for c in range(len(cb) - 1):
        lo, hi = int(cb[c]), int(cb[c + 1])
        written, read = [], []
        for j in range(4):
            sel = np.arange(lo, hi)[nvert[lo:hi] > j]
            lra0 = (t[sel] == TYPE_LRA) & (j == 0)
            written.append(v[sel[~lra0], j])
            read.append(v[sel[lra0], j])
        written, read = np.concatenate(written), np.concatenate(read)
        assert len(np.unique(written)) == len(written), f"colour {c}: a vertex written twice"
        assert not np.isin(read, written).any(), f"colour {c}: a vertex read and written"
        assert len(np.unique(t[lo:hi])) == 1, f"colour {c}: mixed constraint types"
    # the coloured records are the non-parallel group's constraints, each once
    g = s.groups
    prev = g[-2] if len(g) > 1 else np.zeros(1, g.dtype)[0]
    want = sum(int(g[-1][k]) - int(prev[k]) for k in ("volume_end", "bend_end", "edge_end"))
    lo, hi = int(prev["lra_end"]), int(g[-1]["lra_end"])
    want += int((s.lras["v"][lo:hi, 0] != s.lras["v"][lo:hi, 1]).sum())
    assert len(cr) == want


def test_checker_catches_a_reordered_level():
    """The check above fails on a program whose levels 10 and 11 are swapped, and on one with
    a record dropped: it is not vacuous."""
    s = build_settings(MESHES["cloth32_resident"]())
    t, idx, level, verts, p0, p1 = windowed_program(host_schedule(s))
    swapped = level.copy()
    swapped[level == 10], swapped[level == 11] = 11, 10
    with pytest.raises(AssertionError, match="after"):
        check_program(s, t, idx, swapped, verts, p0, p1)
    keep = np.arange(len(t)) != 5
    with pytest.raises(AssertionError, match="missing"):
        check_program(s, t[keep], idx[keep], level[keep], verts[keep], p0[keep], p1[keep])


@pytest.mark.parametrize("name", list(MESHES))
def test_coloured_schedule_is_the_reference_order_of_recoloured_settings(name):
    """XPBD_SCHEDULE_COLOURED: the program follows, as checked above, the sequential order of
    the same settings with each run of each group sorted by colour (tests/helpers.py), so the
    device's results under it are the reference's on those settings bit for bit.  It takes far
    fewer levels than the exact program."""
    from helpers import recoloured_settings
    s = build_settings(MESHES[name]())
    r, perm = recoloured_settings(s)
    inverse = {}
    for ty, key in ((TYPE_VOLUME, "volumes"), (TYPE_BEND, "bends"), (TYPE_EDGE, "edges"), (TYPE_LRA, "lras")):
        inv = np.zeros(len(perm[key]), np.int64)
        inv[perm[key]] = np.arange(len(perm[key]))
        inverse[ty] = inv
    h = host_schedule(s, "coloured")

    def remap(t, idx, level, verts, p0, p1):
        idx = idx.copy()
        for ty, inv in inverse.items():
            sel = t == ty
            idx[sel] = inv[idx[sel]]
        return t, idx, level, verts, p0, p1

    n, nlev = check_program(r, *remap(*windowed_program(h)))
    _, nlev_exact = check_program(s, *windowed_program(host_schedule(s)))
    print(f"{name}: {n} constraints in {nlev} coloured levels ({nlev_exact} exact)", flush=True)
    assert nlev * 3 < nlev_exact
    if len(h["identity_level_begin"]):
        ident = levels_of(h["identity_recs"], h["identity_rec_index"], [(0, h["identity_level_begin"])], [lambda v: v])
        check_program(r, *remap(*ident))


def test_coloured_tail_schedule_colours_only_the_last_group():
    """XPBD_SCHEDULE_COLOURED_TAIL keeps the parallel groups' levels and colours the last group."""
    s = build_settings(MESHES["cloth127_streamed"]())
    exact, tail = host_schedule(s), host_schedule(s, "coloured_tail")
    npw = exact["num_par_windows"]
    assert np.array_equal(exact["windows"][:npw, 3], tail["windows"][:npw, 3])
    assert tail["windows"][npw:, 3].sum() * 10 < exact["windows"][npw:, 3].sum()
