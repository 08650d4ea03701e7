"""The FP64 CPU oracle of the differentiable rollout (oracle/xpbd_adjoint.cpp).

Its adjoint is pinned against central finite differences of its own FP64
forward pass: the derivative is fixed by the mathematics, so a wrong Jacobian
term (a dropped rest-length term, a sign, the compliance factor) fails here.
Its FP32 forward pass is pinned bit for bit against the reference-parity
restatement of the update (the FP32 update of an edge-only body without
colliders).
"""
from __future__ import annotations

import numpy as np
import pytest

from joltphysics_amd import scenes
from joltphysics_amd.records import BodyParams
from joltphysics_amd.xpbd import build_settings
from oracle import RolloutReference, have_restatement

pytestmark = pytest.mark.skipif(not have_restatement(), reason="oracle restatement not built (make -C oracle restatement)")


KINDS = ["edges", "edges_lra", "tet_volumes", "edges_bends", "balloon_pressure"]


def _case(n=6, frames=8, kind="edges"):
    if kind == "tet_volumes":
        case = scenes.rollout_tet_block(num_bodies=1, cells=2, frames=frames)
    elif kind == "balloon_pressure":
        case = scenes.rollout_balloon(num_bodies=1, subdivisions=1, frames=frames)
    else:
        case = scenes.config4_rollout(num_bodies=1, n=n, frames=frames, lra=kind == "edges_lra", bends=kind == "edges_bends")
    s = build_settings(case.mesh)
    ref = RolloutReference(s, iterations=case.iterations, frames=frames, dt=case.dt, gravity=case.gravity,
                           pressure=case.pressure)
    comp = case.compliance.astype(np.float64)
    vel = case.velocity.astype(np.float64)
    target = ref.forward(comp * case.target_scale, vel)
    return case, s, ref, comp, vel, target


def _loss(ref, target, c, v):
    x = ref.forward(c, v)
    return float(np.mean(np.sum((x - target) ** 2, axis=1)))


@pytest.mark.parametrize("kind", KINDS)
def test_adjoint_matches_central_differences_fp64(kind):
    """The FP64 adjoint against central differences of the FP64 forward pass: edge-only, with
    LRA constraints (some of which act: checked), a hanging tet block with volumes, a cloth
    with dihedral bends, and a pinned balloon under pressure (whose volume grows: checked)."""
    case, s, ref, comp, vel, target = _case(kind=kind)
    if kind == "tet_volumes":
        assert len(s.volumes) > 0
    if kind == "edges_bends":
        assert len(s.bends) > 0
    if kind == "balloon_pressure":
        x = ref.forward(comp, vel)
        f = s.faces["v"].astype(np.int64)
        six = lambda p: float(np.einsum("ij,ij->i", np.cross(p[f[:, 0]], p[f[:, 1]]), p[f[:, 2]]).sum())
        p0 = s.vertices["pos"].astype(np.float64)
        assert six(x - x[f].reshape(-1, 3).mean(0)) > six(p0 - p0.mean(0)) * 1.0001, "pressure does not inflate"
    if kind == "edges_lra":
        assert len(s.lras) > 0
        stretched = ref.forward(comp, vel)
        lv = s.lras["v"].astype(np.int64)
        d = np.linalg.norm(stretched[lv[:, 1]] - stretched[lv[:, 0]], axis=1)
        assert (d >= s.lras["max_distance"] * (1 - 1e-6)).any(), "no LRA acts at the end"
        assert np.array_equal(ref.forward(comp, vel), RolloutReference(s, iterations=case.iterations, frames=case.frames, dt=case.dt,
                                                                         gravity=case.gravity).forward(comp, vel))
    x = ref.forward(comp, vel)
    g = 2.0 * (x - target) / ref.nv
    x2, gc, gv = ref.gradient(comp, vel, g)
    assert np.array_equal(x, x2)
    scale_c = np.abs(gc).max()
    for e in range(0, ref.ne, 5):
        h = 1e-4 * comp[e]
        cp, cm = comp.copy(), comp.copy()
        cp[e] += h
        cm[e] -= h
        fd = (_loss(ref, target, cp, vel) - _loss(ref, target, cm, vel)) / (2 * h)
        assert abs(fd - gc[e]) <= 1e-5 * max(abs(fd), 1e-3 * scale_c), (e, fd, gc[e])
    flat = vel.reshape(-1)
    gvf = gv.reshape(-1)
    scale_v = np.abs(gvf).max()
    for k in range(0, flat.size, 7):
        h = 1e-5
        vp, vm = flat.copy(), flat.copy()
        vp[k] += h
        vm[k] -= h
        fd = (_loss(ref, target, comp, vp.reshape(-1, 3)) - _loss(ref, target, comp, vm.reshape(-1, 3))) / (2 * h)
        assert abs(fd - gvf[k]) <= 1e-5 * max(abs(fd), 1e-3 * scale_v), (k, fd, gvf[k])


@pytest.mark.parametrize("kind", KINDS)
def test_fp32_forward_matches_update_restatement(kind):
    """The rollout oracle's FP32 pass equals the parity restatement of the full update
    on the same body (no colliders, identity rotation, no recentring), with LRAs and volumes too."""
    from oracle import Restatement
    case, s, ref, comp, vel, target = _case(n=8, frames=6, kind=kind)
    x32 = ref.forward(comp, vel, precision=32)
    st = s.copy()
    st.edges["compliance"] = comp.astype(np.float32)
    st.vertices["vel"] = vel.astype(np.float32)
    r = Restatement(case.gravity, 32)
    r.add_soft_body(st, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0),
                    BodyParams(num_iterations=case.iterations, allow_sleeping=False, update_position=False,
                               pressure=case.pressure))
    r.update(case.dt, 1, 6)
    out = r.read(0)
    r.close()
    assert np.array_equal(out["pos"].astype(np.float64), x32)


def test_fp32_forward_close_to_fp64():
    case, s, ref, comp, vel, target = _case(n=8, frames=20)
    d = np.abs(ref.forward(comp, vel, precision=32) - ref.forward(comp, vel, precision=64)).max()
    assert d < 1e-4, d
