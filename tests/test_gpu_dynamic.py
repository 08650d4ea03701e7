"""GPU parity with dynamic colliders: the soft body's contact impulses shared with
dynamic rigid bodies and the velocity changes it reports
(SoftBodyMotionProperties.cpp:752-797, 923-935).

One soft body: against the reference's CustomUpdate sequence, bit-exact vertex state
and rigid-body velocities every frame (LDS-resident and streamed).  Two soft bodies on
the same dynamic bodies in one bundle: against the restatement (both read the
velocities of the call's start, as the C ABI documents), bit-exact.
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from joltphysics_amd.records import MOTION_DYNAMIC
from helpers import GpuRunner, RestatementRunner, custom_run, float_equal, reference_custom_run
from oracle import have_reference

pytestmark = pytest.mark.gpu


def _check(out, vel, ref, ref_vel, frames, ncol, nbodies=1):
    for f in frames:
        for b in range(nbodies):
            assert float_equal(out[f][b]["pos"], ref[f][b]["pos"]), f"positions differ at frame {f}, body {b}"
            assert float_equal(out[f][b]["vel"], ref[f][b]["vel"]), f"velocities differ at frame {f}, body {b}"
        for k in range(ncol):
            for j in range(2):
                assert float_equal(vel[f][k][j], ref_vel[f][k][j]), f"collider {k} velocity {j} differs at frame {f}"


@pytest.mark.skipif(not have_reference("nofma"), reason="reference library not built")
@pytest.mark.parametrize("name", list(scene_zoo.CUSTOM))
def test_dynamic_bodies_bit_exact_with_reference(name):
    sc = scene_zoo.CUSTOM[name]()
    frames = tuple(range(1, sc.frames + 1))
    print(f"scene {name}: frames {frames[-1]}", flush=True)
    settings, bounds, ref, ref_vel = reference_custom_run(sc, frames)
    run = GpuRunner(sc, settings, bounds)
    out, vel = custom_run(sc, settings, bounds, frames, run)
    run.close()
    _check(out, vel, ref, ref_vel, frames, len(sc.colliders))
    assert ref_vel[frames[-1]][0] != ref_vel[1][0], "the cloth never reached the dynamic box"


@pytest.mark.skipif(not have_reference("nofma"), reason="reference library not built")
def test_two_bodies_on_dynamic_bodies_match_restatement():
    import copy
    sc = scene_zoo.two_cloths_on_dynamic_bodies()
    frames = tuple(range(1, sc.frames + 1))
    # settings and colliders' device data from the reference (its settings builder)
    from oracle import JoltReference
    ref = JoltReference("nofma")
    settings = [ref.build_settings(b.mesh) for b in sc.bodies]
    single = scene_zoo.cloth_on_dynamic_bodies()
    _, bounds, _, _ = reference_custom_run(single, (1,))
    sc.colliders = single.colliders      # device parameters and initial motion from the reference
    sc_cpu = copy.deepcopy(sc)
    run = RestatementRunner(sc_cpu, settings, bounds)
    want, want_vel = custom_run(sc_cpu, settings, bounds, frames, run)
    run.close()
    run = GpuRunner(sc, settings, bounds)
    out, vel = custom_run(sc, settings, bounds, frames, run)
    run.close()
    _check(out, vel, want, want_vel, frames, len(sc.colliders), nbodies=2)


@pytest.mark.skipif(not have_reference("nofma"), reason="reference library not built")
@pytest.mark.parametrize("name", list(scene_zoo.CUSTOM))
def test_contact_validation_settings_bit_exact(name):
    """Per-pair contact settings as a contact listener's OnSoftBodyContactValidate gives them
    (SoftBodyMotionProperties.cpp:151-216, 763, 798): a dynamic box with all three scales
    changed (soft body mInvMassScale1 0.5, mInvMassScale2 3, mInvInertiaScale2 2), a rejected
    dynamic sphere, and a static box with mInvMassScale1 0 (no interaction: skipped);
    bit-exact against the reference's listener every checked frame, the colliders' linear and
    angular velocities included, LDS-resident and streamed.  A run with unit settings shows
    that the cloth reaches the rejected sphere and the skipped box, so both branches act."""
    sc = scene_zoo.CUSTOM[name]()
    nc = len(sc.colliders)
    dyn = [k for k, c in enumerate(sc.colliders) if c.motion_type == MOTION_DYNAMIC]
    static = [k for k, c in enumerate(sc.colliders) if c.motion_type != MOTION_DYNAMIC]
    assert len(dyn) >= 2 and len(static) >= 1
    frames = tuple(range(1, sc.frames + 1))
    print(f"scene {name}: contact settings, frames {frames[-1]}", flush=True)

    unit = np.tile(np.float32([1.0, 1.0, 1.0, 1.0]), (nc, 1))
    settings, bounds, _, _ = reference_custom_run(sc, (1,))
    run = GpuRunner(sc, settings, bounds, contact_settings=unit)
    out, _ = custom_run(sc, settings, bounds, frames, run)
    run.close()
    touched = set()
    for f in frames:
        touched |= set(int(k) for k in np.unique(out[f][0]["contact_collider"]) if k >= 0)
    print(f"  colliders touched with unit settings: {sorted(touched)}", flush=True)
    assert dyn[0] in touched and dyn[1] in touched and static[0] in touched, touched

    sc = scene_zoo.CUSTOM[name]()
    table = unit.copy()
    table[dyn[0]] = (0.5, 3.0, 2.0, 1.0)
    table[dyn[1]] = (1.0, 1.0, 1.0, 0.0)
    table[static[0]] = (0.0, 1.0, 1.0, 1.0)
    settings, bounds, ref, ref_vel = reference_custom_run(sc, frames, contact_settings=table)
    run = GpuRunner(sc, settings, bounds, contact_settings=table)
    out, vel = custom_run(sc, settings, bounds, frames, run)
    run.close()
    _check(out, vel, ref, ref_vel, frames, nc)
    for f in frames:
        hit = out[f][0]["contact_collider"]
        assert not np.isin(hit, [dyn[1], static[0]]).any(), f"a rejected or skipped collider reports contact, frame {f}"
    assert any(ref_vel[f][dyn[0]] != ref_vel[1][dyn[0]] for f in frames), "the cloth never reached the scaled box"
