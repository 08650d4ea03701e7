"""GPU parity of the coloured (throughput) schedule: the device running settings under
XPBD_SCHEDULE_COLOURED equals the reference (FMA off) running the same settings with each run
of each update group sorted by colour (tests/helpers.py: recoloured_settings), float for float,
on LDS-resident, bundled and streamed bodies; and it differs from the exact schedule's result
(the reordered sweep rounds differently).
"""
from __future__ import annotations

import numpy as np
import pytest

import scene_zoo
from helpers import float_equal, max_abs, recoloured_settings, reference_run, run_gpu

pytestmark = pytest.mark.gpu

SCENES = ("tetcube_plane_box", "pressure_sphere", "lra_cloth", "batched", "cloth100_sphere", "cloth_on_compounds",
          "block_lra_boxes_streamed")


@pytest.mark.parametrize("name", SCENES)
def test_coloured_schedule_equals_reference_on_recoloured_settings(name):
    sc = scene_zoo.ALL[name]()
    frames = tuple(sorted({1, 2, min(sc.frames, 6)}))
    from oracle import JoltReference
    ref = JoltReference("nofma")
    settings = [ref.build_settings(b.mesh) for b in sc.bodies]
    for b, s in zip(sc.bodies, settings):
        if b.post is not None:
            b.post(s)
    recoloured = [recoloured_settings(s)[0] for s in settings]
    print(f"scene {name}: {len(sc.bodies)} bodies, frames {frames}", flush=True)
    _, bounds, want = reference_run(sc, frames, settings=recoloured)
    got = run_gpu(sc, settings, bounds=bounds, frames_to_keep=frames, schedule="coloured")
    _, _, exact = reference_run(sc, frames, settings=settings)
    moved = False
    for f in frames:
        for i in range(len(sc.bodies)):
            g, r = got[f][i], want[f][i]
            print(f"  frame {f} body {i}: max |dpos| {max_abs(g['pos'], r['pos']):.3e}, "
                  f"vs the exact order {max_abs(g['pos'], exact[f][i]['pos']):.3e}", flush=True)
            assert float_equal(g["pos"], r["pos"]), f"{name}: positions differ at frame {f}, body {i}"
            assert float_equal(g["vel"], r["vel"]), f"{name}: velocities differ at frame {f}, body {i}"
            assert float_equal(g["body_pos"], r["body_pos"]), f"{name}: body position differs at frame {f}, body {i}"
            moved |= not float_equal(g["pos"], exact[f][i]["pos"])
    assert moved, "the coloured schedule gave the exact order's results: the recolouring had no effect"
