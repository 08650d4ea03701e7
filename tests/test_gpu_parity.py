"""GPU parity: the HIP path through the C ABI against the reference.

Bar: bit-exact (float equality) vertex positions, velocities and body
positions against the reference compiled with FMA contraction off, for every
frame checked.  The reference's trajectory comes from the committed golden
fixture (tests/golden/make_fixtures.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import float_equal, load_config1_fixture, max_abs, run_gpu
from joltphysics_amd import scenes

pytestmark = pytest.mark.gpu


def test_config1_cloth_on_box_bit_exact_300_frames():
    fx, s = load_config1_fixture()
    sc = scenes.config1_cloth_on_box()
    frames = tuple(int(f) for f in fx["frames"])
    print(f"config1 cloth32 on box: frames {frames}", flush=True)
    out = run_gpu(sc, [s], bounds=fx["collider_bounds"], frames_to_keep=frames)
    for f in frames:
        g = out[f][0]
        dp = max_abs(g["pos"], fx[f"pos_{f}"])
        dv = max_abs(g["vel"], fx[f"vel_{f}"])
        print(f"frame {f}: max |dpos| {dp:.3e} max |dvel| {dv:.3e}", flush=True)
        assert float_equal(g["pos"], fx[f"pos_{f}"]), f"positions differ at frame {f}: {dp}"
        assert float_equal(g["vel"], fx[f"vel_{f}"]), f"velocities differ at frame {f}: {dv}"
        assert float_equal(g["body_pos"], fx[f"body_pos_{f}"]), f"body position differs at frame {f}"
