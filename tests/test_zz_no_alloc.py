"""The allocation-free rollout forward (xpbd_rollout_forward_noalloc) against the
ordinary one (xpbd_rollout_forward), through the C ABI.

Both run the same float forward pass; the allocation-free one keeps the pass's
per-vertex velocity and previous position in the library's fixed device array.
Checked bit for bit (integer-exact comparison of float outputs) on two calls in a
row, the second on a smaller rollout, so the second call runs on an array the first
one left dirty; then a backward pass after an allocation-free forward equals the
backward pass after an ordinary forward, and a rollout above the array's capacity is
refused with XPBD_E_CAPACITY before any launch.
"""
from __future__ import annotations

import numpy as np
import pytest

from joltphysics_amd import scenes
from joltphysics_amd.xpbd import Context, XpbdError, build_settings

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

XPBD_NOALLOC_MAX_VERTICES = 262144


def _rollout(ctx, num_bodies, n, frames, interval):
    case = scenes.config4_rollout(num_bodies=num_bodies, n=n, frames=frames, compliance=1e-3)
    st = ctx.create_settings(build_settings(case.mesh))
    r = ctx.create_rollout([st] * case.num_bodies, num_iterations=case.iterations, dt=case.dt, frames=case.frames,
                           checkpoint_interval=interval)
    return case, r


def _inputs(case, r):
    rng = np.random.default_rng(11)
    comp = torch.from_numpy(case.compliance * rng.uniform(0.5, 2.0, case.compliance.shape)).float().cuda()
    vel = torch.from_numpy(case.velocity.reshape(-1)).float().cuda()
    return comp, vel


def _run(r, comp, vel, noalloc):
    out = torch.empty((r.num_vertices, 3), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    (r.forward_noalloc if noalloc else r.forward)(comp.data_ptr(), vel.data_ptr(), out.data_ptr(), stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _grads(r, comp, vel, gx):
    gc = torch.empty_like(comp)
    gv = torch.empty_like(vel)
    r.backward(gx.data_ptr(), gc.data_ptr(), gv.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return gc.cpu().numpy(), gv.cpu().numpy()


def test_noalloc_forward_matches_forward_on_two_calls():
    ctx = Context(0)
    try:
        ctx.set_gravity((0.0, 0.0, -9.81))
        # first call: 6 cloths 32 x 32 (6,144 vertices); second call: 2 cloths 16 x 16 (512)
        for num_bodies, n, frames, interval in ((6, 32, 12, 5), (2, 16, 9, 4)):
            print(f"noalloc forward: {num_bodies} cloths {n}x{n}, {frames} frames", flush=True)
            case, r = _rollout(ctx, num_bodies, n, frames, interval)
            comp, vel = _inputs(case, r)
            want = _run(r, comp, vel, noalloc=False)
            got = _run(r, comp, vel, noalloc=True)
            assert np.isfinite(want).all()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
                f"max |d| {np.abs(got - want).max():.3e}"
            # the adjoint after the allocation-free forward equals the adjoint after the ordinary one
            gx = torch.from_numpy(np.random.default_rng(3).standard_normal((r.num_vertices, 3))).float().cuda()
            gc1, gv1 = _grads(r, comp, vel, gx)
            _run(r, comp, vel, noalloc=False)
            gc0, gv0 = _grads(r, comp, vel, gx)
            assert np.array_equal(gc1.view(np.uint32), gc0.view(np.uint32))
            assert np.array_equal(gv1.view(np.uint32), gv0.view(np.uint32))
            r.close()
    finally:
        torch.cuda.synchronize()
        ctx.close()


def test_noalloc_forward_rejects_rollouts_beyond_its_array():
    ctx = Context(0)
    try:
        ctx.set_gravity((0.0, 0.0, -9.81))
        # 65 cloths 64 x 64 = 266,240 vertices > 262,144
        case, r = _rollout(ctx, 65, 64, 1, 1)
        assert r.num_vertices > XPBD_NOALLOC_MAX_VERTICES
        comp, vel = _inputs(case, r)
        out = torch.empty((r.num_vertices, 3), dtype=torch.float32, device="cuda")
        with pytest.raises(XpbdError, match="allocation-free"):
            r.forward_noalloc(comp.data_ptr(), vel.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        r.close()
    finally:
        torch.cuda.synchronize()
        ctx.close()
