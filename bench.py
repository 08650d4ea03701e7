#!/usr/bin/env python3
"""Benchmark of the MI355X soft-body XPBD update.

One "step" is one frame: PhysicsSystem::Update(dt = 1/60, collision_steps) for
every soft body of the workload, i.e. collision_steps x iterations XPBD
sub-steps per body.  The default workload is the batched uneven scene (2048
bodies: cloths 8^2-128^2, 5-tet cubes 4^3-16^3 cells, pressurised icospheres;
2 collision steps x 5 iterations) with synthetic geometry.  The K timed frames
run as one launch per kernel variant on the library's stream, with the inputs
resident in HBM.

Prints one JSON line:
  metric/value/unit   constraint solves per second (sum over bodies of
                      steps x iterations x constraints, per frame)
  roofline            algorithmic HBM bytes of the update kernels per launch /
                      launch duration (hipEvents on the library stream)
  cpu_baseline        the reference (compiled from its sources, shipped x86-64
                      flags, PhysicsSystem::Update with JobSystemThreadPool on
                      all host cores) on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from joltphysics_amd import scenes  # noqa: E402
from joltphysics_amd.xpbd import Context, build_settings  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit (fp32 SoA; each constraint gathers position +
# inverse mass of its vertices and scatters positions, its record is read once
# per sub-step; per vertex and sub-step the velocity update + integration read
# x, prev, v, w, plane, shape index and write prev, x, v; per vertex and
# collision step the plane search reads 16 B and writes 24 B).
BYTES_EDGE, BYTES_BEND, BYTES_VOLUME, BYTES_LRA = 72, 136, 136, 48
BYTES_VERTEX_SUBSTEP, BYTES_VERTEX_STEP = 96, 40


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    if name == "batched2048":
        return scenes.config3_batched()
    if name == "cloth512_cube48":
        return scenes.config2_cloth_and_cube()
    if name == "cloth32_on_box":
        return scenes.config1_cloth_on_box()
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes_per_frame(scene, settings):
    total = 0
    for b, s in zip(scene.bodies, settings):
        sub = scene.collision_steps * b.params.num_iterations
        total += sub * (BYTES_EDGE * len(s.edges) + BYTES_BEND * len(s.bends) + BYTES_VOLUME * len(s.volumes)
                        + BYTES_LRA * len(s.lras) + BYTES_VERTEX_SUBSTEP * len(s.vertices))
        total += scene.collision_steps * BYTES_VERTEX_STEP * len(s.vertices)
    return total


def usable_cpus():
    """CPUs this process may run on: affinity mask, capped by a cgroup CPU quota."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()
            if quota != "max":
                n = min(n, max(1, int(int(quota) / int(period))))
    except Exception:
        pass
    return n


def cpu_baseline(scene, settings, target_seconds=15.0, max_bodies=256):
    """Reference CPU path on a bounded sample of the workload (first bodies of the
    same randomised list), all host cores."""
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import JoltReference, JoltWorld, have_reference
        if not have_reference("fma"):
            log("cpu_baseline: oracle/_ref/libjolt_oracle_fma.so not built; skipped")
            return None
        ref = JoltReference("fma")
        threads = max(min(ref.hardware_concurrency(), usable_cpus()) - 1, 0)
        nb = min(max_bodies, len(scene.bodies))
        world = JoltWorld(ref, scene.gravity, max_bodies=nb + len(scene.colliders) + 16, num_threads=threads)
        for c in scene.colliders:
            world.add_collider(c)
        for b, s in zip(scene.bodies[:nb], settings[:nb]):
            world.add_soft_body(s, b.position, b.rotation, b.params)
        world.optimize_broadphase()
        sub_scene_solves = scenes.solves_per_frame(scene, settings[:nb]) if nb == len(scene.bodies) else \
            sum(scene.collision_steps * b.params.num_iterations * s.num_constraints()
                for b, s in zip(scene.bodies[:nb], settings[:nb]))
        world.update(scene.dt, scene.collision_steps, 1)          # warm-up frame
        t1 = world.update(scene.dt, scene.collision_steps, 1)
        frames = int(max(1, min(600, target_seconds / max(t1, 1e-6))))
        t = world.update(scene.dt, scene.collision_steps, frames)
        world.close()
        return {"value": sub_scene_solves * frames / t, "unit": "constraint-solves/s", "cores": threads + 1,
                "kind": "reference",
                "sample": f"first {nb} of {len(scene.bodies)} bodies, {frames} frames, "
                          f"PhysicsSystem::Update on JobSystemThreadPool({threads} workers + caller), "
                          f"reference built with -O3 -mavx2 -mfma -ffp-contract=fast",
                "ms_per_frame": 1000.0 * t / frames}
    except Exception as e:  # the baseline is reported, never required
        log(f"cpu_baseline failed: {e!r}")
        return None


def measured_traffic(workload):
    """HBM bytes per frame of the update measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (calibrated against a plain copy), as recorded in profiles/<workload>_hbm_traffic.md."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_frame:"):
            return float(line.split(":")[1]), os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))
    return None, None


DIGESTS = {"exact": "batched2048_sample_digests.json", "coloured": "batched2048_coloured_sample_digests.json"}


def parity_check(ctx, ids, scene, settings, frame, schedule="exact"):
    """Compares the GPU state of the batched scene's 64 check bodies after `frame` frames
    from the initial state with digests of the reference's state (tests/golden/
    batched2048_sample_digests.json, written from the reference built with FMA
    contraction off; for the coloured schedule batched2048_coloured_sample_digests.json, the
    reference on the same settings with each run of each update group sorted by colour)."""
    name = DIGESTS[schedule]
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return {"checked": False, "reason": "no reference digests for this workload"}
    with open(path) as f:
        fx = json.load(f)
    if fx["scene"] != scene.name or fx["bodies"] != len(scene.bodies):
        return {"checked": False, "reason": "no reference digests for this workload"}
    if str(frame) not in fx["frames"]:
        return {"checked": False, "reason": f"the reference digests cover frames {min(map(int, fx['frames']))}-"
                                            f"{max(map(int, fx['frames']))}, the warm-up ends at frame {frame}"}
    sample = fx["sample"]
    if sample != scenes.batched_check_sample(scene, settings):
        return {"checked": False, "reason": "the workload's check sample differs from the digests'"}
    want = fx["frames"][str(frame)]
    got = [scenes.state_digest(ctx.read(ids[k])) for k in sample]
    mismatched = [k for k, g, w in zip(sample, got, want) if g != w]
    return {"checked": True, "frame": frame, "bodies": len(sample), "bit_exact": not mismatched,
            "mismatched_bodies": mismatched,
            "against": "reference PhysicsSystem::Update (FMA off)" + (" on recoloured settings" if schedule != "exact" else "")
                       + ": tests/golden/" + name}


def timed_run(scene, settings, schedule, warmup, steps):
    """Uploads the workload with the given schedule, runs `warmup` frames, checks the sampled
    bodies against the reference's digests, then times `steps` frames in one xpbd_update call.
    Returns (wall seconds, launches, kernel milliseconds, parity)."""
    ctx = Context(0)
    ctx.set_gravity(scene.gravity)
    ctx.set_colliders(scene.colliders)
    t0 = time.time()
    ids = []
    for b, s in zip(scene.bodies, settings):
        st = ctx.create_settings(s, schedule)
        ids.append(ctx.create_body(st, b.position, b.rotation, b.params))
    log(f"uploaded ({schedule} schedule) in {time.time() - t0:.1f}s")
    if warmup > 0:
        ctx.update(scene.dt, scene.collision_steps, warmup)
    ctx.synchronize()
    parity = parity_check(ctx, ids, scene, settings, warmup, schedule)
    log(f"parity after the warm-up: {parity}")
    ctx.profile_enable(True)
    t0 = time.perf_counter()
    ctx.update(scene.dt, scene.collision_steps, steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    n_launch, kernel_ms = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.close()
    log(f"{schedule} schedule: {1000.0 * (t1 - t0) / steps:.3f} ms per frame")
    return t1 - t0, n_launch, kernel_ms, parity


def differentiation_run(warmup=1, repeats=3):
    """fwd+bwd of the differentiation workload (64 cloths 64 x 64, 60 frames x 5 iterations,
    checkpoints every 10 frames): the float forward pass and the adjoint through the
    torch.autograd front end, timed with device synchronisation around each pass."""
    import torch
    from joltphysics_amd.autograd import rollout
    case = scenes.config4_rollout()
    s = build_settings(case.mesh)
    ctx = Context(0)
    ctx.set_gravity(case.gravity)
    st = ctx.create_settings(s)
    r = ctx.create_rollout([st] * case.num_bodies, num_iterations=case.iterations, dt=case.dt, frames=case.frames,
                           checkpoint_interval=case.checkpoint_interval)
    dev = torch.device("cuda")
    target_c = torch.from_numpy(case.compliance * case.target_scale).to(dev)
    vel0 = torch.from_numpy(case.velocity.reshape(-1)).to(dev)
    with torch.no_grad():
        target = rollout(target_c, vel0, r).clone()
    fwd, tot = [], []
    for k in range(warmup + repeats):
        comp = torch.from_numpy(case.compliance).to(dev).requires_grad_(True)
        vel = vel0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = rollout(comp, vel, r)
        loss = ((x - target) ** 2).sum(dim=1).mean()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if k >= warmup:
            fwd.append(t1 - t0)
            tot.append(t2 - t0)
    out = {"workload": case.name, "bodies": case.num_bodies, "vertices": int(r.num_vertices),
           "edges": int(r.num_edges), "frames": case.frames, "iterations": case.iterations,
           "checkpoint_interval": case.checkpoint_interval,
           "fwd_ms_per_frame": round(1000.0 * min(fwd) / case.frames, 4),
           "fwd_bwd_ms_per_frame": round(1000.0 * min(tot) / case.frames, 4),
           "fwd_bwd_ms_total": round(1000.0 * min(tot), 3),
           "loss": float(loss.detach().cpu()),
           "grad_norms": {"compliance": float(comp.grad.norm().cpu()), "velocity": float(vel.grad.norm().cpu())}}
    r.close()
    ctx.close()
    return out


def large_body_run(steps=3, warmup=1):
    """The large single body (64^3-cell tet block, 274,625 vertices, edge + volume + 270,529 LRA
    constraints among 512 static boxes, 8 collision steps x 10 iterations) on a team of
    workgroups: exact (the non-parallel group in the reference's order, bit-exact) and with
    the re-coloured non-parallel group.  Times `steps` frames after `warmup` with hipEvents on
    the library stream; the coloured run's positions are compared with the exact run's after
    the same frames."""
    scene = scenes.config5_block(frames=warmup + steps)
    t0 = time.time()
    settings = [build_settings(b.mesh) for b in scene.bodies]
    for b, s in zip(scene.bodies, settings):
        if b.post is not None:
            b.post(s)
    log(f"large body {scene.name}: {len(settings[0].vertices)} vertices, {settings[0].num_constraints()} constraints, "
        f"settings built in {time.time() - t0:.1f}s")
    solves = scenes.solves_per_frame(scene, settings)
    bytes_frame = algorithmic_bytes_per_frame(scene, settings)
    out = {"workload": scene.name, "vertices": int(len(settings[0].vertices)),
           "constraints": int(settings[0].num_constraints()), "solves_per_frame": int(solves),
           "collision_steps": scene.collision_steps, "iterations": scene.bodies[0].params.num_iterations,
           "colliders": len(scene.colliders), "frames_timed": steps}
    final = {}
    # exact: the reference's order; coloured: the team's re-coloured non-parallel group;
    # coloured_schedule: settings under XPBD_SCHEDULE_COLOURED (every group coloured), the
    # non-parallel group's colours run as the team's colour passes
    for mode in ("exact", "coloured", "coloured_schedule"):
        ctx = Context(0)
        ctx.set_gravity(scene.gravity)
        ctx.set_colliders(scene.colliders)
        st = ctx.create_settings(settings[0], "coloured" if mode == "coloured_schedule" else "exact")
        body = ctx.create_body(st, scene.bodies[0].position, scene.bodies[0].rotation, scene.bodies[0].params)
        if mode != "exact":
            ctx.set_tail_mode(body, True)
        ctx.update(scene.dt, scene.collision_steps, warmup)
        ctx.synchronize()
        ctx.profile_enable(True)
        t0 = time.perf_counter()
        ctx.update(scene.dt, scene.collision_steps, steps)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        n_launch, kernel_ms = ctx.profile_read()
        ctx.profile_enable(False)
        kind, _, _, team = ctx.placement(body)
        stats = st.stats()
        final[mode] = ctx.read(body)["pos"]
        ctx.close()
        achieved = bytes_frame * steps / (kernel_ms / 1000.0) / 1e9
        out[mode] = {"ms_per_frame": round(1000.0 * wall / steps, 3), "solves_per_s": solves * steps / wall,
                     "placement": kind, "team_workgroups": team,
                     "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(achieved / HBM_PEAK_GBS, 4),
                                  "algorithmic_bytes_per_launch": bytes_frame * steps,
                                  "launch_ms": round(kernel_ms / max(n_launch, 1), 3)}}
        if mode == "coloured":
            out[mode]["colours"] = stats["tail_colours"]
        log(f"large body, {mode}: {out[mode]['ms_per_frame']} ms per frame ({kind}, {team} workgroups)")
    out["parallel_windows"], out["tail_levels"], out["tail_windows"] = (stats["parallel_windows"], stats["tail_levels"],
                                                                        stats["tail_windows"])
    for mode in ("coloured", "coloured_schedule"):
        d = np.abs(final[mode].astype(np.float64) - final["exact"])
        out[mode]["vs_exact_after_frames"] = {"frames": warmup + steps, "max_abs_m": float(d.max()), "mean_abs_m": float(d.mean())}
        out[mode]["speedup"] = round(out["exact"]["ms_per_frame"] / out[mode]["ms_per_frame"], 2)
    return scene, settings, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batched2048")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-diff", action="store_true", help="skip the fwd+bwd differentiation measurement")
    ap.add_argument("--no-large", action="store_true", help="skip the large single body measurement")
    ap.add_argument("--only-large", action="store_true", help="only the large single body measurement (profiling)")
    ap.add_argument("--schedule", default="exact", choices=("exact", "coloured", "coloured_tail"),
                    help="constraint schedule of the timed workload (exact: the reference's order)")
    ap.add_argument("--no-throughput", action="store_true", help="skip the coloured-schedule measurement")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the known-size reference "
                         "for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "independent bodies would shard across GPUs by cost with no exchange", file=sys.stderr)
        sys.exit(2)

    if args.only_large:
        _, _, large = large_body_run(steps=args.steps, warmup=args.warmup)
        print(json.dumps({"large_body": large}), flush=True)
        return

    scene = make_workload(args.workload)
    t0 = time.time()
    settings = scenes.resolve_meshes(scene, build_settings)
    log(f"workload {scene.name}: {len(settings)} bodies, {sum(len(s.vertices) for s in settings)} vertices, "
        f"{sum(s.num_constraints() for s in settings)} constraints, settings built in {time.time() - t0:.1f}s")

    wall, n_launch, kernel_ms, parity = timed_run(scene, settings, args.schedule, args.warmup, args.steps)
    solves = scenes.solves_per_frame(scene, settings)
    value = solves * args.steps / wall
    throughput = None
    if args.schedule == "exact" and not args.no_throughput:
        # the coloured schedule on the same workload: fewer, wider level passes
        t_wall, t_launch, t_kernel_ms, t_parity = timed_run(scene, settings, "coloured", args.warmup, args.steps)
        t_achieved = algorithmic_bytes_per_frame(scene, settings) * args.steps / (t_kernel_ms / 1000.0) / 1e9
        throughput = {"schedule": "coloured", "value": solves * args.steps / t_wall, "unit": "constraint-solves/s",
                      "ms_per_step": 1000.0 * t_wall / args.steps, "launch_ms": round(t_kernel_ms / max(t_launch, 1), 3),
                      "roofline": {"bound": "hbm", "achieved": round(t_achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(t_achieved / HBM_PEAK_GBS, 4)},
                      "parity": t_parity}
    bytes_frame = algorithmic_bytes_per_frame(scene, settings)
    achieved = bytes_frame * args.steps / (kernel_ms / 1000.0) / 1e9
    traffic, traffic_src = measured_traffic(scene.name)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if traffic is None else traffic * args.steps,
                "traffic_source": traffic_src,
                "kernel": "one xpbd_update: bundle_kernel launches (one per launch class) + update_kernel (streamed "
                          "bodies), concurrent on one stream per hardware queue, timed fork to join",
                "algorithmic_bytes_per_launch": bytes_frame * args.steps,
                "launch_ms": round(kernel_ms / max(n_launch, 1), 3)}
    diff = None if args.no_diff else differentiation_run()
    large = None
    if not args.no_large:
        large_scene, large_settings, large = large_body_run()
    cpu = None if args.no_cpu_baseline else cpu_baseline(scene, settings, args.cpu_seconds)
    if large is not None and not args.no_cpu_baseline:
        large_cpu = cpu_baseline(large_scene, large_settings, target_seconds=8.0)
        large["cpu_baseline"] = large_cpu
        if large_cpu:
            large["gpu_over_cpu"] = {m: round(large[m]["solves_per_s"] / large_cpu["value"], 2)
                                     for m in ("exact", "coloured", "coloured_schedule")}
    out = {
        "metric": "constraint-solves/sec", "value": value, "unit": "constraint-solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": scene.name, "bodies": len(settings),
                   "vertices": int(sum(len(s.vertices) for s in settings)),
                   "constraints": int(sum(s.num_constraints() for s in settings)),
                   "solves_per_frame": int(solves), "collision_steps": scene.collision_steps,
                   "iterations": int(scene.bodies[0].params.num_iterations), "dt": scene.dt,
                   "schedule": args.schedule, "parity": parity},
        "roofline": roofline, "cpu_baseline": cpu, "differentiation": diff, "large_body": large,
        "throughput_schedule": throughput,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
        if throughput:
            throughput["gpu_over_cpu"] = round(throughput["value"] / cpu["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
