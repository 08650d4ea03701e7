// Soft-body XPBD update kernels for gfx950 (MI355X).
//
// One workgroup owns one soft body for a whole xpbd_update call (all frames,
// collision steps and iterations).
//
// Constraint projection follows the reference's update groups
// (SoftBodyMotionProperties.cpp:1030-1115).  Inside a group the reference
// applies its constraints sequentially (volume, bend, edge, LRA order); the
// host turns that sequence into levels in which no two constraints touch the
// same vertex, ordered so that every read-after-write / write-after-read pair
// keeps its sequential order.  The parallel groups are vertex-disjoint, so
// their levels are merged (level L of every group runs together) and the
// non-parallel group runs after them.  A level runs lane-parallel across the
// workgroup with a barrier after it, which gives bit-identical results to the
// reference's sequential loops.
//
// Two storage modes, chosen per body by the host:
//   RESIDENT  positions + inverse masses stay in LDS (float4) for the whole
//             launch; the level windows index body vertices directly;
//   streamed  positions live in global memory; each window of levels gathers
//             its vertex subset into LDS, runs, and scatters it back.
// STAGED additionally copies the constraint records and level offsets into LDS
// once per launch (they are reused every iteration of every step).
//
// All arithmetic follows the reference's operation order (jolt_math.h); the
// file is compiled with -ffp-contract=off and IEEE division / sqrt.

#include <cfloat>

#include "xpbd_device.h"

namespace xpbd {

constexpr int kMaxWaves = 16;       // 1024-thread workgroups

// Dynamic LDS layout:
//   float4 X[xcount]            positions (resident: all vertices, streamed: window)
//   uint4  recs[nrec]           staged records
//   uint32 lvl[nlvl]            staged level offsets
//   Leaf   leaves[max_leaves]
//   int    collider_slot[num_colliders]
//   float4 scratch[num_faces]   (pressure)
//   float  red[kMaxWaves * 32]  (reductions)
struct LdsLayout
{
    uint32_t recs, lvl, leaves, slots, scratch, red, total;
};

__host__ __device__ inline LdsLayout lds_layout(uint32_t xcount, uint32_t nrec, uint32_t nlvl, uint32_t max_leaves, uint32_t nc, uint32_t nf)
{
    LdsLayout l;
    uint32_t off = xcount * 16u;
    l.recs = off; off += nrec * 16u;
    l.lvl = off; off += nlvl * 4u;
    off = (off + 15u) & ~15u; l.leaves = off; off += max_leaves * (uint32_t)sizeof(Leaf);
    off = (off + 15u) & ~15u; l.slots = off; off += nc * 4u;
    off = (off + 15u) & ~15u; l.scratch = off; off += nf * 16u;
    off = (off + 15u) & ~15u; l.red = off; off += kMaxWaves * 32u * 4u;
    l.total = (off + 15u) & ~15u;
    return l;
}

// Runs `nlev` levels on the window X.  Records of one level touch disjoint
// vertices, so the workgroup projects a level lane-parallel and a barrier
// separates consecutive levels.
//
// STAGED: records and level offsets are in LDS.  Otherwise they stream from
// global memory: each thread holds its record of the next two levels in
// registers (issued two levels ahead), so the load latency overlaps two levels
// of projection.
template <bool STAGED, int BLOCK, bool RODS>
__device__ __forceinline__ void run_levels(const uint4 *recs, const uint32_t *lvl, uint32_t nlev, float4 *X, float inv_dt_sq, int tid,
                                           const RecordCtx &sk, bool gsync)
{
    if (STAGED)
        run_levels_lds<BLOCK, RODS>(recs, lvl, nlev, X, inv_dt_sq, (uint32_t)tid, NoSkip(), sk, gsync);
    else
        run_levels_global<BLOCK, RODS>(recs, lvl, nlev, X, inv_dt_sq, (uint32_t)tid, NoSkip(), sk, gsync);
}

// ---------------------------------------------------------------- kernel
// Advances each body of `body_list` by `frames` x `steps` collision steps.  RODS: the
// variant for bodies with Cosserat rods.
template <bool RESIDENT, bool STAGED, int kBlock, bool RODS, bool EXT>
__global__ __launch_bounds__(kBlock) void update_kernel(
    const BodyDev *__restrict__ bodies, BodyStateDev *__restrict__ states, const uint32_t *__restrict__ body_list,
    const ColliderDev *__restrict__ colliders, uint32_t num_colliders, uint32_t max_leaves,
    float gx, float gy, float gz, float dt, int32_t steps, int32_t frames, int32_t *__restrict__ error_flag)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ BodyStateDev st;
    __shared__ float sh_scalars[8];
    __shared__ uint32_t sh_num_leaves;
    __shared__ uint32_t sh_num_dyn_leaves;
    __shared__ Leaf sh_sensor_leaf;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid >> 6;
    const uint32_t bi = body_list[blockIdx.x];
    const BodyDev &B = bodies[bi];
    const SettingsDev &S = B.s;
    const uint32_t nv = S.num_vertices;
    const bool has_pressure = B.pressure > 0.0f && S.num_faces > 0;
    const LdsLayout lay = lds_layout(RESIDENT ? nv : S.window_capacity, STAGED ? S.num_records : 0u,
                                     STAGED ? S.num_levels_total : 0u, max_leaves, num_colliders,
                                     has_pressure ? S.num_faces : 0u);
    float4 *X = reinterpret_cast<float4 *>(lds_raw);
    const uint4 *recs = STAGED ? reinterpret_cast<const uint4 *>(lds_raw + lay.recs) : S.recs;
    const uint32_t *lvl = STAGED ? reinterpret_cast<const uint32_t *>(lds_raw + lay.lvl) : S.lvl;
    Leaf *leaves = reinterpret_cast<Leaf *>(lds_raw + lay.leaves);
    int32_t *slots = reinterpret_cast<int32_t *>(lds_raw + lay.slots);
    float4 *scratch = reinterpret_cast<float4 *>(lds_raw + lay.scratch);
    float *red = reinterpret_cast<float *>(lds_raw + lay.red);
    // Per-vertex phases address positions here
    float4 *P = RESIDENT ? X : B.x;
    // Cosserat rods: their states stay in global memory (2 float4 per rod)
    const uint32_t num_rods = RODS && B.rods != nullptr ? S.num_rods : 0u;
    const bool has_rods = num_rods > 0;
    // barrier after a per-vertex phase: streamed positions (and rod states) are handed
    // between wavefronts through global memory
    auto sync_positions = [&]() {
        if (RESIDENT && !has_rods)
            __syncthreads();
        else
            sync_global();
    };

    if (tid == 0)
        st = states[bi];
    if (STAGED)
    {
        uint4 *rd = reinterpret_cast<uint4 *>(lds_raw + lay.recs);
        for (uint32_t i = tid; i < S.num_records; i += kBlock)
            rd[i] = S.recs[i];
        uint32_t *ld = reinterpret_cast<uint32_t *>(lds_raw + lay.lvl);
        for (uint32_t i = tid; i < S.num_levels_total; i += kBlock)
            ld[i] = S.lvl[i];
    }
    if (RESIDENT)
        for (uint32_t i = tid; i < nv; i += kBlock)
            X[i] = B.x[i];
    // dynamic colliders: the call's velocity-change sums start at zero
    DynStateDev *const dyn = B.dyn;
    const bool has_dynamic = dyn != nullptr && B.num_dynamic > 0;
    if (has_dynamic)
        for (uint32_t k = tid; k < B.num_dynamic; k += kBlock)
            clear_dynamic(dyn[k]);
    __syncthreads();

    const V3 gravity_world = v3(gx, gy, gz);
    // PhysicsSystem::Update: step delta time (PhysicsSystem.cpp:215)
    const float step_dt = dt / (float)steps;
    const uint32_t iterations = B.iterations;
    const uint32_t num_windows = S.num_par_windows + S.num_tail_windows;

    bool report_cleared = false;
    for (int32_t frame = 0; frame < frames; ++frame)
        for (int32_t step = 0; step < steps; ++step)
        {
            if (!st.active)
            {
                // a body that does not run in a step reports no contacts for it
                if (!report_cleared)
                {
                    for (uint32_t i = tid; i < nv; i += kBlock)
                        B.report[i] = contact_report(v3(0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f), false, -1);
                    if (B.sensor_hits)
                        for (uint32_t k = tid; k < B.num_sensors; k += kBlock)
                            B.sensor_hits[k] = 0;
                }
                report_cleared = true;
                break;
            }

            // InitializeUpdateContext (SoftBodyMotionProperties.cpp:937-960)
            const M44 com = rotation_translation(st.rot[0], st.rot[1], st.rot[2], st.rot[3], v3(st.pos[0], st.pos[1], st.pos[2]));
            const V3 g_local = mul3x3_transposed(com, B.gravity_factor * gravity_world);
            const float dt_sub = step_dt / (float)iterations;
            const V3 disp_gravity = (((0.5f * (float)iterations) * (float)(iterations + 1)) * (dt_sub * dt_sub)) * g_local;
            const float inv_dt_sq = 1.0f / (dt_sub * dt_sub);

            // DetermineCollidingShapes (SoftBodyMotionProperties.cpp:236-253), static colliders
            {
                V3 lmin = vmin(v3(st.bmin[0], st.bmin[1], st.bmin[2]), v3(st.pmin[0], st.pmin[1], st.pmin[2]));
                V3 lmax = vmax(v3(st.bmax[0], st.bmax[1], st.bmax[2]), v3(st.pmax[0], st.pmax[1], st.pmax[2]));
                V3 r = v3(B.vertex_radius, B.vertex_radius, B.vertex_radius);
                lmin = lmin - r;
                lmax = lmax + r;
                // AABox::Transformed (Geometry/AABox.h:193-213)
                V3 wmin = col3(com, 3), wmax = col3(com, 3);
                for (int c = 0; c < 3; ++c)
                {
                    V3 col = col3(com, c);
                    V3 a = col * comp(lmin, c);
                    V3 b = col * comp(lmax, c);
                    wmin = wmin + vmin(a, b);
                    wmax = wmax + vmax(a, b);
                }
                const M44 inv_com = inversed_rotation_translation(com);
                for (uint32_t ci = tid; ci < num_colliders; ci += kBlock)
                {
                    const ColliderDev &cd = colliders[ci];
                    // AABox::Overlaps (Geometry/AABox.h:164-167)
                    bool overlap = !(wmin.x > cd.bounds[3] || wmin.y > cd.bounds[4] || wmin.z > cd.bounds[5]
                                     || wmax.x < cd.bounds[0] || wmax.y < cd.bounds[1] || wmax.z < cd.bounds[2]);
                    slots[ci] = overlap && !pair_skipped(pair_settings(B.pair, cd.owner), cd.sensor >= 0, cd.motion == 2)
                                && part_collected<EXT>(cd, inv_com, lmin, lmax) ? 1 : 0;
                }
                __syncthreads();
                if (tid == 0)
                {
                    uint32_t n = 0, nd = 0;
                    for (uint32_t ci = 0; ci < num_colliders; ++ci)
                    {
                        if (colliders[ci].sensor >= 0)
                        {
                            slots[ci] = slots[ci] ? -2 : -1;    // sensors never collide; -2: overlapping
                            continue;
                        }
                        if (slots[ci])
                        {
                            if (n < max_leaves)
                            {
                                slots[ci] = (int32_t)n;
                                nd += colliders[ci].motion == 2 ? 1u : 0u;
                            }
                            else
                            {
                                slots[ci] = -1;
                                atomicOr(error_flag, 1);
                            }
                            ++n;
                        }
                        else
                            slots[ci] = -1;
                    }
                    sh_num_leaves = n < max_leaves ? n : max_leaves;
                    sh_num_dyn_leaves = has_dynamic ? nd : 0u;
                }
                __syncthreads();
                for (uint32_t ci = tid; ci < num_colliders; ci += kBlock)
                    if (slots[ci] >= 0)
                    {
                        const ColliderDev &cd = colliders[ci];
                        make_leaf<EXT>(cd, inv_com, B.friction, B.restitution, leaves[slots[ci]], dyn, pair_settings(B.pair, cd.owner));
                        leaves[slots[ci]].collider = cd.owner;
                    }
                // dynamic states are read by the in-order contact pass of another wavefront
                if (has_dynamic)
                    sync_global();
                else
                    __syncthreads();
            }
            const uint32_t num_leaves = sh_num_leaves;
            const uint32_t num_dyn_leaves = sh_num_dyn_leaves;

            // DetermineSensorCollisions (SoftBodyMotionProperties.cpp:267-289): a sensor
            // reports contact when a vertex penetrates it
            if (B.sensor_hits)
            {
                const M44 inv_com_s = inversed_rotation_translation(com);
                for (uint32_t ci = 0; ci < num_colliders; ++ci)
                {
                    const int32_t rank = colliders[ci].sensor;
                    if (rank < 0)
                        continue;
                    float largest = -FLT_MAX;
                    if (slots[ci] == -2)
                    {
                        if (tid == 0)
                            make_leaf<EXT>(colliders[ci], inv_com_s, B.friction, B.restitution, sh_sensor_leaf, nullptr);
                        __syncthreads();
                        for (uint32_t i = tid; i < nv; i += kBlock)
                            largest = smax(largest, sensor_penetration<EXT>(sh_sensor_leaf, P[i]));
                    }
                    largest = wave_max(largest);
                    if (lane == 0)
                        red[wave * 32] = largest;
                    __syncthreads();
                    if (tid == 0)
                    {
                        for (int w = 1; w < kBlock / kWave; ++w)
                            largest = smax(largest, red[w * 32]);
                        // a compound sensor is hit when one of its leaves is
                        B.sensor_hits[rank] = largest > 0.0f || (!colliders[ci].first && B.sensor_hits[rank]) ? 1 : 0;
                    }
                    __syncthreads();
                }
            }

            // DetermineCollisionPlanes (SoftBodyMotionProperties.cpp:257-265)
            if (num_leaves > 0)
                for (uint32_t i = tid; i < nv; i += kBlock)
                {
                    float4 p = P[i];
                    if (p.w > 0.0f)
                    {
                        float4 pl = B.plane[i];
                        Pl plane; plane.n = v3(pl.x, pl.y, pl.z); plane.c = pl.w;
                        float largest = -FLT_MAX;
                        int32_t idx = -1;
                        collide_vertex<EXT>(leaves, num_leaves, xyz(p), largest, idx, plane);
                        if (idx >= 0)
                        {
                            B.idx[i] = idx;
                            B.plane[i] = make_float4(plane.n.x, plane.n.y, plane.n.z, plane.c);
                        }
                    }
                }

            // ApplyPressure (SoftBodyMotionProperties.cpp:291-323)
            auto apply_pressure = [&]() {
                const uint32_t nf = S.num_faces;
                for (uint32_t f = tid; f < nf; f += kBlock)
                {
                    V3 x1 = xyz(P[S.faces[3 * f]]), x2 = xyz(P[S.faces[3 * f + 1]]), x3 = xyz(P[S.faces[3 * f + 2]]);
                    scratch[f] = make_float4(0.0f, 0.0f, 0.0f, dot(cross(x1, x2), x3));
                }
                __syncthreads();
                if (tid == 0)
                {
                    // GetVolumeTimesSix: sequential sum in face order (SoftBodyMotionProperties.cpp:107-118)
                    float six_volume = 0.0f;
                    for (uint32_t f = 0; f < nf; ++f)
                        six_volume += scratch[f].w;
                    sh_scalars[0] = six_volume;
                }
                __syncthreads();
                const float six_volume = sh_scalars[0];
                if (six_volume > 0.0f)
                {
                    const float coefficient = B.pressure * dt_sub / six_volume;
                    for (uint32_t f = tid; f < nf; f += kBlock)
                    {
                        V3 x1 = xyz(P[S.faces[3 * f]]), x2 = xyz(P[S.faces[3 * f + 1]]), x3 = xyz(P[S.faces[3 * f + 2]]);
                        V3 impulse = coefficient * cross(x2 - x1, x3 - x1);
                        scratch[f] = make_float4(impulse.x, impulse.y, impulse.z, 0.0f);
                    }
                    __syncthreads();
                    // v_i += w_i * impulse_f for the faces of vertex i in face order
                    for (uint32_t i = tid; i < nv; i += kBlock)
                    {
                        uint32_t fb = S.vf_begin[i], fe = S.vf_begin[i + 1];
                        if (fb == fe)
                            continue;
                        float w = P[i].w;
                        V3 v = xyz(B.v[i]);
                        for (uint32_t k = fb; k < fe; ++k)
                            v = v + w * xyz(scratch[S.vf_list[k]]);
                        B.v[i] = f4(v, 0.0f);
                    }
                }
                __syncthreads();
            };

            // IntegratePositions (SoftBodyMotionProperties.cpp:325-350); accumulated force is zero
            const float damping = smax(0.0f, 1.0f - B.linear_damping * dt_sub);
            const V3 sub_step_gravity = g_local * dt_sub;
            const V3 sub_step_impulse = (v3(0.0f, 0.0f, 0.0f) * dt_sub) / smax((float)nv, 1.0f);
            auto integrate_vertex = [&](uint32_t i, float4 p, V3 v) {
                if (p.w > 0.0f)
                {
                    v = v + (sub_step_gravity + sub_step_impulse * p.w);
                    v = v * damping;
                }
                V3 position = xyz(p);
                B.prev[i] = f4(position, 0.0f);
                P[i] = f4(position + v * dt_sub, p.w);
                B.v[i] = f4(v, 0.0f);
            };

            // IntegratePositions, rod orientations (SoftBodyMotionProperties.cpp:352-364): the
            // angular velocity is damped and integrated; the slot keeps the previous rotation
            const float half_dt = 0.5f * dt_sub;
            auto integrate_rod = [&](uint32_t r, Q4 rotation, V3 angular_velocity) {
                angular_velocity = angular_velocity * damping;
                const Q4 delta_rotation = mul_imaginary(angular_velocity, rotation) * half_dt;
                B.rods[2u * r + 1u] = make_float4(rotation.x, rotation.y, rotation.z, rotation.w);
                const Q4 n = normalized(rotation + delta_rotation);
                B.rods[2u * r] = make_float4(n.x, n.y, n.z, n.w);
            };
            // ApplyCollisionConstraintsAndUpdateVelocities, rods (SoftBodyMotionProperties.cpp:821-824):
            // the angular velocity from the rotation change, then the next sub-step's integration
            const float two_div_dt = 2.0f / dt_sub;
            auto update_rods = [&](bool integrate_next) {
                for (uint32_t r = tid; r < num_rods; r += kBlock)
                {
                    const float4 s = B.rods[2u * r], p = B.rods[2u * r + 1u];
                    const Q4 rotation = q4(s.x, s.y, s.z, s.w);
                    const Q4 d = quat_mul(rotation, conjugated(q4(p.x, p.y, p.z, p.w)));
                    const V3 angular_velocity = two_div_dt * v3(d.x, d.y, d.z);
                    if (integrate_next)
                        integrate_rod(r, rotation, angular_velocity);
                    else
                        B.rods[2u * r + 1u] = make_float4(angular_velocity.x, angular_velocity.y, angular_velocity.z, 0.0f);
                }
            };

            if (has_pressure)
                apply_pressure();
            for (uint32_t i = tid; i < nv; i += kBlock)
                integrate_vertex(i, P[i], xyz(B.v[i]));
            for (uint32_t r = tid; r < num_rods; r += kBlock)
            {
                const float4 s = B.rods[2u * r], w = B.rods[2u * r + 1u];
                integrate_rod(r, q4(s.x, s.y, s.z, s.w), v3(w.x, w.y, w.z));
            }
            sync_positions();

            const float restitution_threshold = -2.0f * len(g_local) * dt_sub;
            const float max_lin_sq = B.max_linear_velocity * B.max_linear_velocity;
            float max_v_sq = 0.0f;
            V3 bmin = v3(FLT_MAX, FLT_MAX, FLT_MAX), bmax = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            V3 pmin = bmin, pmax = bmax;
            V3 lin = v3(0.0f, 0.0f, 0.0f), ang = v3(0.0f, 0.0f, 0.0f);

            for (uint32_t it = 0; it < iterations; ++it)
            {
                const bool last = it + 1 == iterations;

                // Parallel-group windows, then the non-parallel group's windows
                for (uint32_t w = 0; w < num_windows; ++w)
                {
                    const WindowDev W = S.windows[w];
                    if (!RESIDENT && W.vcount)
                    {
                        for (uint32_t j = tid; j < W.vcount; j += kBlock)
                            X[j] = B.x[S.vlist[W.vbegin + j]];
                        __syncthreads();
                    }
                    run_levels<STAGED, kBlock, RODS>(recs, lvl + W.lbegin, W.nlev, X, inv_dt_sq, tid,
                                               RecordCtx{ B.skin, B.skin_base, (float)(it + 1) / (float)iterations,
                                                          B.rods, S.rod_const, S.bt_const },
                                               has_rods);
                    if (!RESIDENT && W.vcount)
                    {
                        for (uint32_t j = tid; j < W.vcount; j += kBlock)
                            B.x[S.vlist[W.vbegin + j]] = X[j];
                        sync_global();
                    }
                }

                // ApplyCollisionConstraintsAndUpdateVelocities (SoftBodyMotionProperties.cpp:700-825);
                // fused with the next iteration's integration when there is no pressure, or with
                // UpdateSoftBodyState's per-vertex work after the last one.  A vertex in contact
                // with a dynamic collider changes that body's velocities, which the next such
                // vertex reads: the parallel pass marks these vertices (contact 2) and leaves
                // them; thread 0 then runs them in vertex order.
                auto collide_update = [&](uint32_t i, bool in_order) {
                    float4 p = P[i];
                    V3 x = xyz(p);
                    V3 v = xyz(B.v[i]);
                    if (p.w > 0.0f)
                    {
                        V3 prev_v = v;
                        v = (x - xyz(B.prev[i])) / dt_sub;
                        int32_t idx = B.idx[i];
                        if (idx >= 0)
                        {
                            float4 pl = B.plane[i];
                            Pl plane; plane.n = v3(pl.x, pl.y, pl.z); plane.c = pl.w;
                            float projected_distance = -signed_distance(plane, x) + B.vertex_radius;
                            if (projected_distance > 0.0f)
                            {
                                const Leaf &cs = leaves[idx];
                                if (cs.dyn >= 0 && !in_order)
                                {
                                    B.contact[i] = 2;
                                    return;
                                }
                                B.contact[i] = 1;
                                V3 n = plane.n;
                                x = x + n * projected_distance;
                                if (cs.dyn >= 0)
                                    dynamic_contact(dyn[cs.dyn], ld3(dyn[cs.dyn].com), x, v, prev_v, n, projected_distance, p.w,
                                                    cs.friction, cs.restitution, dt_sub, restitution_threshold,
                                                    cs.inv_mass_scale1);
                                else if (cs.inv_mass_scale1 > 0.0f)
                                {
                                    V3 v_normal = n * dot(n, v);
                                    V3 v_tangential = v - v_normal;
                                    float vtl = len(v_tangential);
                                    if (vtl > 0.0f)
                                    {
                                        float f = cs.friction * projected_distance / (vtl * dt_sub);
                                        v = v - v_tangential * (f < 1.0f ? f : 1.0f);
                                    }
                                    v = v - v_normal;
                                    float prev_v_normal = dot(prev_v, n);
                                    if (prev_v_normal < restitution_threshold)
                                        v = v - (cs.restitution * prev_v_normal) * n;
                                }
                            }
                        }
                    }
                    if (!last)
                    {
                        if (has_pressure)
                        {
                            P[i] = f4(x, p.w);
                            B.v[i] = f4(v, 0.0f);
                        }
                        else
                            integrate_vertex(i, f4(x, p.w), v);
                    }
                    else
                    {
                        // UpdateSoftBodyState per-vertex part (SoftBodyMotionProperties.cpp:852-880)
                        float v_sq = len_sq(v);
                        max_v_sq = smax(max_v_sq, v_sq);
                        if (v_sq > max_lin_sq)
                            v = v * sqrtf(max_lin_sq / v_sq);
                        lin = lin + v;
                        ang = ang + cross(x, v);
                        bmin = vmin(bmin, x);
                        bmax = vmax(bmax, x);
                        V3 pp = (x + v * step_dt) + disp_gravity;
                        pmin = vmin(pmin, pp);
                        pmax = vmax(pmax, pp);
                        P[i] = f4(x, p.w);
                        B.v[i] = f4(v, 0.0f);
                        const int32_t hit = B.idx[i];
                        B.report[i] = contact_report(x, B.plane[i], B.contact[i] != 0 && hit >= 0,
                                                     hit >= 0 ? leaves[hit].collider : -1);
                        B.idx[i] = -1;
                        B.contact[i] = 0;
                    }
                };
                for (uint32_t i = tid; i < nv; i += kBlock)
                    collide_update(i, false);
                if (num_dyn_leaves > 0)
                {
                    // the marks were written by other wavefronts
                    sync_global();
                    if (wave == 0)
                        for (uint32_t k0 = 0; k0 < nv; k0 += kWave)
                        {
                            const uint32_t k = k0 + (uint32_t)lane;
                            unsigned long long marked = __ballot(k < nv && B.contact[k] == 2);
                            while (marked)
                            {
                                const uint32_t j = (uint32_t)__builtin_ctzll(marked);
                                marked &= marked - 1ull;
                                if (lane == 0)
                                    collide_update(k0 + j, true);
                            }
                        }
                    sync_global();
                }
                if (has_rods)
                    update_rods(!last);
                sync_positions();
                if (!last && has_pressure)
                {
                    apply_pressure();
                    for (uint32_t i = tid; i < nv; i += kBlock)
                        integrate_vertex(i, P[i], xyz(B.v[i]));
                    sync_positions();
                }
            }

            // UpdateSoftBodyState reductions (SoftBodyMotionProperties.cpp:846-921)
            {
                float vals[13] = { bmin.x, bmin.y, bmin.z, pmin.x, pmin.y, pmin.z,
                                   bmax.x, bmax.y, bmax.z, pmax.x, pmax.y, pmax.z, max_v_sq };
#pragma unroll
                for (int k = 0; k < 6; ++k) vals[k] = wave_min(vals[k]);
#pragma unroll
                for (int k = 6; k < 13; ++k) vals[k] = wave_max(vals[k]);
                float sums[6] = { wave_sum(lin.x), wave_sum(lin.y), wave_sum(lin.z), wave_sum(ang.x), wave_sum(ang.y), wave_sum(ang.z) };
                if (lane == 0)
                {
#pragma unroll
                    for (int k = 0; k < 13; ++k) red[wave * 32 + k] = vals[k];
#pragma unroll
                    for (int k = 0; k < 6; ++k) red[wave * 32 + 16 + k] = sums[k];
                }
                __syncthreads();
                if (tid == 0)
                {
                    for (int w = 1; w < kBlock / kWave; ++w)
                    {
#pragma unroll
                        for (int k = 0; k < 6; ++k) vals[k] = smin(vals[k], red[w * 32 + k]);
#pragma unroll
                        for (int k = 6; k < 13; ++k) vals[k] = smax(vals[k], red[w * 32 + k]);
#pragma unroll
                        for (int k = 0; k < 6; ++k) sums[k] += red[w * 32 + 16 + k];
                    }
                    V3 bmn = v3(vals[0], vals[1], vals[2]), pmn = v3(vals[3], vals[4], vals[5]);
                    V3 bmx = v3(vals[6], vals[7], vals[8]), pmx = v3(vals[9], vals[10], vals[11]);
                    float mvsq = vals[12];
                    float divider = (float)(nv > 1 ? nv : 1u);
                    // body velocity (sums in tree order, the reference sums sequentially)
                    V3 lv = mul3x3(com, v3(sums[0], sums[1], sums[2]) / divider);
                    V3 av = mul3x3(com, v3(sums[3], sums[4], sums[5]) / divider);
                    st.lin[0] = lv.x; st.lin[1] = lv.y; st.lin[2] = lv.z;
                    st.ang[0] = av.x; st.ang[1] = av.y; st.ang[2] = av.z;
                    if (has_dynamic)
                        for (uint32_t k = 0; k < B.num_dynamic; ++k)
                            finish_dynamic(dyn[k], com);
                    V3 delta = v3(0.0f, 0.0f, 0.0f);
                    if (B.update_position)
                    {
                        delta = 0.5f * (bmn + bmx);                   // AABox::GetCenter
                        V3 dw = mul3x3(com, delta);
                        V3 nd = neg(delta);
                        bmn = bmn + nd; bmx = bmx + nd; pmn = pmn + nd; pmx = pmx + nd;
                        // JobSoftBodyFinalize: GetPosition() + mDeltaPosition (PhysicsSystem.cpp:2864)
                        st.pos[0] = st.pos[0] + dw.x; st.pos[1] = st.pos[1] + dw.y; st.pos[2] = st.pos[2] + dw.z;
                    }
                    st.bmin[0] = bmn.x; st.bmin[1] = bmn.y; st.bmin[2] = bmn.z;
                    st.bmax[0] = bmx.x; st.bmax[1] = bmx.y; st.bmax[2] = bmx.z;
                    st.pmin[0] = pmn.x; st.pmin[1] = pmn.y; st.pmin[2] = pmn.z;
                    st.pmax[0] = pmx.x; st.pmax[1] = pmx.y; st.pmax[2] = pmx.z;
                    sh_scalars[1] = delta.x; sh_scalars[2] = delta.y; sh_scalars[3] = delta.z;
                    if (B.allow_sleeping)
                    {
                        // sleep test (SoftBodyMotionProperties.cpp:903-914, MotionProperties.inl:169-173)
                        if (mvsq > 0.03f)
                            st.sleep_timer = 0.0f;
                        else
                        {
                            st.sleep_timer += step_dt;
                            if (st.sleep_timer >= 0.5f)
                                st.active = 0;
                        }
                    }
                }
                __syncthreads();
                if (B.update_position)
                {
                    V3 delta = v3(sh_scalars[1], sh_scalars[2], sh_scalars[3]);
                    for (uint32_t i = tid; i < nv; i += kBlock)
                    {
                        float4 p = P[i];
                        P[i] = f4(xyz(p) - delta, p.w);
                    }
                }
                if (B.skin)
                {
                    // skin states (SoftBodyMotionProperties.cpp:890-917): the skinned positions
                    // follow the recentring; previous positions are invalid until the next SkinVertices
                    float4 *e = const_cast<float4 *>(B.skin) + 4u * (size_t)B.skin_base;
                    const V3 delta = v3(sh_scalars[1], sh_scalars[2], sh_scalars[3]);
                    for (uint32_t i = tid; i < nv; i += kBlock)
                    {
                        if (B.update_position)
                        {
                            const float4 sp = e[4u * i + 1u];
                            e[4u * i + 1u] = f4(xyz(sp) - delta, sp.w);
                        }
                        e[4u * i + 3u].w = 0.0f;
                    }
                }
                __syncthreads();
            }
        }

    if (RESIDENT)
        for (uint32_t i = tid; i < nv; i += kBlock)
            B.x[i] = X[i];
    if (tid == 0)
        states[bi] = st;
}

} // namespace xpbd

// ---------------------------------------------------------------- launchers
namespace xpbd {

uint32_t kernel_lds_bytes(bool resident, bool staged, const SettingsDev &s, uint32_t max_leaves, uint32_t nc, uint32_t nf)
{
    return lds_layout(resident ? s.num_vertices : s.window_capacity, staged ? s.num_records : 0u,
                      staged ? s.num_levels_total : 0u, max_leaves, nc, nf).total;
}

template <bool R, bool S, int BLOCK, bool RODS, bool EXT>
static hipError_t launch_variant(const BodyDev *bodies, BodyStateDev *states, const uint32_t *body_list, uint32_t num_blocks,
                                 const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves, uint32_t lds_bytes,
                                 float gx, float gy, float gz, float dt, int32_t steps, int32_t frames, int32_t *error_flag,
                                 hipStream_t stream)
{
    static bool configured = false;
    if (!configured)
    {
        hipError_t e = hipFuncSetAttribute((const void *)update_kernel<R, S, BLOCK, RODS, EXT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
        if (e != hipSuccess)
            return e;
        configured = true;
    }
    hipLaunchKernelGGL((update_kernel<R, S, BLOCK, RODS, EXT>), dim3(num_blocks), dim3(BLOCK), lds_bytes, stream,
                       bodies, states, body_list, colliders, num_colliders, max_leaves, gx, gy, gz, dt, steps, frames, error_flag);
    return hipGetLastError();
}

template <bool RODS, bool EXT>
static hipError_t launch_update_rods(bool resident, bool staged, const BodyDev *bodies, BodyStateDev *states, const uint32_t *body_list,
                                     uint32_t num_blocks, const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves,
                                     uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                                     int32_t *error_flag, hipStream_t stream)
{
    if (resident && staged)
        return launch_variant<true, true, 256, RODS, EXT>(bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream);
    if (resident)
        return launch_variant<true, false, 256, RODS, EXT>(bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream);
    return launch_variant<false, false, 1024, RODS, EXT>(bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream);
}

// Resident single-body launches use 256-thread workgroups; streamed bodies use
// 1024 threads so one pass covers the wide merged levels of an 8192-vertex window.
// Bodies with Cosserat rods run in the variants that hold the rod code.
hipError_t launch_update(bool resident, bool staged, bool rods, bool ext, const BodyDev *bodies, BodyStateDev *states, const uint32_t *body_list,
                         uint32_t num_blocks, const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves,
                         uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                         int32_t *error_flag, hipStream_t stream)
{
    if (num_blocks == 0)
        return hipSuccess;
    if (rods)
        return ext ? launch_update_rods<true, true>(resident, staged, bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream)
                   : launch_update_rods<true, false>(resident, staged, bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream);
    return ext ? launch_update_rods<false, true>(resident, staged, bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream)
               : launch_update_rods<false, false>(resident, staged, bodies, states, body_list, num_blocks, colliders, num_colliders, max_leaves, lds_bytes, gx, gy, gz, dt, steps, frames, error_flag, stream);
}

} // namespace xpbd
