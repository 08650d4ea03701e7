// Soft-body XPBD update for large bodies on a team of workgroups (gfx950).
//
// A body whose parallel update groups fill several LDS windows runs on a team of
// 1024-thread workgroups, one per CU.  Per iteration of a collision step:
//   1. each member gathers its share of the parallel-group windows into LDS, runs
//      their levels and scatters them back (the groups are vertex-disjoint, so the
//      windows of different members never share a vertex; SoftBodyMotionProperties.cpp:
//      1057-1115 spreads the same groups over its job threads);
//   2. the non-parallel group: member 0 runs its windows in the reference's order
//      (exact), or, in the coloured mode, every member projects its share of one
//      colour of the re-coloured group at a time, straight in global memory;
//   3. the per-vertex phases (collision planes, integration, collision and velocity
//      update, UpdateSoftBodyState) run over each member's slice of the vertices.
// Members hand positions to each other through global memory across a team
// barrier: every wavefront's stores complete, one thread releases at agent scope
// (the members may sit on different XCDs, whose L2 caches are not coherent with each
// other) and counts itself in, waits for the team's count, and acquires at agent scope.
// Every wait is bounded: a barrier that waits longer than kTeamSpinTicks sets an
// error bit, and from then on no barrier of the launch waits, so the launch always
// drains.
//
// The team path covers static and kinematic colliders, contact reports, sleeping,
// skinned constraints and the position update; bodies with rods, pressure, dynamic or
// sensor colliders run in the single-workgroup kernel.  All arithmetic follows the
// reference's operation order (jolt_math.h); compiled with -ffp-contract=off.

#include <cfloat>

#include "xpbd_device.h"

namespace xpbd {

constexpr int kTeamBlock = 1024;
constexpr int kTeamWaves = kTeamBlock / kWave;
// error bit of a barrier that waited too long (the realtime counter runs at 100 MHz: 2 s)
constexpr int32_t kErrTeamTimeout = 4;
constexpr uint64_t kTeamSpinTicks = 200000000ull;

struct TeamLds
{
    uint32_t leaves, slots, red, total;
};

__host__ __device__ inline TeamLds team_lds_layout(uint32_t window_capacity, uint32_t max_leaves, uint32_t nc)
{
    TeamLds l;
    uint32_t off = window_capacity * 16u;
    off = (off + 15u) & ~15u; l.leaves = off; off += max_leaves * (uint32_t)sizeof(Leaf);
    off = (off + 15u) & ~15u; l.slots = off; off += nc * 4u;
    off = (off + 15u) & ~15u; l.red = off; off += kTeamWaves * 32u * 4u;
    l.total = (off + 15u) & ~15u;
    return l;
}

// Team barrier.  `target` = team size x number of this barrier in the launch.
__device__ __forceinline__ void team_barrier(uint32_t *counter, uint32_t target, int32_t *error_flag)
{
    __builtin_amdgcn_s_waitcnt(kWaitVmcntZero);     // this wavefront's global stores have completed
    __syncthreads();
    if (threadIdx.x == 0)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while ((int32_t)(__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0)
        {
            if (__hip_atomic_load(error_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kErrTeamTimeout)
                break;
            __builtin_amdgcn_s_sleep(1);
            if (__builtin_amdgcn_s_memrealtime() - t0 > kTeamSpinTicks)
            {
                __hip_atomic_fetch_or(error_flag, kErrTeamTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Positions in global memory (the coloured non-parallel group): records of one colour
// never share a written vertex, so plain loads and stores suffice within a colour.
struct GlobalPositions
{
    float4 *x;
};
__device__ __forceinline__ float4 pget(const GlobalPositions &X, uint32_t i) { return X.x[i]; }
__device__ __forceinline__ void pset(const GlobalPositions &X, uint32_t i, float4 p) { X.x[i] = p; }

// One record of the coloured non-parallel group: a = the four body vertex indices,
// b = {type, p0, p1, 0} (the parameters of the packed 16 B record).
__device__ __forceinline__ void project_coloured(const GlobalPositions &X, uint4 a, uint4 b, float inv_dt_sq)
{
    const float p0 = __uint_as_float(b.y), p1 = __uint_as_float(b.z);
    switch (b.x)
    {
    case kTypeVolume: project_volume(X, a.x, a.y, a.z, a.w, p0, p1, inv_dt_sq); break;
    case kTypeBend: project_bend(X, a.x, a.y, a.z, a.w, p0, p1, inv_dt_sq); break;
    case kTypeEdge: project_edge(X, a.x, a.y, p0, p1, inv_dt_sq); break;
    case kTypeLRA: project_lra(X, a.x, a.y, p0); break;
    default: break;
    }
}

template <bool EXT>
__global__ __launch_bounds__(kTeamBlock) void team_kernel(
    const BodyDev *__restrict__ bodies, BodyStateDev *__restrict__ states, const TeamDev *__restrict__ teams,
    const uint2 *__restrict__ block_map, const ColliderDev *__restrict__ colliders, uint32_t num_colliders,
    uint32_t max_leaves, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames, int32_t *__restrict__ error_flag)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ BodyStateDev st;
    __shared__ float sh_delta[4];
    __shared__ uint32_t sh_num_leaves;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid >> 6;
    const uint2 bm = block_map[blockIdx.x];
    const TeamDev &TD = teams[bm.x];
    const uint32_t member = bm.y, T = TD.size;
    const uint32_t bi = TD.body;
    const BodyDev &B = bodies[bi];
    const SettingsDev &S = B.s;
    const uint32_t nv = S.num_vertices;
    const TeamLds lay = team_lds_layout(S.window_capacity, max_leaves, num_colliders);
    float4 *X = reinterpret_cast<float4 *>(lds_raw);
    Leaf *const lds_leaves = reinterpret_cast<Leaf *>(lds_raw + lay.leaves);
    Leaf *leaves = lds_leaves;
    int32_t *slots = reinterpret_cast<int32_t *>(lds_raw + lay.slots);
    float *red = reinterpret_cast<float *>(lds_raw + lay.red);
    float4 *P = B.x;
    // this member's slice of the vertices (whole wavefronts)
    const uint32_t chunk = ((nv + T - 1u) / T + (uint32_t)kWave - 1u) / (uint32_t)kWave * (uint32_t)kWave;
    const uint32_t vb = min(nv, member * chunk), ve = min(nv, vb + chunk);
    uint32_t epoch = 0;
    auto barrier = [&]() { team_barrier(TD.sync, T * ++epoch, error_flag); };

    if (tid == 0)
        st = states[bi];
    __syncthreads();

    const V3 gravity_world = v3(gx, gy, gz);
    const float step_dt = dt / (float)steps;        // PhysicsSystem.cpp:215
    const uint32_t iterations = B.iterations;

    bool report_cleared = false;
    for (int32_t frame = 0; frame < frames; ++frame)
        for (int32_t step = 0; step < steps; ++step)
        {
            if (!st.active)
            {
                // a body that does not run in a step reports no contacts for it
                if (!report_cleared)
                    for (uint32_t i = vb + tid; i < ve; i += kTeamBlock)
                        B.report[i] = contact_report(v3(0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f), false, -1);
                report_cleared = true;
                break;
            }

            // InitializeUpdateContext (SoftBodyMotionProperties.cpp:937-960)
            const M44 com = rotation_translation(st.rot[0], st.rot[1], st.rot[2], st.rot[3], v3(st.pos[0], st.pos[1], st.pos[2]));
            const V3 g_local = mul3x3_transposed(com, B.gravity_factor * gravity_world);
            const float dt_sub = step_dt / (float)iterations;
            const V3 disp_gravity = (((0.5f * (float)iterations) * (float)(iterations + 1)) * (dt_sub * dt_sub)) * g_local;
            const float inv_dt_sq = 1.0f / (dt_sub * dt_sub);

            // DetermineCollidingShapes (SoftBodyMotionProperties.cpp:236-253): every member
            // builds the same colliding list in its own LDS
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
                for (uint32_t ci = tid; ci < num_colliders; ci += kTeamBlock)
                {
                    const ColliderDev &cd = colliders[ci];
                    // AABox::Overlaps (Geometry/AABox.h:164-167)
                    bool overlap = !(wmin.x > cd.bounds[3] || wmin.y > cd.bounds[4] || wmin.z > cd.bounds[5]
                                     || wmax.x < cd.bounds[0] || wmax.y < cd.bounds[1] || wmax.z < cd.bounds[2]);
                    slots[ci] = overlap && !pair_skipped(pair_settings(B.pair, cd.owner), false, false)
                                && part_collected<EXT>(cd, inv_com, lmin, lmax) ? 1 : 0;
                }
                __syncthreads();
                if (tid == 0)
                {
                    uint32_t n = 0;
                    for (uint32_t ci = 0; ci < num_colliders; ++ci)
                        slots[ci] = slots[ci] ? (int32_t)n++ : -1;
                    sh_num_leaves = n;
                }
                __syncthreads();
                // the colliding list in LDS when it fits, else in this member's slice of the
                // team's global leaf table (a body among more colliders than LDS holds)
                leaves = sh_num_leaves <= max_leaves ? lds_leaves : TD.leaves + (size_t)member * num_colliders;
                for (uint32_t ci = tid; ci < num_colliders; ci += kTeamBlock)
                    if (slots[ci] >= 0)
                    {
                        const ColliderDev &cd = colliders[ci];
                        make_leaf<EXT>(cd, inv_com, B.friction, B.restitution, leaves[slots[ci]], nullptr, pair_settings(B.pair, cd.owner));
                        leaves[slots[ci]].collider = cd.owner;
                    }
                if (sh_num_leaves <= max_leaves)
                    __syncthreads();
                else
                    sync_global();
            }
            const uint32_t num_leaves = sh_num_leaves;

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

            // DetermineCollisionPlanes (SoftBodyMotionProperties.cpp:257-265), then the first
            // iteration's integration, over this member's vertices
            for (uint32_t i = vb + tid; i < ve; i += kTeamBlock)
            {
                float4 p = P[i];
                if (num_leaves > 0 && p.w > 0.0f)
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
                integrate_vertex(i, p, xyz(B.v[i]));
            }
            barrier();

            const float restitution_threshold = -2.0f * len(g_local) * dt_sub;
            const float max_lin_sq = B.max_linear_velocity * B.max_linear_velocity;
            float max_v_sq = 0.0f;
            V3 bmin = v3(FLT_MAX, FLT_MAX, FLT_MAX), bmax = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            V3 pmin = bmin, pmax = bmax;
            V3 lin = v3(0.0f, 0.0f, 0.0f), ang = v3(0.0f, 0.0f, 0.0f);

            for (uint32_t it = 0; it < iterations; ++it)
            {
                const bool last = it + 1 == iterations;
                const RecordCtx rc{ B.skin, B.skin_base, (float)(it + 1) / (float)iterations, nullptr, nullptr, nullptr };
                auto run_window = [&](uint32_t w) {
                    const WindowDev W = S.windows[w];
                    for (uint32_t j = tid; j < W.vcount; j += kTeamBlock)
                        X[j] = P[S.vlist[W.vbegin + j]];
                    __syncthreads();
                    run_levels_global<kTeamBlock>(S.recs, S.lvl + W.lbegin, W.nlev, X, inv_dt_sq, (uint32_t)tid, NoSkip(), rc);
                    for (uint32_t j = tid; j < W.vcount; j += kTeamBlock)
                        P[S.vlist[W.vbegin + j]] = X[j];
                    __syncthreads();
                };

                // 1. this member's parallel-group windows
                for (uint32_t w = TD.par_split[member]; w < TD.par_split[member + 1]; ++w)
                    run_window(w);
                barrier();

                // 2. the non-parallel group
                if (TD.num_colours == 0)
                {
                    if (member == 0)
                        for (uint32_t w = S.num_par_windows; w < S.num_par_windows + S.num_tail_windows; ++w)
                        {
                            run_window(w);
                            // the next window may gather what this one scattered
                            sync_global();
                        }
                    barrier();
                }
                else
                {
                    const GlobalPositions G{ P };
                    for (uint32_t c = 0; c < TD.num_colours; ++c)
                    {
                        const uint32_t lo = TD.colour_begin[c], hi = TD.colour_begin[c + 1];
                        for (uint32_t k = lo + member * kTeamBlock + tid; k < hi; k += T * kTeamBlock)
                            project_coloured(G, TD.colour_recs[2u * k], TD.colour_recs[2u * k + 1u], inv_dt_sq);
                        barrier();
                    }
                }

                // 3. ApplyCollisionConstraintsAndUpdateVelocities (SoftBodyMotionProperties.cpp:700-825),
                // fused with the next iteration's integration, or with UpdateSoftBodyState's
                // per-vertex work after the last one
                for (uint32_t i = vb + tid; i < ve; i += kTeamBlock)
                {
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
                                B.contact[i] = 1;
                                V3 n = plane.n;
                                x = x + n * projected_distance;
                                if (cs.inv_mass_scale1 > 0.0f)
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
                        integrate_vertex(i, f4(x, p.w), v);
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
                }
                if (!last)
                    barrier();
            }

            // UpdateSoftBodyState reductions (SoftBodyMotionProperties.cpp:846-921): each member
            // reduces its vertices, member 0 combines the members in order
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
                    for (int w = 1; w < kTeamWaves; ++w)
                    {
#pragma unroll
                        for (int k = 0; k < 6; ++k) vals[k] = smin(vals[k], red[w * 32 + k]);
#pragma unroll
                        for (int k = 6; k < 13; ++k) vals[k] = smax(vals[k], red[w * 32 + k]);
#pragma unroll
                        for (int k = 0; k < 6; ++k) sums[k] += red[w * 32 + 16 + k];
                    }
                    float *out = TD.red + 32u * member;
#pragma unroll
                    for (int k = 0; k < 13; ++k) out[k] = vals[k];
#pragma unroll
                    for (int k = 0; k < 6; ++k) out[16 + k] = sums[k];
                }
                barrier();
                if (member == 0 && tid == 0)
                {
                    for (uint32_t m = 1; m < T; ++m)
                    {
                        const float *in = TD.red + 32u * m;
#pragma unroll
                        for (int k = 0; k < 6; ++k) vals[k] = smin(vals[k], in[k]);
#pragma unroll
                        for (int k = 6; k < 13; ++k) vals[k] = smax(vals[k], in[k]);
#pragma unroll
                        for (int k = 0; k < 6; ++k) sums[k] += in[16 + k];
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
                    states[bi] = st;
                    float *d = TD.red + 32u * T;
                    d[0] = delta.x; d[1] = delta.y; d[2] = delta.z;
                }
                barrier();
                // every member takes the step's result from member 0
                if (tid == 0)
                {
                    st = states[bi];
                    const float *d = TD.red + 32u * T;
                    sh_delta[0] = d[0]; sh_delta[1] = d[1]; sh_delta[2] = d[2];
                }
                __syncthreads();
                const V3 delta = v3(sh_delta[0], sh_delta[1], sh_delta[2]);
                if (B.update_position)
                    for (uint32_t i = vb + tid; i < ve; i += kTeamBlock)
                    {
                        float4 p = P[i];
                        P[i] = f4(xyz(p) - delta, p.w);
                    }
                if (B.skin)
                {
                    // skin states (SoftBodyMotionProperties.cpp:890-917): the skinned positions
                    // follow the recentring; previous positions are invalid until the next SkinVertices
                    float4 *e = const_cast<float4 *>(B.skin) + 4u * (size_t)B.skin_base;
                    for (uint32_t i = vb + tid; i < ve; i += kTeamBlock)
                    {
                        if (B.update_position)
                        {
                            const float4 sp = e[4u * i + 1u];
                            e[4u * i + 1u] = f4(xyz(sp) - delta, sp.w);
                        }
                        e[4u * i + 3u].w = 0.0f;
                    }
                }
                // the next step's first phase touches only this member's vertices; the
                // member-0 reductions slot is rewritten only after the next step's barriers
                __syncthreads();
            }
        }
}

uint32_t team_lds_bytes(uint32_t window_capacity, uint32_t max_leaves, uint32_t nc)
{
    return team_lds_layout(window_capacity, max_leaves, nc).total;
}

template <bool EXT>
static hipError_t launch_team_variant(const BodyDev *bodies, BodyStateDev *states, const TeamDev *teams, const uint2 *block_map,
                                      uint32_t num_blocks, const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves,
                                      uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                                      int32_t *error_flag, hipStream_t stream)
{
    static const hipError_t configured =
        hipFuncSetAttribute((const void *)team_kernel<EXT>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (configured != hipSuccess)
        return configured;
    if (num_blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL((team_kernel<EXT>), dim3(num_blocks), dim3(kTeamBlock), lds_bytes, stream, bodies, states, teams, block_map,
                       colliders, num_colliders, max_leaves, gx, gy, gz, dt, steps, frames, error_flag);
    return hipGetLastError();
}

// ext: the collider list holds triangle trees or scaled convex hulls
hipError_t launch_team(bool ext, const BodyDev *bodies, BodyStateDev *states, const TeamDev *teams, const uint2 *block_map,
                       uint32_t num_blocks, const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves,
                       uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                       int32_t *error_flag, hipStream_t stream)
{
    if (ext)
        return launch_team_variant<true>(bodies, states, teams, block_map, num_blocks, colliders, num_colliders, max_leaves, lds_bytes,
                                         gx, gy, gz, dt, steps, frames, error_flag, stream);
    return launch_team_variant<false>(bodies, states, teams, block_map, num_blocks, colliders, num_colliders, max_leaves, lds_bytes,
                                      gx, gy, gz, dt, steps, frames, error_flag, stream);
}

} // namespace xpbd
