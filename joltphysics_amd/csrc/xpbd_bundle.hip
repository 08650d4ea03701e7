// Bundle kernel: several LDS-resident soft bodies per workgroup.
//
// One 256-thread workgroup owns a bundle of bodies for a whole xpbd_update call
// (all frames x collision steps x iterations).  All positions + inverse masses
// of the bundle live in LDS as float4.  Each body's constraint program -- the
// merged levels of its parallel update groups followed by the levels of its
// non-parallel group (SoftBodyMotionProperties.cpp:1030-1115) -- is merged with
// the other bodies' programs by level index, so the lanes of a level are filled
// by constraints of several bodies.  Bodies are independent (they share no
// vertices), so merging keeps every body's own sequential order and the results
// stay bit-identical to the reference.  Within a level the host groups records
// by constraint type and pads the groups to whole wavefronts, so a wavefront
// runs one constraint path instead of the union of all of them.
//
// Per-vertex phases look up their body through a per-vertex body slot in LDS
// and use that body's per-step constants (gravity, sub-step, damping, ...).

#include "xpbd_device.h"

namespace xpbd {


// Per-body constants of the bundle (LDS), refreshed every collision step.
struct BodyLocal
{
    // creation settings
    float linear_damping, max_linear_velocity, restitution, friction, pressure, gravity_factor, vertex_radius;
    int32_t update_position, allow_sleeping;
    // per step
    float gx, gy, gz;           // gravity in body space
    float dt_sub, inv_dt_sq, damping;
    float sgx, sgy, sgz;        // gravity * dt_sub
    float dgx, dgy, dgz;        // displacement due to gravity
    float restitution_threshold, max_lin_sq;
    float delta[3];             // recentring of this step
    float six_volume;           // pressure
    uint32_t num_leaves;
    uint32_t num_dyn_leaves;    // leaves of dynamic colliders this step
    uint32_t num_dynamic;       // dynamic colliders of the context
    uint32_t report_cleared;    // asleep: its contact report was cleared
    const float4 *skin_entries; // the body's skin entries (4 float4 per vertex), or null
    int32_t *sensor_hits;       // the body's sensor contact row, or null
    uint32_t num_sensors;
    uint32_t ran;               // the body was active at the start of this step
    DynStateDev *dyn;           // the body's row of dynamic-collider states
    const float4 *pair;         // the body's per-collider contact settings, or null
    uint32_t nf;
    const uint32_t *faces;
    const uint32_t *vf_begin;
    const uint32_t *vf_list;
};

struct BundleLds
{
    uint32_t recs, lvl, vb, locals, states, total;
};

// pos_bytes: 16 (float4 position + inverse mass) or 12 (float3; one inverse mass per bundle)
__host__ __device__ inline BundleLds bundle_layout(uint32_t nv, uint32_t nrec, uint32_t nlvl, uint32_t nb, uint32_t pos_bytes)
{
    BundleLds l;
    uint32_t off = (nv * pos_bytes + 15u) & ~15u;
    l.recs = off; off += nrec * 16u;
    l.lvl = off; off += nlvl * 4u;
    off = (off + 15u) & ~15u; l.vb = off; off += nv;
    off = (off + 15u) & ~15u; l.locals = off; off += nb * (uint32_t)sizeof(BodyLocal);
    off = (off + 15u) & ~15u; l.states = off; off += nb * (uint32_t)sizeof(BodyStateDev);
    l.total = (off + 15u) & ~15u;
    return l;
}

// Position container of a bundle: float4 per vertex, or float3 per vertex plus the
// inverse mass all vertices of the bundle share.
template <bool UNIFORM> struct BundlePositions;
template <> struct BundlePositions<false>
{
    using Type = float4 *;
    static __device__ Type make(unsigned char *lds, float) { return reinterpret_cast<float4 *>(lds); }
};
template <> struct BundlePositions<true>
{
    using Type = PositionsF3;
    static __device__ Type make(unsigned char *lds, float w)
    {
        PositionsF3 p;
        p.x = reinterpret_cast<float *>(lds);
        p.w = w;
        return p;
    }
};

template <bool STAGED, int BLOCK, bool UNIFORM, bool EXT>
__device__ __forceinline__ void bundle_update(
    const BundleDev &G, const BodyDev *__restrict__ bodies, BodyStateDev *__restrict__ states,
    const ColliderDev *__restrict__ colliders, uint32_t num_colliders,
    float gx, float gy, float gz, float dt, int32_t steps, int32_t frames)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const unsigned long long t_begin = __builtin_amdgcn_s_memrealtime();
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid >> 6;
    const uint32_t nv = G.nv;
    const uint32_t nb = G.num_bodies;
    const uint32_t nlev = G.nlev;
    const bool has_sleepers = G.has_sleepers != 0;
    const BundleLds lay = bundle_layout(nv, STAGED ? G.num_records : 0u, STAGED ? G.nlev + 1 : 0u, nb, UNIFORM ? 12u : 16u);
    const typename BundlePositions<UNIFORM>::Type X = BundlePositions<UNIFORM>::make(lds_raw, G.uniform_inv_mass);
    const uint4 *recs = STAGED ? reinterpret_cast<const uint4 *>(lds_raw + lay.recs) : G.recs;
    const uint32_t *lvl = STAGED ? reinterpret_cast<const uint32_t *>(lds_raw + lay.lvl) : G.lvl;
    uint8_t *vb = reinterpret_cast<uint8_t *>(lds_raw + lay.vb);
    BodyLocal *loc = reinterpret_cast<BodyLocal *>(lds_raw + lay.locals);
    BodyStateDev *st = reinterpret_cast<BodyStateDev *>(lds_raw + lay.states);
    Leaf *leaves_full = reinterpret_cast<Leaf *>(G.leaves_full);
    Leaf *leaves = reinterpret_cast<Leaf *>(G.leaves);

    // ---- launch prologue
    if (STAGED)
    {
        uint4 *rd = reinterpret_cast<uint4 *>(lds_raw + lay.recs);
        for (uint32_t i = tid; i < G.num_records; i += BLOCK)
            rd[i] = G.recs[i];
        uint32_t *ld = reinterpret_cast<uint32_t *>(lds_raw + lay.lvl);
        for (uint32_t i = tid; i <= G.nlev; i += BLOCK)
            ld[i] = G.lvl[i];
    }
    for (uint32_t i = tid; i < nv; i += BLOCK)
        pset(X, i, G.x[i]);
    for (uint32_t b = 0; b < nb; ++b)
    {
        const BundleBodyDev bb = G.bodies[b];
        for (uint32_t i = tid; i < bb.nv; i += BLOCK)
            vb[bb.voff + i] = (uint8_t)b;
    }
    for (uint32_t b = tid; b < nb; b += BLOCK)
    {
        const BodyDev &B = bodies[G.bodies[b].body];
        BodyLocal &L = loc[b];
        L.linear_damping = B.linear_damping; L.max_linear_velocity = B.max_linear_velocity;
        L.restitution = B.restitution; L.friction = B.friction; L.pressure = B.pressure;
        L.gravity_factor = B.gravity_factor; L.vertex_radius = B.vertex_radius;
        L.update_position = B.update_position; L.allow_sleeping = B.allow_sleeping;
        L.nf = B.pressure > 0.0f ? B.s.num_faces : 0u;
        L.faces = B.s.faces; L.vf_begin = B.s.vf_begin; L.vf_list = B.s.vf_list;
        L.dyn = B.dyn;
        L.pair = B.pair;
        L.report_cleared = 0;
        L.skin_entries = B.skin ? B.skin + 4u * (size_t)B.skin_base : nullptr;
        L.sensor_hits = B.sensor_hits;
        L.num_sensors = B.sensor_hits ? B.num_sensors : 0u;
        L.num_dynamic = B.dyn ? B.num_dynamic : 0u;
        for (uint32_t k = 0; k < L.num_dynamic; ++k)
            clear_dynamic(L.dyn[k]);
        st[b] = states[G.bodies[b].body];
    }
    __syncthreads();
    // dynamic colliders anywhere in the bundle (workgroup-uniform)
    bool any_dynamic = false;
    for (uint32_t b = 0; b < nb; ++b)
        any_dynamic |= loc[b].num_dynamic > 0;

    const V3 gravity_world = v3(gx, gy, gz);
    const float step_dt = dt / (float)steps;                        // PhysicsSystem.cpp:215
    const uint32_t iterations = G.iterations;
    const uint32_t nc = num_colliders;
    // every body of a bundle has the same dt, steps and iterations, so the same sub-step
    const float bundle_dt_sub = step_dt / (float)iterations;
    const float inv_dt_sq = 1.0f / (bundle_dt_sub * bundle_dt_sub);

    auto body_of = [&](uint32_t i) -> uint32_t { return vb[i]; };

    for (int32_t frame = 0; frame < frames; ++frame)
        for (int32_t step = 0; step < steps; ++step)
        {
            // ---- a body that does not run in a step reports no contacts for it
            if (has_sleepers)
                for (uint32_t b = wave; b < nb; b += (BLOCK / kWave))
                    if (!st[b].active && !loc[b].report_cleared)
                    {
                        const BundleBodyDev bb = G.bodies[b];
                        for (uint32_t k = lane; k < bb.nv; k += kWave)
                            G.report[bb.voff + k] = contact_report(v3(0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f), false, -1);
                        if (lane == 0)
                            loc[b].report_cleared = 1;
                    }
            // ---- InitializeUpdateContext per body (SoftBodyMotionProperties.cpp:937-960)
            for (uint32_t b = tid; b < nb; b += BLOCK)
            {
                BodyLocal &L = loc[b];
                const BodyStateDev &s = st[b];
                L.ran = s.active ? 1u : 0u;
                M44 com = rotation_translation(s.rot[0], s.rot[1], s.rot[2], s.rot[3], v3(s.pos[0], s.pos[1], s.pos[2]));
                V3 g_local = mul3x3_transposed(com, L.gravity_factor * gravity_world);
                float dt_sub = step_dt / (float)iterations;
                V3 disp = (((0.5f * (float)iterations) * (float)(iterations + 1)) * (dt_sub * dt_sub)) * g_local;
                L.gx = g_local.x; L.gy = g_local.y; L.gz = g_local.z;
                L.dt_sub = dt_sub;
                L.inv_dt_sq = 1.0f / (dt_sub * dt_sub);
                L.damping = smax(0.0f, 1.0f - L.linear_damping * dt_sub);
                V3 sg = g_local * dt_sub;
                L.sgx = sg.x; L.sgy = sg.y; L.sgz = sg.z;
                L.dgx = disp.x; L.dgy = disp.y; L.dgz = disp.z;
                L.restitution_threshold = -2.0f * len(g_local) * dt_sub;
                L.max_lin_sq = L.max_linear_velocity * L.max_linear_velocity;
            }
            // ---- DetermineCollidingShapes: candidate leaves per (body, collider)
            for (uint32_t p = tid; p < nb * nc; p += BLOCK)
            {
                const uint32_t b = p / nc, c = p - b * nc;
                const BodyStateDev &s = st[b];
                Leaf &out = leaves_full[p];
                if (!s.active)
                {
                    out.type = -1;
                    continue;
                }
                const BodyLocal &L = loc[b];
                M44 com = rotation_translation(s.rot[0], s.rot[1], s.rot[2], s.rot[3], v3(s.pos[0], s.pos[1], s.pos[2]));
                V3 lmin = vmin(v3(s.bmin[0], s.bmin[1], s.bmin[2]), v3(s.pmin[0], s.pmin[1], s.pmin[2]));
                V3 lmax = vmax(v3(s.bmax[0], s.bmax[1], s.bmax[2]), v3(s.pmax[0], s.pmax[1], s.pmax[2]));
                V3 r = v3(L.vertex_radius, L.vertex_radius, L.vertex_radius);
                lmin = lmin - r;
                lmax = lmax + r;
                V3 wmin = col3(com, 3), wmax = col3(com, 3);               // AABox::Transformed
                for (int k = 0; k < 3; ++k)
                {
                    V3 col = col3(com, k);
                    V3 a = col * comp(lmin, k);
                    V3 bq = col * comp(lmax, k);
                    wmin = wmin + vmin(a, bq);
                    wmax = wmax + vmax(a, bq);
                }
                const ColliderDev &cd = colliders[c];
                const float4 pr = pair_settings(L.pair, cd.owner);
                bool overlap = !(wmin.x > cd.bounds[3] || wmin.y > cd.bounds[4] || wmin.z > cd.bounds[5]
                                 || wmax.x < cd.bounds[0] || wmax.y < cd.bounds[1] || wmax.z < cd.bounds[2]);
                const M44 inv_com = inversed_rotation_translation(com);
                if (!overlap || pair_skipped(pr, cd.sensor >= 0, cd.motion == 2) || !part_collected<EXT>(cd, inv_com, lmin, lmax))
                {
                    out.type = -1;
                    continue;
                }
                make_leaf<EXT>(cd, inv_com, L.friction, L.restitution, out, L.dyn, pr);
                out.collider = cd.owner;
            }
            sync_global();
            // colliding list of each body in collider (broadphase) order
            for (uint32_t b = tid; b < nb; b += BLOCK)
            {
                uint32_t n = 0, nd = 0;
                for (uint32_t c = 0; c < nc; ++c)
                    if (leaves_full[b * nc + c].type >= 0 && colliders[c].sensor < 0)
                    {
                        nd += leaves_full[b * nc + c].dyn >= 0 ? 1u : 0u;
                        leaves[b * nc + n++] = leaves_full[b * nc + c];
                    }
                loc[b].num_leaves = n;
                loc[b].num_dyn_leaves = nd;
            }
            sync_global();

            // ---- DetermineSensorCollisions (SoftBodyMotionProperties.cpp:267-289), one wavefront
            // per body: a sensor reports contact when a vertex penetrates it; sensors never collide
            for (uint32_t b = wave; b < nb; b += (BLOCK / kWave))
            {
                const BodyLocal &L = loc[b];
                if (L.num_sensors == 0)
                    continue;
                const BundleBodyDev bb = G.bodies[b];
                for (uint32_t c = 0; c < nc; ++c)
                {
                    const int32_t rank = colliders[c].sensor;
                    if (rank < 0)
                        continue;
                    const Leaf &leaf = leaves_full[b * nc + c];
                    float largest = -FLT_MAX;
                    if (st[b].active && leaf.type >= 0)
                        for (uint32_t k = lane; k < bb.nv; k += kWave)
                            largest = smax(largest, sensor_penetration<EXT>(leaf, pget(X, bb.voff + k)));
                    largest = wave_max(largest);
                    // a compound sensor is hit when one of its leaves is
                    if (lane == 0)
                        L.sensor_hits[rank] = largest > 0.0f || (!colliders[c].first && L.sensor_hits[rank]) ? 1 : 0;
                }
            }

            // ---- DetermineCollisionPlanes (SoftBodyMotionProperties.cpp:257-265)
            for (uint32_t i = tid; i < nv; i += BLOCK)
            {
                const uint32_t b = body_of(i);
                const uint32_t n = loc[b].num_leaves;
                float4 p = pget(X, i);
                if (n > 0 && st[b].active && p.w > 0.0f)
                {
                    Pl plane; plane.n = v3(0.0f, 0.0f, 0.0f); plane.c = 0.0f;
                    float largest = -FLT_MAX;
                    int32_t idx = -1;
                    collide_vertex<EXT>(leaves + b * nc, n, xyz(p), largest, idx, plane);
                    if (idx >= 0)
                    {
                        G.idx[i] = idx;
                        G.plane[i] = make_float4(plane.n.x, plane.n.y, plane.n.z, plane.c);
                    }
                }
            }

            // ---- ApplyPressure (SoftBodyMotionProperties.cpp:291-323)
            auto apply_pressure = [&]() {
                // GetVolumeTimesSix: a sequential sum in face order, one wavefront per
                // body.  The lanes compute the terms of 64 faces at a time and the sum
                // takes them in lane order, so the additions happen in face order.
                for (uint32_t b = wave; b < nb; b += (BLOCK / kWave))
                {
                    const BodyLocal &L = loc[b];
                    if (L.nf == 0 || !st[b].active)
                        continue;
                    const uint32_t voff = G.bodies[b].voff;
                    float six = 0.0f;
                    for (uint32_t f0 = 0; f0 < L.nf; f0 += kWave)
                    {
                        const uint32_t f = f0 + (uint32_t)lane;
                        float term = 0.0f;
                        if (f < L.nf)
                        {
                            V3 x1 = xyz(pget(X, voff + L.faces[3 * f])), x2 = xyz(pget(X, voff + L.faces[3 * f + 1])), x3 = xyz(pget(X, voff + L.faces[3 * f + 2]));
                            term = dot(cross(x1, x2), x3);
                        }
                        const uint32_t count = L.nf - f0 < (uint32_t)kWave ? L.nf - f0 : (uint32_t)kWave;
                        for (uint32_t k = 0; k < count; ++k)
                            six += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(term), k));
                    }
                    if (lane == 0)
                        loc[b].six_volume = six;
                }
                __syncthreads();
                for (uint32_t b = 0; b < nb; ++b)
                {
                    const BodyLocal &L = loc[b];
                    if (L.nf == 0 || !st[b].active || !(L.six_volume > 0.0f))
                        continue;
                    const BundleBodyDev bb = G.bodies[b];
                    const float coefficient = L.pressure * L.dt_sub / L.six_volume;
                    for (uint32_t f = tid; f < L.nf; f += BLOCK)
                    {
                        V3 x1 = xyz(pget(X, bb.voff + L.faces[3 * f])), x2 = xyz(pget(X, bb.voff + L.faces[3 * f + 1])), x3 = xyz(pget(X, bb.voff + L.faces[3 * f + 2]));
                        V3 impulse = coefficient * cross(x2 - x1, x3 - x1);
                        G.face_scratch[bb.face_off + f] = make_float4(impulse.x, impulse.y, impulse.z, 0.0f);
                    }
                }
                sync_global();
                for (uint32_t i = tid; i < nv; i += BLOCK)
                {
                    const uint32_t b = body_of(i);
                    const BodyLocal &L = loc[b];
                    if (L.nf == 0 || !st[b].active || !(L.six_volume > 0.0f))
                        continue;
                    const BundleBodyDev bb = G.bodies[b];
                    const uint32_t li = i - bb.voff;
                    uint32_t fb = L.vf_begin[li], fe = L.vf_begin[li + 1];
                    if (fb == fe)
                        continue;
                    float w = pget(X, i).w;
                    V3 v = xyz(G.v[i]);
                    for (uint32_t k = fb; k < fe; ++k)
                        v = v + w * xyz(G.face_scratch[bb.face_off + L.vf_list[k]]);
                    G.v[i] = f4(v, 0.0f);
                }
                __syncthreads();
            };

            // ---- IntegratePositions (SoftBodyMotionProperties.cpp:325-350)
            auto integrate_vertex = [&](uint32_t i, const BodyLocal &L, float4 p, V3 v) {
                if (p.w > 0.0f)
                {
                    const V3 sub_step_impulse = (v3(0.0f, 0.0f, 0.0f) * L.dt_sub) / smax((float)G.bodies[vb[i]].nv, 1.0f);
                    v = v + (v3(L.sgx, L.sgy, L.sgz) + sub_step_impulse * p.w);
                    v = v * L.damping;
                }
                V3 position = xyz(p);
                G.prev[i] = f4(position, 0.0f);
                pset(X, i, f4(position + v * L.dt_sub, p.w));
                G.v[i] = f4(v, 0.0f);
            };

            if (G.has_pressure)
                apply_pressure();
            for (uint32_t i = tid; i < nv; i += BLOCK)
            {
                const uint32_t b = body_of(i);
                if (st[b].active)
                    integrate_vertex(i, loc[b], pget(X, i), xyz(G.v[i]));
            }
            __syncthreads();

            for (uint32_t it = 0; it < iterations; ++it)
            {
                const bool last = it + 1 == iterations;

                // ---- merged level program of all bodies (skin records carry their entries)
                const RecordCtx sk{ G.skin, 0u, (float)(it + 1) / (float)iterations };
                if (has_sleepers)
                {
                    auto skip = [&](const uint4 &r) { return r.x == kNopRecord || !st[vb[r.x & 0x3fffu]].active; };
                    if (STAGED) run_levels_lds<BLOCK>(recs, lvl, nlev, X, inv_dt_sq, tid, skip, sk);
                    else run_levels_global<BLOCK>(recs, lvl, nlev, X, inv_dt_sq, tid, skip, sk);
                }
                else
                {
                    if (STAGED) run_levels_lds<BLOCK>(recs, lvl, nlev, X, inv_dt_sq, tid, NoSkip(), sk);
                    else run_levels_global<BLOCK>(recs, lvl, nlev, X, inv_dt_sq, tid, NoSkip(), sk);
                }

                // ---- collision + velocity (SoftBodyMotionProperties.cpp:700-825), fused with the
                // next integration (no pressure) or the per-vertex part of UpdateSoftBodyState.
                // A vertex in contact with a dynamic collider changes that body's velocities,
                // which the next such vertex reads: the parallel pass marks these vertices
                // (contact 2) and leaves them; one lane per body then runs them in vertex order.
                auto collide_update = [&](uint32_t i, bool in_order) {
                    const uint32_t b = body_of(i);
                    const BodyLocal &L = loc[b];
                    float4 p = pget(X, i);
                    V3 x = xyz(p);
                    V3 v = xyz(G.v[i]);
                    if (p.w > 0.0f)
                    {
                        V3 prev_v = v;
                        v = (x - xyz(G.prev[i])) / L.dt_sub;
                        int32_t idx = G.idx[i];
                        if (idx >= 0)
                        {
                            float4 pl = G.plane[i];
                            Pl plane; plane.n = v3(pl.x, pl.y, pl.z); plane.c = pl.w;
                            float projected_distance = -signed_distance(plane, x) + L.vertex_radius;
                            if (projected_distance > 0.0f)
                            {
                                const Leaf &cs = leaves[b * nc + idx];
                                if (cs.dyn >= 0 && !in_order)
                                {
                                    G.contact[i] = 2;
                                    return;
                                }
                                G.contact[i] = 1;
                                V3 n = plane.n;
                                x = x + n * projected_distance;
                                if (cs.dyn >= 0)
                                    dynamic_contact(L.dyn[cs.dyn], ld3(L.dyn[cs.dyn].com), x, v, prev_v, n, projected_distance, p.w,
                                                    cs.friction, cs.restitution, L.dt_sub, L.restitution_threshold,
                                                    cs.inv_mass_scale1);
                                else if (cs.inv_mass_scale1 > 0.0f)
                                {
                                    V3 v_normal = n * dot(n, v);
                                    V3 v_tangential = v - v_normal;
                                    float vtl = len(v_tangential);
                                    if (vtl > 0.0f)
                                    {
                                        float f = cs.friction * projected_distance / (vtl * L.dt_sub);
                                        v = v - v_tangential * (f < 1.0f ? f : 1.0f);
                                    }
                                    v = v - v_normal;
                                    float prev_v_normal = dot(prev_v, n);
                                    if (prev_v_normal < L.restitution_threshold)
                                        v = v - (cs.restitution * prev_v_normal) * n;
                                }
                            }
                        }
                    }
                    if (!last)
                    {
                        if (L.nf > 0)
                        {
                            pset(X, i, f4(x, p.w));
                            G.v[i] = f4(v, 0.0f);
                        }
                        else
                            integrate_vertex(i, L, f4(x, p.w), v);
                    }
                    else
                    {
                        // UpdateSoftBodyState per-vertex part (SoftBodyMotionProperties.cpp:852-880)
                        float v_sq = len_sq(v);
                        if (v_sq > L.max_lin_sq)
                            v = v * sqrtf(L.max_lin_sq / v_sq);
                        pset(X, i, f4(x, p.w));
                        G.v[i] = f4(v, 0.0f);
                        G.prev[i].w = v_sq;        // pre-clamp |v|^2 for the sleep test
                        const int32_t hit = G.idx[i];
                        G.report[i] = contact_report(x, G.plane[i], G.contact[i] != 0 && hit >= 0,
                                                     hit >= 0 ? leaves[b * nc + hit].collider : -1);
                        G.idx[i] = -1;
                        G.contact[i] = 0;
                    }
                };
                for (uint32_t i = tid; i < nv; i += BLOCK)
                    if (st[body_of(i)].active)
                        collide_update(i, false);
                if (any_dynamic)
                {
                    // the marks and the dynamic states were written by other wavefronts
                    sync_global();
                    for (uint32_t b = wave; b < nb; b += (BLOCK / kWave))
                    {
                        if (loc[b].num_dyn_leaves == 0 || !st[b].active)
                            continue;
                        const BundleBodyDev bb = G.bodies[b];
                        for (uint32_t k0 = 0; k0 < bb.nv; k0 += kWave)
                        {
                            const uint32_t k = k0 + (uint32_t)lane;
                            unsigned long long marked = __ballot(k < bb.nv && G.contact[bb.voff + k] == 2);
                            while (marked)
                            {
                                const uint32_t j = (uint32_t)__builtin_ctzll(marked);
                                marked &= marked - 1ull;
                                if (lane == 0)
                                    collide_update(bb.voff + k0 + j, true);
                            }
                        }
                    }
                    sync_global();
                }
                __syncthreads();
                if (!last && G.has_pressure)
                {
                    apply_pressure();
                    for (uint32_t i = tid; i < nv; i += BLOCK)
                    {
                        const uint32_t b = body_of(i);
                        const BodyLocal &L = loc[b];
                        if (st[b].active && L.nf > 0)
                            integrate_vertex(i, L, pget(X, i), xyz(G.v[i]));
                    }
                    __syncthreads();
                }
            }

            // the reductions read velocities other wavefronts stored
            sync_global();
            // ---- UpdateSoftBodyState reductions, one wavefront per body (SoftBodyMotionProperties.cpp:846-921)
            for (uint32_t b = wave; b < nb; b += (BLOCK / kWave))
            {
                BodyStateDev &s = st[b];
                if (!s.active)
                    continue;
                const BodyLocal &L = loc[b];
                const BundleBodyDev bb = G.bodies[b];
                const V3 disp = v3(L.dgx, L.dgy, L.dgz);
                float vals[13] = { FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX,
                                   -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, 0.0f };
                float sums[6] = { 0, 0, 0, 0, 0, 0 };
                for (uint32_t k = lane; k < bb.nv; k += kWave)
                {
                    const uint32_t i = bb.voff + k;
                    V3 x = xyz(pget(X, i));
                    V3 v = xyz(G.v[i]);
                    vals[12] = smax(vals[12], G.prev[i].w);
                    V3 pp = (x + v * step_dt) + disp;
                    vals[0] = smin(vals[0], x.x); vals[1] = smin(vals[1], x.y); vals[2] = smin(vals[2], x.z);
                    vals[3] = smin(vals[3], pp.x); vals[4] = smin(vals[4], pp.y); vals[5] = smin(vals[5], pp.z);
                    vals[6] = smax(vals[6], x.x); vals[7] = smax(vals[7], x.y); vals[8] = smax(vals[8], x.z);
                    vals[9] = smax(vals[9], pp.x); vals[10] = smax(vals[10], pp.y); vals[11] = smax(vals[11], pp.z);
                    V3 av = cross(x, v);
                    sums[0] += v.x; sums[1] += v.y; sums[2] += v.z;
                    sums[3] += av.x; sums[4] += av.y; sums[5] += av.z;
                }
#pragma unroll
                for (int k = 0; k < 6; ++k) vals[k] = wave_min(vals[k]);
#pragma unroll
                for (int k = 6; k < 13; ++k) vals[k] = wave_max(vals[k]);
#pragma unroll
                for (int k = 0; k < 6; ++k) sums[k] = wave_sum(sums[k]);
                if (lane == 0)
                {
                    M44 com = rotation_translation(s.rot[0], s.rot[1], s.rot[2], s.rot[3], v3(s.pos[0], s.pos[1], s.pos[2]));
                    for (uint32_t k = 0; k < L.num_dynamic; ++k)
                        finish_dynamic(L.dyn[k], com);
                    V3 bmn = v3(vals[0], vals[1], vals[2]), pmn = v3(vals[3], vals[4], vals[5]);
                    V3 bmx = v3(vals[6], vals[7], vals[8]), pmx = v3(vals[9], vals[10], vals[11]);
                    float divider = (float)(bb.nv > 1 ? bb.nv : 1u);
                    // body velocity (sums in tree order, the reference sums sequentially)
                    V3 lv = mul3x3(com, v3(sums[0], sums[1], sums[2]) / divider);
                    V3 av = mul3x3(com, v3(sums[3], sums[4], sums[5]) / divider);
                    s.lin[0] = lv.x; s.lin[1] = lv.y; s.lin[2] = lv.z;
                    s.ang[0] = av.x; s.ang[1] = av.y; s.ang[2] = av.z;
                    V3 delta = v3(0.0f, 0.0f, 0.0f);
                    if (L.update_position)
                    {
                        delta = 0.5f * (bmn + bmx);                   // AABox::GetCenter
                        V3 dw = mul3x3(com, delta);
                        V3 nd = neg(delta);
                        bmn = bmn + nd; bmx = bmx + nd; pmn = pmn + nd; pmx = pmx + nd;
                        s.pos[0] = s.pos[0] + dw.x; s.pos[1] = s.pos[1] + dw.y; s.pos[2] = s.pos[2] + dw.z;   // PhysicsSystem.cpp:2864
                    }
                    s.bmin[0] = bmn.x; s.bmin[1] = bmn.y; s.bmin[2] = bmn.z;
                    s.bmax[0] = bmx.x; s.bmax[1] = bmx.y; s.bmax[2] = bmx.z;
                    s.pmin[0] = pmn.x; s.pmin[1] = pmn.y; s.pmin[2] = pmn.z;
                    s.pmax[0] = pmx.x; s.pmax[1] = pmx.y; s.pmax[2] = pmx.z;
                    loc[b].delta[0] = delta.x; loc[b].delta[1] = delta.y; loc[b].delta[2] = delta.z;
                    if (L.allow_sleeping)
                    {
                        if (vals[12] > 0.03f)
                            s.sleep_timer = 0.0f;
                        else
                        {
                            s.sleep_timer += step_dt;
                            if (s.sleep_timer >= 0.5f)
                                s.active = 0;
                        }
                    }
                }
            }
            __syncthreads();
            for (uint32_t i = tid; i < nv; i += BLOCK)
            {
                const uint32_t b = body_of(i);
                const BodyLocal &L = loc[b];
                // recentre (a body that fell asleep this step still recentres: the reference
                // deactivates it after UpdateSoftBodyState)
                if (L.update_position && G.bodies[b].nv > 0)
                {
                    V3 delta = v3(L.delta[0], L.delta[1], L.delta[2]);
                    float4 p = pget(X, i);
                    pset(X, i, f4(xyz(p) - delta, p.w));
                }
                // skin states (SoftBodyMotionProperties.cpp:890-917): the skinned positions
                // follow the recentring; the previous positions stay valid only until the
                // next SkinVertices
                if (L.skin_entries && L.ran)
                {
                    float4 *e = const_cast<float4 *>(L.skin_entries) + 4u * (size_t)(i - G.bodies[b].voff);
                    if (L.update_position)
                    {
                        const float4 sp = e[1];
                        e[1] = f4(xyz(sp) - v3(L.delta[0], L.delta[1], L.delta[2]), sp.w);
                    }
                    e[3].w = 0.0f;
                }
            }
            // every wavefront has read the deltas before they are cleared for the next step
            __syncthreads();
            for (uint32_t b = tid; b < nb; b += BLOCK)
            {
                loc[b].delta[0] = 0.0f; loc[b].delta[1] = 0.0f; loc[b].delta[2] = 0.0f;
            }
            __syncthreads();
        }

    for (uint32_t i = tid; i < nv; i += BLOCK)
        G.x[i] = pget(X, i);
    for (uint32_t b = tid; b < nb; b += BLOCK)
        states[G.bodies[b].body] = st[b];
    // duration of the whole launch for this bundle: the host orders the bundles of
    // later updates by it, longest first
    if (tid == 0 && G.elapsed)
        *G.elapsed = __builtin_amdgcn_s_memrealtime() - t_begin;
}

// One workgroup per bundle.  STAGED: the records and level offsets are copied
// into LDS at launch (small programs only, 256-thread launches); UNIFORM: three
// floats per vertex in LDS and the one inverse mass of the bundle.
template <bool STAGED, int BLOCK, bool UNIFORM, bool EXT>
__global__ __launch_bounds__(BLOCK) void bundle_kernel(
    const BundleDev *__restrict__ bundles, const BodyDev *__restrict__ bodies, BodyStateDev *__restrict__ states,
    const ColliderDev *__restrict__ colliders, uint32_t num_colliders,
    float gx, float gy, float gz, float dt, int32_t steps, int32_t frames)
{
    bundle_update<STAGED, BLOCK, UNIFORM, EXT>(bundles[blockIdx.x], bodies, states, colliders, num_colliders, gx, gy, gz, dt, steps, frames);
}

uint32_t bundle_lds_bytes(uint32_t nv, uint32_t nrec, uint32_t nlvl, uint32_t nb, bool uniform)
{
    return bundle_layout(nv, nrec, nlvl, nb, uniform ? 12u : 16u).total;
}

uint32_t bundle_leaf_bytes() { return (uint32_t)sizeof(Leaf); }

template <bool S, int BLOCK, bool U, bool EXT>
static hipError_t launch_bundles_variant(const BundleDev *bundles, uint32_t num_bundles, const BodyDev *bodies, BodyStateDev *states,
                                         const ColliderDev *colliders, uint32_t num_colliders, uint32_t lds_bytes,
                                         float gx, float gy, float gz, float dt, int32_t steps, int32_t frames, hipStream_t stream)
{
    static const hipError_t configured = hipFuncSetAttribute((const void *)bundle_kernel<S, BLOCK, U, EXT>,
                                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024);
    if (configured != hipSuccess)
        return configured;
    hipLaunchKernelGGL((bundle_kernel<S, BLOCK, U, EXT>), dim3(num_bundles), dim3(BLOCK), lds_bytes, stream,
                       bundles, bodies, states, colliders, num_colliders, gx, gy, gz, dt, steps, frames);
    return hipGetLastError();
}

template <bool U, bool EXT>
static hipError_t launch_bundles_uniform(bool staged, uint32_t block, const BundleDev *bundles, uint32_t num_bundles,
                                         const BodyDev *bodies, BodyStateDev *states, const ColliderDev *colliders,
                                         uint32_t num_colliders, uint32_t lds_bytes, float gx, float gy, float gz, float dt,
                                         int32_t steps, int32_t frames, hipStream_t stream)
{
    if (staged && block == 256)
        return launch_bundles_variant<true, 256, U, EXT>(bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
    if (staged)
        return hipErrorInvalidValue;
    if (block == 1024)
        return launch_bundles_variant<false, 1024, U, EXT>(bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
    if (block == 512)
        return launch_bundles_variant<false, 512, U, EXT>(bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
    if (block == 256)
        return launch_bundles_variant<false, 256, U, EXT>(bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
    return hipErrorInvalidValue;
}

// block: 256, 512 or 1024 threads (staged bundles only in 256-thread launches); uniform:
// the bundles' vertices share one inverse mass and keep three floats per vertex in LDS;
// ext: the collider list holds triangle trees or scaled convex hulls
hipError_t launch_bundles(bool staged, uint32_t block, bool uniform, bool ext, const BundleDev *bundles, uint32_t num_bundles,
                          const BodyDev *bodies, BodyStateDev *states, const ColliderDev *colliders, uint32_t num_colliders,
                          uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                          hipStream_t stream)
{
    if (num_bundles == 0)
        return hipSuccess;
    if (ext)
    {
        if (uniform)
            return launch_bundles_uniform<true, true>(staged, block, bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
        return launch_bundles_uniform<false, true>(staged, block, bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
    }
    if (uniform)
        return launch_bundles_uniform<true, false>(staged, block, bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
    return launch_bundles_uniform<false, false>(staged, block, bundles, num_bundles, bodies, states, colliders, num_colliders, lds_bytes, gx, gy, gz, dt, steps, frames, stream);
}

} // namespace xpbd
