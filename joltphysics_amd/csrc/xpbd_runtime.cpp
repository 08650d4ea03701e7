// Host runtime of the C ABI declared in include/jolt_xpbd.h: settings upload
// with the level schedule, soft-body state, colliders and kernel launches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/jolt_xpbd.h"
#include "jolt_math.h"
#include "xpbd_types.h"

namespace xpbd {
uint32_t resident_lds_bytes(uint32_t nv, uint32_t max_leaves, uint32_t nc, uint32_t nf);
uint32_t leaf_bytes();
hipError_t launch_resident(const BodyDev *bodies, BodyStateDev *states, const uint32_t *body_list, uint32_t num_blocks,
                           const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves, uint32_t lds_bytes,
                           float gx, float gy, float gz, float dt, int32_t steps, int32_t frames, int32_t *error_flag,
                           hipStream_t stream);
} // namespace xpbd

using namespace xpbd;

namespace {

thread_local std::string g_last_error = "no error";

int Fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t _e = (expr);                                                                             \
        if (_e != hipSuccess)                                                                               \
            return Fail(XPBD_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

constexpr uint32_t kMaxLdsBytes = 160 * 1024 - 1024;

// Reference record layouts (SoftBodySharedSettings.h)
struct VertexRec { float pos[3], vel[3], inv_mass; };
struct FaceRec { uint32_t v[3], material; };
struct EdgeRec { uint32_t v[2]; float rest, compliance; };
struct BendRec { uint32_t v[4]; float compliance, initial_angle; };
struct VolumeRec { uint32_t v[4]; float six_rest_volume, compliance; };
struct LRARec { uint32_t v[2]; float max_distance; };
struct GroupRec { uint32_t edge_end, lra_end, rod_ss_end, rod_bt_end, bend_end, volume_end, skinned_end; };
static_assert(sizeof(VertexRec) == 28 && sizeof(FaceRec) == 16 && sizeof(EdgeRec) == 16, "record size");
static_assert(sizeof(BendRec) == 24 && sizeof(VolumeRec) == 24 && sizeof(LRARec) == 12 && sizeof(GroupRec) == 28, "record size");

template <class T>
struct DevBuf
{
    T *p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t count)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count == 0)
            return hipSuccess;
        return hipMalloc((void **)&p, sizeof(T) * count);
    }
    hipError_t upload(const std::vector<T> &v)
    {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess || v.empty())
            return e;
        return hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
    }
};

} // namespace

// ---------------------------------------------------------------------------
struct xpbd_settings
{
    xpbd_context *ctx = nullptr;
    int refcount = 1;
    uint32_t num_vertices = 0, num_faces = 0;
    std::vector<VertexRec> vertices;        // host copy for body creation
    DevBuf<uint4> recs;
    DevBuf<uint32_t> lvl;
    DevBuf<uint2> groups;
    DevBuf<uint32_t> faces, vf_begin, vf_list;
    uint32_t num_parallel = 0;
    uint32_t tail_one_wave = 0;
    uint32_t stats[5] = { 0, 0, 0, 0, 0 };
};

struct Body
{
    xpbd_settings *settings = nullptr;
    uint32_t nv = 0;
    DevBuf<float4> x, v, prev, plane;
    DevBuf<int32_t> idx, contact;
    BodyDev dev{};
    BodyStateDev state{};
};

struct xpbd_context
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Body>> bodies;
    DevBuf<BodyDev> d_bodies;
    DevBuf<BodyStateDev> d_states;
    DevBuf<uint32_t> d_list;
    bool tables_dirty = true;
    bool states_on_device = false;
    std::vector<ColliderDev> colliders;
    DevBuf<ColliderDev> d_colliders;
    DevBuf<int32_t> d_error;
    float gravity[3] = { 0.0f, -9.81f, 0.0f };       // PhysicsSystem.h:392 default
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double profiled_ms = 0.0;
    int32_t profiled_launches = 0;
};

namespace {

// ------------------------------------------------------------------ schedule
// Converts the reference's sequential group order into levels.  For each
// constraint c in processing order (ProcessGroup: volume, bend, edge, LRA;
// SoftBodyMotionProperties.cpp:1030-1055) with read set R and write set W:
//     level(c) = 1 + max( last_write[r] for r in R,
//                         max(last_write[w], last_read[w]) for w in W )
// so every pair that shares a vertex with at least one writer keeps its
// order, and constraints of one level touch disjoint write sets.
struct Item { uint32_t type; uint32_t v[4]; uint32_t nv; float p0, p1; };

void LevelGroup(const std::vector<Item> &seq, std::vector<uint32_t> &last_write, std::vector<uint32_t> &last_read,
                std::vector<uint4> &recs, std::vector<uint32_t> &lvl, uint32_t &num_levels, uint32_t &max_width)
{
    std::vector<uint32_t> level(seq.size());
    uint32_t nl = 0;
    for (size_t k = 0; k < seq.size(); ++k)
    {
        const Item &it = seq[k];
        uint32_t L = 0;
        bool lra = it.type == kTypeLRA;
        for (uint32_t j = 0; j < it.nv; ++j)
        {
            uint32_t v = it.v[j];
            L = std::max(L, last_write[v]);
            bool writes = !lra || j == 1;
            if (writes)
                L = std::max(L, last_read[v]);
        }
        ++L;
        level[k] = L;
        for (uint32_t j = 0; j < it.nv; ++j)
        {
            uint32_t v = it.v[j];
            bool writes = !lra || j == 1;
            if (writes)
                last_write[v] = L;
            last_read[v] = std::max(last_read[v], L);
        }
        nl = std::max(nl, L);
    }
    // counting sort by level (stable)
    std::vector<uint32_t> count(nl + 2, 0);
    for (uint32_t L : level) ++count[L];
    std::vector<uint32_t> start(nl + 2, 0);
    for (uint32_t L = 1; L <= nl; ++L) start[L + 1] = start[L] + count[L];
    uint32_t base = (uint32_t)recs.size();
    recs.resize(base + seq.size());
    max_width = 0;
    for (uint32_t L = 1; L <= nl; ++L)
    {
        lvl.push_back(base + start[L]);
        max_width = std::max(max_width, count[L]);
    }
    lvl.push_back(base + (uint32_t)seq.size());
    std::vector<uint32_t> fill(start.begin(), start.end());
    for (size_t k = 0; k < seq.size(); ++k)
    {
        const Item &it = seq[k];
        uint4 r;
        r.x = (it.v[0] & 0x3fffu) | (it.type << 14) | ((it.nv > 1 ? it.v[1] : 0u) << 16);
        r.y = (it.nv > 2 ? it.v[2] : 0u) | ((it.nv > 3 ? it.v[3] : 0u) << 16);
        float p0 = it.p0, p1 = it.p1;
        memcpy(&r.z, &p0, 4);
        memcpy(&r.w, &p1, 4);
        recs[base + fill[level[k]]++] = r;
    }
    num_levels = nl;
    // reset the per-vertex trackers touched by this group
    for (const Item &it : seq)
        for (uint32_t j = 0; j < it.nv; ++j)
            last_write[it.v[j]] = last_read[it.v[j]] = 0;
}

int BuildSchedule(const xpbd_settings_desc &d, xpbd_settings &s)
{
    const uint32_t nv = d.num_vertices;
    if (nv == 0 || d.vertices == nullptr)
        return Fail(XPBD_E_INVALID, "settings have no vertices");
    if (d.num_skinned || d.num_rod_stretch_shear || d.num_rod_bend_twist)
        return Fail(XPBD_E_UNSUPPORTED, "skinned constraints and Cosserat rods are not supported by this build");
    if (nv >= kMaxWindowVertices)
        return Fail(XPBD_E_CAPACITY, "more than 16383 vertices in one body is not supported by the LDS-resident kernel");
    if (d.num_update_groups == 0 || d.update_groups == nullptr)
        return Fail(XPBD_E_INVALID, "settings have no update groups (SoftBodySharedSettings::Optimize was not run)");

    const EdgeRec *edges = (const EdgeRec *)d.edges;
    const BendRec *bends = (const BendRec *)d.dihedral_bends;
    const VolumeRec *vols = (const VolumeRec *)d.volumes;
    const LRARec *lras = (const LRARec *)d.lras;
    const GroupRec *groups = (const GroupRec *)d.update_groups;

    // validate group ranges
    GroupRec prev{ 0, 0, 0, 0, 0, 0, 0 };
    for (uint32_t g = 0; g < d.num_update_groups; ++g)
    {
        const GroupRec &c = groups[g];
        if (c.edge_end < prev.edge_end || c.lra_end < prev.lra_end || c.bend_end < prev.bend_end || c.volume_end < prev.volume_end
            || c.rod_ss_end != 0 || c.rod_bt_end != 0 || c.skinned_end != 0)
            return Fail(XPBD_E_INVALID, "update group end indices are not monotone or reference unsupported constraint types");
        prev = c;
    }
    if (prev.edge_end != d.num_edges || prev.lra_end != d.num_lras || prev.bend_end != d.num_dihedral_bends || prev.volume_end != d.num_volumes)
        return Fail(XPBD_E_INVALID, "the last update group does not end at the constraint counts");

    auto check = [nv](const uint32_t *v, int n) {
        for (int i = 0; i < n; ++i)
            if (v[i] >= nv)
                return false;
        return true;
    };

    std::vector<uint4> recs;
    std::vector<uint32_t> lvl;
    std::vector<uint2> gtab;
    std::vector<uint32_t> last_write(nv, 0), last_read(nv, 0);
    std::vector<Item> seq;
    prev = GroupRec{ 0, 0, 0, 0, 0, 0, 0 };
    uint32_t max_par_levels = 0, tail_levels = 0, tail_count = 0, tail_width = 0;
    for (uint32_t g = 0; g < d.num_update_groups; ++g)
    {
        const GroupRec &c = groups[g];
        seq.clear();
        for (uint32_t k = prev.volume_end; k < c.volume_end; ++k)
        {
            if (!check(vols[k].v, 4)) return Fail(XPBD_E_INVALID, "volume constraint vertex out of range");
            seq.push_back({ kTypeVolume, { vols[k].v[0], vols[k].v[1], vols[k].v[2], vols[k].v[3] }, 4, vols[k].six_rest_volume, vols[k].compliance });
        }
        for (uint32_t k = prev.bend_end; k < c.bend_end; ++k)
        {
            if (!check(bends[k].v, 4)) return Fail(XPBD_E_INVALID, "bend constraint vertex out of range");
            seq.push_back({ kTypeBend, { bends[k].v[0], bends[k].v[1], bends[k].v[2], bends[k].v[3] }, 4, bends[k].compliance, bends[k].initial_angle });
        }
        for (uint32_t k = prev.edge_end; k < c.edge_end; ++k)
        {
            if (!check(edges[k].v, 2)) return Fail(XPBD_E_INVALID, "edge constraint vertex out of range");
            seq.push_back({ kTypeEdge, { edges[k].v[0], edges[k].v[1], 0, 0 }, 2, edges[k].rest, edges[k].compliance });
        }
        for (uint32_t k = prev.lra_end; k < c.lra_end; ++k)
        {
            if (!check(lras[k].v, 2)) return Fail(XPBD_E_INVALID, "LRA constraint vertex out of range");
            seq.push_back({ kTypeLRA, { lras[k].v[0], lras[k].v[1], 0, 0 }, 2, lras[k].max_distance, 0.0f });
        }
        uint2 gt;
        gt.x = (uint32_t)lvl.size();
        uint32_t nl = 0, width = 0;
        LevelGroup(seq, last_write, last_read, recs, lvl, nl, width);
        gt.y = nl;
        gtab.push_back(gt);
        if (g + 1 < d.num_update_groups)
            max_par_levels = std::max(max_par_levels, nl);
        else
        {
            tail_levels = nl;
            tail_count = (uint32_t)seq.size();
            tail_width = width;
        }
        prev = c;
    }

    s.num_parallel = d.num_update_groups - 1;
    s.tail_one_wave = tail_width <= 64 ? 1u : 0u;
    s.stats[0] = s.num_parallel;
    s.stats[1] = max_par_levels;
    s.stats[2] = tail_levels;
    s.stats[3] = tail_count;
    s.stats[4] = (uint32_t)recs.size();

    // faces and vertex -> face lists (pressure)
    const FaceRec *faces = (const FaceRec *)d.faces;
    std::vector<uint32_t> f3(3ull * d.num_faces);
    std::vector<uint32_t> vf_begin(nv + 1, 0), vf_list;
    for (uint32_t f = 0; f < d.num_faces; ++f)
    {
        if (!check(faces[f].v, 3)) return Fail(XPBD_E_INVALID, "face vertex out of range");
        for (int j = 0; j < 3; ++j)
        {
            f3[3 * f + j] = faces[f].v[j];
            ++vf_begin[faces[f].v[j] + 1];
        }
    }
    for (uint32_t i = 0; i < nv; ++i) vf_begin[i + 1] += vf_begin[i];
    vf_list.resize(vf_begin[nv]);
    {
        std::vector<uint32_t> fill(vf_begin.begin(), vf_begin.end() - 1);
        for (uint32_t f = 0; f < d.num_faces; ++f)
            for (int j = 0; j < 3; ++j)
                vf_list[fill[faces[f].v[j]]++] = f;
    }

    hipError_t e;
    if ((e = s.recs.upload(recs)) != hipSuccess || (e = s.lvl.upload(lvl)) != hipSuccess || (e = s.groups.upload(gtab)) != hipSuccess
        || (e = s.faces.upload(f3)) != hipSuccess || (e = s.vf_begin.upload(vf_begin)) != hipSuccess || (e = s.vf_list.upload(vf_list)) != hipSuccess)
        return Fail(XPBD_E_DEVICE, std::string("settings upload: ") + hipGetErrorString(e));
    s.num_vertices = nv;
    s.num_faces = d.num_faces;
    const VertexRec *vr = (const VertexRec *)d.vertices;
    s.vertices.assign(vr, vr + nv);
    return XPBD_OK;
}

int SyncTables(xpbd_context *ctx)
{
    if (!ctx->tables_dirty)
        return XPBD_OK;
    std::vector<BodyDev> devs;
    std::vector<BodyStateDev> states;
    for (auto &b : ctx->bodies)
    {
        devs.push_back(b->dev);
        states.push_back(b->state);
    }
    // states already on the device are kept (they are pulled before adding bodies)
    HIP_TRY(ctx->d_bodies.upload(devs));
    HIP_TRY(ctx->d_states.upload(states));
    ctx->tables_dirty = false;
    ctx->states_on_device = true;
    return XPBD_OK;
}

int PullStates(xpbd_context *ctx)
{
    if (!ctx->states_on_device || ctx->d_states.n == 0)
        return XPBD_OK;
    std::vector<BodyStateDev> states(ctx->d_states.n);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(states.data(), ctx->d_states.p, sizeof(BodyStateDev) * states.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < states.size() && i < ctx->bodies.size(); ++i)
        ctx->bodies[i]->state = states[i];
    return XPBD_OK;
}

} // namespace

// =========================================================================== C ABI
extern "C" {

const char *xpbd_last_error(void) { return g_last_error.c_str(); }

int xpbd_context_create(int device, xpbd_context **out_ctx)
{
    if (!out_ctx)
        return Fail(XPBD_E_INVALID, "out_ctx is null");
    *out_ctx = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return Fail(XPBD_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= count)
        return Fail(XPBD_E_INVALID, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<xpbd_context> ctx(new xpbd_context);
    ctx->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIP_TRY(ctx->d_error.alloc(1));
    HIP_TRY(hipMemset(ctx->d_error.p, 0, sizeof(int32_t)));
    *out_ctx = ctx.release();
    return XPBD_OK;
}

int xpbd_context_destroy(xpbd_context *ctx)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &ev : ctx->events)
    {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    ctx->bodies.clear();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return XPBD_OK;
}

void *xpbd_context_stream(xpbd_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int xpbd_settings_create(xpbd_context *ctx, const xpbd_settings_desc *desc, xpbd_settings **out_settings)
{
    if (!ctx || !desc || !out_settings)
        return Fail(XPBD_E_INVALID, "null argument");
    *out_settings = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<xpbd_settings> s(new xpbd_settings);
    s->ctx = ctx;
    int rc = BuildSchedule(*desc, *s);
    if (rc != XPBD_OK)
        return rc;
    *out_settings = s.release();
    return XPBD_OK;
}

int xpbd_settings_release(xpbd_settings *s)
{
    if (!s)
        return Fail(XPBD_E_INVALID, "null settings");
    if (--s->refcount == 0)
        delete s;
    return XPBD_OK;
}

int xpbd_settings_stats(const xpbd_settings *s, uint32_t out[5])
{
    if (!s || !out)
        return Fail(XPBD_E_INVALID, "null argument");
    memcpy(out, s->stats, sizeof(s->stats));
    return XPBD_OK;
}

int xpbd_body_create(xpbd_context *ctx, const xpbd_body_desc *d, int32_t *out_body)
{
    if (!ctx || !d || !out_body || !d->settings)
        return Fail(XPBD_E_INVALID, "null argument");
    if (d->settings->ctx != ctx)
        return Fail(XPBD_E_INVALID, "settings belong to another context");
    if (d->num_iterations == 0)
        return Fail(XPBD_E_INVALID, "num_iterations must be > 0");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = PullStates(ctx);
    if (rc != XPBD_OK)
        return rc;
    xpbd_settings *s = d->settings;
    std::unique_ptr<Body> b(new Body);
    b->settings = s;
    b->nv = s->num_vertices;
    const uint32_t nv = b->nv;

    // SoftBodyMotionProperties::Initialize (SoftBodyMotionProperties.cpp:59-105)
    const float *q = d->rotation;
    M44 rot = d->make_rotation_identity ? rotation_translation(q[0], q[1], q[2], q[3], v3(0, 0, 0))
                                        : rotation_translation(0, 0, 0, 1, v3(0, 0, 0));
    std::vector<float4> x(nv), v(nv), plane(nv, make_float4(0, 0, 0, 0));
    V3 bmin = v3(FLT_MAX, FLT_MAX, FLT_MAX), bmax = v3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (uint32_t i = 0; i < nv; ++i)
    {
        const VertexRec &in = s->vertices[i];
        V3 p = mul_point(rot, v3(in.pos[0], in.pos[1], in.pos[2]));
        V3 vel = mul3x3(rot, v3(in.vel[0], in.vel[1], in.vel[2]));
        x[i] = make_float4(p.x, p.y, p.z, in.inv_mass);
        v[i] = make_float4(vel.x, vel.y, vel.z, 0.0f);
        bmin = vmin(bmin, p);
        bmax = vmax(bmax, p);
    }
    std::vector<int32_t> idx(nv, -1), contact(nv, 0);
    HIP_TRY(b->x.upload(x));
    HIP_TRY(b->v.upload(v));
    HIP_TRY(b->prev.upload(x));
    HIP_TRY(b->plane.upload(plane));
    HIP_TRY(b->idx.upload(idx));
    HIP_TRY(b->contact.upload(contact));

    BodyDev &dv = b->dev;
    dv.s.recs = s->recs.p;
    dv.s.lvl = s->lvl.p;
    dv.s.groups = s->groups.p;
    dv.s.faces = s->faces.p;
    dv.s.vf_begin = s->vf_begin.p;
    dv.s.vf_list = s->vf_list.p;
    dv.s.num_vertices = nv;
    dv.s.num_parallel = s->num_parallel;
    dv.s.num_faces = s->num_faces;
    dv.s.tail_one_wave = s->tail_one_wave;
    dv.x = b->x.p; dv.v = b->v.p; dv.prev = b->prev.p; dv.plane = b->plane.p; dv.idx = b->idx.p; dv.contact = b->contact.p;
    dv.iterations = d->num_iterations;
    dv.linear_damping = d->linear_damping;
    dv.max_linear_velocity = d->max_linear_velocity;
    dv.restitution = d->restitution;
    dv.friction = d->friction;
    dv.pressure = d->pressure;
    dv.gravity_factor = d->gravity_factor;
    dv.vertex_radius = d->vertex_radius;
    dv.update_position = d->update_position ? 1 : 0;
    dv.allow_sleeping = d->allow_sleeping ? 1 : 0;

    BodyStateDev &st = b->state;
    memset(&st, 0, sizeof(st));
    for (int k = 0; k < 3; ++k) st.pos[k] = d->position[k];
    if (d->make_rotation_identity)
    {
        st.rot[0] = st.rot[1] = st.rot[2] = 0.0f;
        st.rot[3] = 1.0f;
    }
    else
        for (int k = 0; k < 4; ++k) st.rot[k] = d->rotation[k];
    st.bmin[0] = bmin.x; st.bmin[1] = bmin.y; st.bmin[2] = bmin.z;
    st.bmax[0] = bmax.x; st.bmax[1] = bmax.y; st.bmax[2] = bmax.z;
    memcpy(st.pmin, st.bmin, sizeof(st.pmin));
    memcpy(st.pmax, st.bmax, sizeof(st.pmax));
    st.sleep_timer = 0.0f;
    st.active = 1;

    s->refcount++;
    ctx->bodies.push_back(std::move(b));
    ctx->tables_dirty = true;
    *out_body = (int32_t)(ctx->bodies.size() - 1);
    return XPBD_OK;
}

int xpbd_set_colliders(xpbd_context *ctx, const xpbd_collider_desc *c, uint32_t n)
{
    if (!ctx || (n && !c))
        return Fail(XPBD_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<ColliderDev> v(n);
    for (uint32_t i = 0; i < n; ++i)
    {
        if (c[i].type < XPBD_SHAPE_PLANE || c[i].type > XPBD_SHAPE_CAPSULE)
            return Fail(XPBD_E_UNSUPPORTED, "collider shape type not supported (plane, sphere, box, capsule only)");
        v[i].type = c[i].type;
        memcpy(v[i].params, c[i].params, sizeof(v[i].params));
        memcpy(v[i].pos, c[i].position, sizeof(v[i].pos));
        memcpy(v[i].rot, c[i].rotation, sizeof(v[i].rot));
        v[i].friction = c[i].friction;
        v[i].restitution = c[i].restitution;
        memcpy(v[i].bounds, c[i].world_bounds, sizeof(v[i].bounds));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->colliders = v;
    HIP_TRY(ctx->d_colliders.upload(v));
    return XPBD_OK;
}

int xpbd_set_gravity(xpbd_context *ctx, const float g[3])
{
    if (!ctx || !g)
        return Fail(XPBD_E_INVALID, "null argument");
    memcpy(ctx->gravity, g, sizeof(ctx->gravity));
    return XPBD_OK;
}

int xpbd_update(xpbd_context *ctx, float dt, int32_t collision_steps, int32_t frames)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    if (collision_steps <= 0 || frames < 0 || !(dt > 0.0f))
        return Fail(XPBD_E_INVALID, "dt must be > 0, collision_steps > 0, frames >= 0");
    if (frames == 0 || ctx->bodies.empty())
        return XPBD_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = SyncTables(ctx);
    if (rc != XPBD_OK)
        return rc;

    const uint32_t nc = (uint32_t)ctx->colliders.size();
    // Largest bodies first so the long blocks start early.
    std::vector<uint32_t> order(ctx->bodies.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [ctx](uint32_t a, uint32_t b) { return ctx->bodies[a]->nv > ctx->bodies[b]->nv; });
    uint32_t lds = 0;
    for (uint32_t i : order)
    {
        const Body &b = *ctx->bodies[i];
        uint32_t nf = b.dev.pressure > 0.0f ? b.settings->num_faces : 0u;
        uint32_t need = resident_lds_bytes(b.nv, nc, nc, nf);
        if (need > kMaxLdsBytes)
            return Fail(XPBD_E_CAPACITY, "body " + std::to_string(i) + " needs " + std::to_string(need)
                                             + " bytes of LDS; the LDS-resident kernel allows " + std::to_string(kMaxLdsBytes));
        lds = std::max(lds, need);
    }
    if (ctx->d_list.n != order.size())
        HIP_TRY(ctx->d_list.alloc(order.size()));
    HIP_TRY(hipMemcpyAsync(ctx->d_list.p, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice, ctx->stream));

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profiling)
    {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, ctx->stream));
    }
    HIP_TRY(launch_resident(ctx->d_bodies.p, ctx->d_states.p, ctx->d_list.p, (uint32_t)order.size(), ctx->d_colliders.p, nc, nc, lds,
                            ctx->gravity[0], ctx->gravity[1], ctx->gravity[2], dt, collision_steps, frames, ctx->d_error.p, ctx->stream));
    if (ctx->profiling)
    {
        HIP_TRY(hipEventRecord(e1, ctx->stream));
        ctx->events.emplace_back(e0, e1);
    }
    return XPBD_OK;
}

int xpbd_synchronize(xpbd_context *ctx)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int32_t err = 0;
    HIP_TRY(hipMemcpy(&err, ctx->d_error.p, sizeof(err), hipMemcpyDeviceToHost));
    if (err)
    {
        HIP_TRY(hipMemset(ctx->d_error.p, 0, sizeof(int32_t)));
        return Fail(XPBD_E_CAPACITY, "a body collided with more colliders than the leaf capacity");
    }
    return XPBD_OK;
}

int xpbd_body_read(xpbd_context *ctx, int32_t body, float *pos, float *prev, float *vel, float *plane, int32_t *shape_idx,
                   int32_t *contact, float *body_pos, float *body_rot)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    rc = PullStates(ctx);
    if (rc != XPBD_OK)
        return rc;
    Body &b = *ctx->bodies[body];
    const uint32_t nv = b.nv;
    std::vector<float4> tmp(nv);
    auto get3 = [&](DevBuf<float4> &src, float *dst, int comps) -> int {
        if (!dst)
            return XPBD_OK;
        HIP_TRY(hipMemcpy(tmp.data(), src.p, sizeof(float4) * nv, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nv; ++i)
        {
            dst[comps * i] = tmp[i].x; dst[comps * i + 1] = tmp[i].y; dst[comps * i + 2] = tmp[i].z;
            if (comps == 4) dst[comps * i + 3] = tmp[i].w;
        }
        return XPBD_OK;
    };
    if ((rc = get3(b.x, pos, 3)) != XPBD_OK || (rc = get3(b.prev, prev, 3)) != XPBD_OK || (rc = get3(b.v, vel, 3)) != XPBD_OK
        || (rc = get3(b.plane, plane, 4)) != XPBD_OK)
        return rc;
    if (shape_idx)
        HIP_TRY(hipMemcpy(shape_idx, b.idx.p, sizeof(int32_t) * nv, hipMemcpyDeviceToHost));
    if (contact)
        HIP_TRY(hipMemcpy(contact, b.contact.p, sizeof(int32_t) * nv, hipMemcpyDeviceToHost));
    if (body_pos)
        memcpy(body_pos, b.state.pos, sizeof(b.state.pos));
    if (body_rot)
        memcpy(body_rot, b.state.rot, sizeof(b.state.rot));
    return XPBD_OK;
}

int xpbd_body_write(xpbd_context *ctx, int32_t body, const float *pos, const float *vel)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    Body &b = *ctx->bodies[body];
    const uint32_t nv = b.nv;
    std::vector<float4> tmp(nv);
    if (pos)
    {
        HIP_TRY(hipMemcpy(tmp.data(), b.x.p, sizeof(float4) * nv, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nv; ++i)
            tmp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], tmp[i].w);
        HIP_TRY(hipMemcpy(b.x.p, tmp.data(), sizeof(float4) * nv, hipMemcpyHostToDevice));
    }
    if (vel)
    {
        for (uint32_t i = 0; i < nv; ++i)
            tmp[i] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f);
        HIP_TRY(hipMemcpy(b.v.p, tmp.data(), sizeof(float4) * nv, hipMemcpyHostToDevice));
    }
    return XPBD_OK;
}

int xpbd_body_state(xpbd_context *ctx, int32_t body, float out[14])
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size() || !out)
        return Fail(XPBD_E_INVALID, "invalid body");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    rc = PullStates(ctx);
    if (rc != XPBD_OK)
        return rc;
    const BodyStateDev &s = ctx->bodies[body]->state;
    memcpy(out, s.lin, 12);
    memcpy(out + 3, s.ang, 12);
    memcpy(out + 6, s.bmin, 12);
    memcpy(out + 9, s.bmax, 12);
    out[12] = (float)s.active;
    out[13] = s.sleep_timer;
    return XPBD_OK;
}

int xpbd_profile_enable(xpbd_context *ctx, int32_t enable)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    ctx->profiling = enable != 0;
    return XPBD_OK;
}

int xpbd_profile_read(xpbd_context *ctx, int32_t *num_launches, double *total_ms)
{
    if (!ctx || !num_launches || !total_ms)
        return Fail(XPBD_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    double total = 0.0;
    for (auto &ev : ctx->events)
    {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second));
        total += ms;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    *num_launches = (int32_t)ctx->events.size();
    *total_ms = total;
    ctx->events.clear();
    return XPBD_OK;
}

} // extern "C"
