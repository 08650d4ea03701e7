// Host runtime of the C ABI declared in include/jolt_xpbd.h: settings upload
// with the level schedule, soft-body state, colliders and kernel launches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <limits>

#include "../../include/jolt_xpbd.h"
#include "jolt_math.h"
#include "xpbd_types.h"

namespace xpbd {
uint32_t kernel_lds_bytes(bool resident, bool staged, const SettingsDev &s, uint32_t max_leaves, uint32_t nc, uint32_t nf);
uint32_t bundle_lds_bytes(uint32_t nv, uint32_t nrec, uint32_t nlvl, uint32_t nb, bool uniform);
uint32_t bundle_leaf_bytes();
hipError_t launch_bundles(bool staged, uint32_t block, bool uniform, bool ext, const BundleDev *bundles, uint32_t num_bundles,
                          const BodyDev *bodies, BodyStateDev *states, const ColliderDev *colliders, uint32_t num_colliders,
                          uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                          hipStream_t stream);
hipError_t launch_update(bool resident, bool staged, bool rods, bool ext, const BodyDev *bodies, BodyStateDev *states, const uint32_t *body_list,
                         uint32_t num_blocks, const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves,
                         uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                         int32_t *error_flag, hipStream_t stream);
uint32_t team_lds_bytes(uint32_t window_capacity, uint32_t max_leaves, uint32_t nc);
hipError_t launch_team(bool ext, const BodyDev *bodies, BodyStateDev *states, const TeamDev *teams, const uint2 *block_map,
                       uint32_t num_blocks, const ColliderDev *colliders, uint32_t num_colliders, uint32_t max_leaves,
                       uint32_t lds_bytes, float gx, float gy, float gz, float dt, int32_t steps, int32_t frames,
                       int32_t *error_flag, hipStream_t stream);
hipError_t launch_skin(const float *joints, const SkinInvBindDev *inv_binds, uint32_t num_inv_binds, float *skin,
                       const SkinnedDev *recs, uint32_t num_recs, const uint32_t *normal_faces, const uint32_t *faces,
                       float4 *states, uint32_t nv, float4 *x, float4 *v, int32_t hard_skin_all, hipStream_t stream);
uint32_t rollout_forward_lds_bytes(uint32_t nv, uint32_t nf, bool f64);
uint32_t rollout_backward_lds_bytes(uint32_t nv, uint32_t ne, uint32_t nf);
hipError_t launch_rollout_forward_f32(const RolloutDev &R, const float *compliance, const float *velocity, float *positions,
                                      uint32_t lds, hipStream_t stream);
hipError_t launch_rollout_forward_f64(const RolloutDev &R, const double *compliance, const double *velocity, double *positions,
                                      uint32_t lds, hipStream_t stream);
hipError_t launch_rollout_backward(const RolloutDev &R, const float *grad_positions, float *grad_compliance, float *grad_velocity,
                                   uint32_t lds, hipStream_t stream);
hipError_t rollout_scratch_setup(int device);
uint32_t rollout_noalloc_max_vertices();
hipError_t launch_rollout_forward_f32_noalloc(int device, RolloutDev R, const float *compliance, const float *velocity,
                                              float *positions, uint32_t lds, hipStream_t stream);
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
constexpr uint32_t kNopRecordHost = 0xffffffffu;    // padding record (skipped by the kernels)
constexpr uint32_t kSecondHalf = 1u << 15;          // bend record: the second lane's half (in the vertex-2 field)

// Reference record layouts (SoftBodySharedSettings.h)
struct VertexRec { float pos[3], vel[3], inv_mass; };
struct FaceRec { uint32_t v[3], material; };
struct EdgeRec { uint32_t v[2]; float rest, compliance; };
struct BendRec { uint32_t v[4]; float compliance, initial_angle; };
struct VolumeRec { uint32_t v[4]; float six_rest_volume, compliance; };
struct LRARec { uint32_t v[2]; float max_distance; };
struct GroupRec { uint32_t edge_end, lra_end, rod_ss_end, rod_bt_end, bend_end, volume_end, skinned_end; };
struct SkinWeightRec { uint32_t inv_bind; float weight; };
struct SkinnedRec { uint32_t vertex; SkinWeightRec w[4]; float max_distance, back_stop_distance, back_stop_radius; uint32_t normal_info; };
struct InvBindRec { uint32_t joint; uint32_t pad[3]; float m[16]; };
static_assert(sizeof(SkinnedRec) == 52 && sizeof(InvBindRec) == 80, "skin record size");
// Cosserat rods as the reference holds them (SoftBodySharedSettings.h:312-342; Quat is 16-byte aligned)
struct RodSSRec { uint32_t v[2]; float length, inv_mass, compliance; uint32_t pad[3]; float bishop[4]; };
struct RodBTRec { uint32_t rod[2]; float compliance; uint32_t pad; float omega0[4]; };
static_assert(sizeof(RodSSRec) == 48 && sizeof(RodBTRec) == 32, "rod record size");
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
    DevBuf<WindowDev> windows;
    DevBuf<uint32_t> vlist;
    DevBuf<uint32_t> faces, vf_begin, vf_list;
    SettingsDev dev{};
    uint32_t stats[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    // host copy of a resident body's level sequence (merged parallel levels,
    // then the non-parallel group's levels): records with body vertex indices
    std::vector<uint4> host_recs;
    std::vector<uint32_t> host_level_begin;     // nlev + 1 offsets into host_recs
    std::vector<uint32_t> host_rec_index;       // constraint index of each host record (no-op: ~0u)
    // skinning: device tables for SkinVertices, host copies for the checks and state streams
    DevBuf<SkinnedDev> d_skinned;
    DevBuf<SkinInvBindDev> d_inv_binds;
    DevBuf<uint32_t> d_skinned_normals;
    std::vector<SkinnedRec> skinned;
    std::vector<uint32_t> skinned_normals;
    std::vector<InvBindRec> inv_binds;
    std::vector<FaceRec> face_recs;
    // Cosserat rods: host copy (Bishop frames for body creation) and device constants
    std::vector<RodSSRec> rods;
    DevBuf<float4> rod_const, bt_const;
    bool uniform_mass = false;                  // every vertex has the same inverse mass
    float inv_mass = 0.0f;                      // that inverse mass
    uint32_t num_edges = 0, num_bends = 0, num_volumes = 0, num_lras = 0;
    // streamed bodies: the parallel-group windows (host copy, for splitting them over a team)
    // and the coloured non-parallel group (empty when it holds skinned or rod constraints)
    std::vector<WindowDev> par_windows;
    DevBuf<uint4> colour_recs;
    DevBuf<uint32_t> colour_begin;
    uint32_t num_colours = 0;
    uint32_t colour_records = 0;
    bool colourable = false;
    // settings built on the host only (xpbd_schedule_build): the programs the device would get
    struct HostSchedule
    {
        std::vector<uint4> recs;
        std::vector<uint32_t> rec_index, lvl, vlist, colour_begin;
        std::vector<WindowDev> windows;
        std::vector<uint4> colour_recs;
    };
    std::unique_ptr<HostSchedule> host_schedule;
};

struct Body
{
    xpbd_settings *settings = nullptr;
    uint32_t nv = 0;
    bool owns_storage = true;       // false once the state lives in a bundle's arrays
    DevBuf<float4> x, v, prev, plane, report;
    DevBuf<int32_t> idx, contact;
    DevBuf<float4> rods;            // 2 per rod (the body keeps them; rod bodies never join a bundle)
    DevBuf<float4> pair;            // per collider contact settings (xpbd_body_set_contact_settings)
    uint32_t pair_n = 0;            // colliders the table was given for (0: no table)
    std::vector<uint32_t> pair_sig; // the collider list's signature when the table was given
    int32_t tail_mode = 0;          // XPBD_TAIL_EXACT or XPBD_TAIL_COLOURED
    BodyDev dev{};
    BodyStateDev state{};
    bool skinned = false;           // has skin entries at dev.skin_base (one per vertex)
    DevBuf<float> d_joints, d_skin_m;   // SkinVertices: joint matrices, skin matrices (16 floats each)
};

struct BundleHost
{
    std::vector<uint32_t> bodies;
    uint32_t nv = 0;
    bool staged = false;
    bool uniform = false;
    float inv_mass = 0.0f;
    uint32_t lds = 0;
    DevBuf<uint4> recs;
    DevBuf<uint32_t> lvl;
    DevBuf<BundleBodyDev> table;
    DevBuf<float4> x, v, prev, plane, report, face_scratch;
    DevBuf<int32_t> idx, contact;
    DevBuf<unsigned char> leaves_full, leaves;
    BundleDev dev{};
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
    std::vector<ColliderDev> colliders;     // leaf shapes of the collider bodies, body by body
    std::vector<uint32_t> collider_first;  // per collider body: its first leaf in `colliders`
    std::vector<uint32_t> collider_sig;    // per collider body: type | motion << 8 | sensor << 16 | parts << 17, user id
    DevBuf<ColliderDev> d_colliders;
    DevBuf<PartOpDev> d_part_ops;          // placement steps of compound leaves
    DevBuf<float4> d_hull_vec;      // convex hulls: planes, then points, per hull
    DevBuf<uint32_t> d_hull_idx;    // convex hulls: face first, face count, vertex index, per hull
    DevBuf<HullDev> d_hulls;
    DevBuf<TreeNodeDev> d_tree_nodes;   // triangle trees: nodes of every tree, references rebased
    DevBuf<float> d_tree_tris;          // triangle trees: 9 floats per triangle, every tree
    DevBuf<TreeDev> d_trees;
    bool ext_colliders = false;         // a triangle tree, a scaled convex hull or a compound leaf among the colliders
    // skinned vertex states of all skinned bodies, 4 float4 per vertex: previous and current
    // skinned position, normal, {max distance multiplier, back stop radius, enabled, previous
    // position valid}; kept on the host, uploaded before an update when changed
    std::vector<float4> skin_host;
    DevBuf<float4> d_skin;
    bool skin_dirty = false;        // host copy newer: upload before the next update
    bool skin_on_device = false;    // device copy newer (updates move and invalidate skin states)
    uint32_t num_sensors = 0;       // sensor colliders of the last xpbd_set_colliders
    DevBuf<int32_t> d_sensor_hits;  // bodies x num_sensors
    uint32_t sensor_stride = 0;
    uint32_t num_dynamic = 0;       // dynamic colliders of the last xpbd_set_colliders
    DevBuf<DynStateDev> d_dyn;      // bodies x num_dynamic
    uint32_t dyn_stride = 0;        // num_dynamic the rows of d_dyn are laid out for
    DevBuf<int32_t> d_error;
    float gravity[3] = { 0.0f, -9.81f, 0.0f };       // PhysicsSystem.h:392 default
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double profiled_ms = 0.0;
    int32_t profiled_launches = 0;
    std::vector<uint32_t> host_list;   // sources of the asynchronous table uploads
    std::vector<BodyDev> host_devs;
    // bundles of LDS-resident bodies and the streamed bodies
    bool layout_dirty = true;
    uint32_t layout_colliders = 0;
    std::vector<std::unique_ptr<BundleHost>> bundles;
    std::vector<uint32_t> streamed;
    // large streamed bodies run by teams of workgroups (xpbd_team.hip)
    uint32_t team_min_windows = 3;  // a streamed body with this many parallel-group windows gets a team (0: never)
    uint32_t team_size = 0;         // workgroups per team at least (0: one per parallel-group window)
    std::vector<TeamDev> teams;
    std::vector<uint2> team_blocks; // per workgroup of the team launch: team, member
    DevBuf<TeamDev> d_teams;
    DevBuf<uint2> d_team_blocks;
    DevBuf<uint32_t> d_team_split;
    DevBuf<uint32_t> d_team_sync;   // one 256-byte line per team
    DevBuf<float> d_team_red;
    DevBuf<unsigned char> d_team_leaves;   // per team member, one leaf per collider
    uint32_t team_lds = 0, team_max_leaves = 0;
    struct BundleLaunch
    {
        bool staged = false;    // records staged in LDS (256-thread launches)
        bool uniform = false;   // 12 B LDS positions, one inverse mass per bundle
        uint32_t block = 256;   // threads per workgroup
        uint32_t lds = 0;       // LDS bytes per workgroup
        DevBuf<BundleDev> table;
        uint32_t count = 0;
        double cost = 0.0;      // estimated duration of its longest workgroup (µs per collision step)
        double work = 0.0;      // estimated CU-microseconds per collision step
        std::vector<BundleDev> host;    // the table in launch order
    };
    std::vector<std::unique_ptr<BundleLaunch>> bundle_launches;
    // Measured duration of every bundle's workgroup (realtime-clock ticks, one slot
    // per bundle).  The first update of a layout reads them back; the next one
    // orders each launch's bundles longest first by them (workgroups are
    // dispatched in index order, so the longest ones start first).
    DevBuf<unsigned long long> d_elapsed;
    unsigned long long *h_elapsed = nullptr;    // pinned host copy
    size_t h_elapsed_n = 0;
    hipEvent_t elapsed_ready = nullptr;
    enum { kOrderMeasure, kOrderPending, kOrderMeasured } order_state = kOrderMeasure;
    // The context stream plus auxiliary streams: one per hardware queue the HIP
    // runtime gives the process (GPU_MAX_HW_QUEUES, 4 by default; at most 8 used).
    // A stream beyond that would share a queue and serialise behind it.
    static constexpr int kMaxAux = 7;
    int num_aux = 3;
    hipStream_t aux[kMaxAux] = {};
    hipEvent_t fork = nullptr, join[kMaxAux] = {};
    bool use_bundles = true;
};

namespace {

// ------------------------------------------------------------------ schedule
// Converts the reference's sequential group order into levels.  For each
// constraint c in processing order (ProcessGroup: volume, bend, edge, LRA;
// SoftBodyMotionProperties.cpp:1030-1055) with read set R and write set W:
//     level(c) = 1 + max( last_write[r] for r in R,
//                         max(last_write[w], last_read[w]) for w in W )
// so every pair that shares a vertex with at least one writer keeps its
// order, and the constraints of one level have disjoint write sets that no
// other constraint of the level reads.
struct Item { uint32_t type; uint32_t v[4]; uint32_t nv; float p0, p1; uint32_t index; uint32_t rod[2]; uint32_t nrod; };   // index: in its type's array
// Skinned constraint item (host only): packed as an LRA record of vertex 0 to itself
constexpr uint32_t kItemSkin = 4;
// Cosserat rod items (host only): packed as LRA-type records flagged in y.  Their rod
// states are resources of the schedule besides the vertices (Item::rod).
constexpr uint32_t kItemRodSS = 5, kItemRodBT = 6;

// Bodies up to this many vertices keep all positions in LDS for the launch.
constexpr uint32_t kResidentMaxVertices = 8192;
// Vertex capacity of one gathered window of a streamed body.
constexpr uint32_t kWindowCapacity = 8192;

// Rod states are read and written by every rod item that names them: rod_last keeps
// the level of their last use.
std::vector<std::vector<uint32_t>> LevelGroup(const std::vector<Item> &seq, std::vector<uint32_t> &last_write, std::vector<uint32_t> &last_read,
                                              std::vector<uint32_t> &rod_last)
{
    std::vector<uint32_t> level(seq.size());
    uint32_t nl = 0;
    for (size_t k = 0; k < seq.size(); ++k)
    {
        const Item &it = seq[k];
        const bool lra = it.type == kTypeLRA;
        uint32_t L = 0;
        for (uint32_t j = 0; j < it.nrod; ++j)
            L = std::max(L, rod_last[it.rod[j]]);
        for (uint32_t j = 0; j < it.nv; ++j)
        {
            uint32_t v = it.v[j];
            L = std::max(L, last_write[v]);
            if (!lra || j == 1)
                L = std::max(L, last_read[v]);
        }
        ++L;
        level[k] = L;
        for (uint32_t j = 0; j < it.nrod; ++j)
            rod_last[it.rod[j]] = L;
        for (uint32_t j = 0; j < it.nv; ++j)
        {
            uint32_t v = it.v[j];
            if (!lra || j == 1)
                last_write[v] = L;
            last_read[v] = std::max(last_read[v], L);
        }
        nl = std::max(nl, L);
    }
    std::vector<std::vector<uint32_t>> out(nl);
    for (size_t k = 0; k < seq.size(); ++k)
        out[level[k] - 1].push_back((uint32_t)k);
    for (const Item &it : seq)
    {
        for (uint32_t j = 0; j < it.nv; ++j)
            last_write[it.v[j]] = last_read[it.v[j]] = 0;
        for (uint32_t j = 0; j < it.nrod; ++j)
            rod_last[it.rod[j]] = 0;
    }
    return out;
}

// The throughput schedules (XPBD_SCHEDULE_COLOURED, XPBD_SCHEDULE_COLOURED_TAIL): within each
// run of one constraint type (ProcessGroup's order of the types is kept), a constraint takes the
// lowest colour that no earlier constraint of the run holds on a vertex they share where either
// of them writes (an LRA reads its vertex 0); each colour becomes one level, colours in order.
// Constraints of one colour touch disjoint written vertices, so the levels give exactly what the
// reference computes for settings whose constraints are sorted by colour (stably) within each
// run of each update group.  False for a group with rod constraints or a run needing more than
// 64 * kColourBanks colours (the caller keeps the exact levels).  wmask / rmask: nv * kColourBanks
// words, zero on entry and on return.
constexpr uint32_t kColourBanks = 4;
bool ColourLevels(const std::vector<Item> &seq, std::vector<uint64_t> &wmask, std::vector<uint64_t> &rmask,
                  std::vector<std::vector<uint32_t>> &out)
{
    out.clear();
    std::vector<uint32_t> touched;
    bool ok = true;
    size_t k = 0;
    while (k < seq.size() && ok)
    {
        const uint32_t type = seq[k].type;
        if (type == kItemRodSS || type == kItemRodBT)
        {
            ok = false;
            break;
        }
        size_t e = k;
        while (e < seq.size() && seq[e].type == type)
            ++e;
        const bool lra = type == kTypeLRA;
        const size_t base = out.size();
        for (size_t i = k; i < e && ok; ++i)
        {
            const Item &it = seq[i];
            uint32_t c = ~0u;
            for (uint32_t bank = 0; bank < kColourBanks && c == ~0u; ++bank)
            {
                uint64_t forbidden = 0;
                for (uint32_t j = 0; j < it.nv; ++j)
                {
                    const size_t m = (size_t)it.v[j] * kColourBanks + bank;
                    forbidden |= wmask[m];
                    if (!lra || j == 1)
                        forbidden |= rmask[m];
                }
                if (~forbidden)
                    c = bank * 64u + (uint32_t)__builtin_ctzll(~forbidden);
            }
            if (c == ~0u)
            {
                ok = false;
                break;
            }
            for (uint32_t j = 0; j < it.nv; ++j)
            {
                const size_t m = (size_t)it.v[j] * kColourBanks + c / 64u;
                ((!lra || j == 1) ? wmask : rmask)[m] |= 1ull << (c & 63u);
                touched.push_back(it.v[j]);
            }
            if (out.size() <= base + c)
                out.resize(base + c + 1);
            out[base + c].push_back((uint32_t)i);
        }
        for (uint32_t v : touched)
            for (uint32_t bank = 0; bank < kColourBanks; ++bank)
                wmask[(size_t)v * kColourBanks + bank] = rmask[(size_t)v * kColourBanks + bank] = 0;
        touched.clear();
        k = e;
    }
    if (!ok)
    {
        for (const Item &it : seq)
            for (uint32_t j = 0; j < it.nv; ++j)
                for (uint32_t bank = 0; bank < kColourBanks; ++bank)
                    wmask[(size_t)it.v[j] * kColourBanks + bank] = rmask[(size_t)it.v[j] * kColourBanks + bank] = 0;
        out.clear();
    }
    return ok;
}

uint4 Pack(const Item &it, const uint32_t *local)
{
    auto m = [&](uint32_t j) { return j < it.nv ? (local ? local[it.v[j]] : it.v[j]) : 0u; };
    uint4 r;
    if (it.type == kItemRodSS || it.type == kItemRodBT)
    {
        // stretch-shear: vertices 0, 1 and the rod; bend-twist: no vertices, the constraint
        r.x = it.type == kItemRodSS ? ((m(0) & 0x3fffu) | (kTypeLRA << 14) | (m(1) << 16)) : (kTypeLRA << 14);
        r.y = (it.type == kItemRodSS ? kRodStretchShear : kRodBendTwist) | it.index;
        r.z = r.w = 0u;
        return r;
    }
    if (it.type == kItemSkin)
    {
        // LRA type, vertex 0 twice; y = the body vertex (the skin entry, before the body's base)
        r.x = (m(0) & 0x3fffu) | (kTypeLRA << 14) | (m(0) << 16);
        r.y = it.v[0];
        float p0 = it.p0, p1 = it.p1;
        memcpy(&r.z, &p0, 4);
        memcpy(&r.w, &p1, 4);
        return r;
    }
    r.x = (m(0) & 0x3fffu) | (it.type << 14) | (m(1) << 16);
    r.y = m(2) | (m(3) << 16);
    float p0 = it.p0, p1 = it.p1;
    memcpy(&r.z, &p0, 4);
    memcpy(&r.w, &p1, 4);
    return r;
}

struct GroupLevels
{
    std::vector<Item> seq;
    std::vector<std::vector<uint32_t>> levels;
};

// Appends one level: records grouped by constraint type (bend, volume, edge,
// LRA), each group padded with no-op records to whole wavefronts when that does
// not add a pass of the `block`-thread workgroup, so a wavefront runs one
// constraint path.  `allow_pad` false keeps the groups unpadded.
void AppendTypedLevel(std::vector<uint4> bucket[4], std::vector<uint4> &recs, bool allow_pad = true,
                      const std::vector<uint32_t> *ibucket = nullptr, std::vector<uint32_t> *rec_index = nullptr,
                      uint32_t block = 256)
{
    size_t n = 0, padded = 0;
    for (int t = 0; t < 4; ++t)
    {
        const size_t slots = bucket[t].size() * (t == (int)kTypeBend ? 2 : 1);
        n += slots;
        padded += (slots + 63) / 64 * 64;
    }
    const bool pad = allow_pad && (padded + block - 1) / block <= (n + block - 1) / block;
    const size_t base = recs.size();
    const uint4 nop = make_uint4(0xffffffffu, 0u, 0u, 0u);
    const int order[4] = { (int)kTypeBend, (int)kTypeVolume, (int)kTypeEdge, (int)kTypeLRA };
    for (int t : order)
    {
        if (t == (int)kTypeBend)
        {
            // a bend occupies two adjacent lanes: the record, then its second half
            // (bit 15 of the vertex-2 field set); the group starts the level, so the
            // halves sit at an even and the following odd lane
            for (size_t k = 0; k < bucket[t].size(); ++k)
            {
                uint4 second = bucket[t][k];
                second.y |= kSecondHalf;
                recs.push_back(bucket[t][k]);
                recs.push_back(second);
                if (rec_index)
                {
                    rec_index->push_back(ibucket[t][k]);
                    rec_index->push_back(ibucket[t][k]);
                }
            }
        }
        else
        {
            recs.insert(recs.end(), bucket[t].begin(), bucket[t].end());
            if (rec_index)
                rec_index->insert(rec_index->end(), ibucket[t].begin(), ibucket[t].end());
        }
        if (pad)
            while ((recs.size() - base) % 64 != 0)
            {
                recs.push_back(nop);
                if (rec_index)
                    rec_index->push_back(~0u);
            }
    }
}

struct ProgramBuilder
{
    std::vector<uint4> recs;
    std::vector<uint32_t> rec_index;        // constraint index of each record (no-op: ~0u)
    std::vector<uint32_t> lvl;
    std::vector<WindowDev> windows;
    std::vector<uint32_t> vlist;

    void AppendLevel(const std::vector<std::pair<const struct GroupLevels *, uint32_t>> &lv, const uint32_t *local);

    // Appends one window: levels[L] = list of (group, item) pairs; local maps
    // body vertex -> window slot (null: identity window).
    void AddWindow(const std::vector<std::vector<std::pair<const GroupLevels *, uint32_t>>> &levels, const uint32_t *local,
                   const std::vector<uint32_t> &verts)
    {
        WindowDev w;
        w.vbegin = (uint32_t)vlist.size();
        w.vcount = local ? (uint32_t)verts.size() : 0u;
        if (local)
            vlist.insert(vlist.end(), verts.begin(), verts.end());
        w.lbegin = (uint32_t)lvl.size();
        w.nlev = (uint32_t)levels.size();
        for (const auto &lv : levels)
        {
            lvl.push_back((uint32_t)recs.size());
            AppendLevel(lv, local);
        }
        lvl.push_back((uint32_t)recs.size());
        windows.push_back(w);
    }
};

void ProgramBuilder::AppendLevel(const std::vector<std::pair<const GroupLevels *, uint32_t>> &lv, const uint32_t *local)
{
    std::vector<uint4> bucket[4];
    std::vector<uint32_t> ibucket[4];
    for (const auto &gi : lv)
    {
        const Item &it = gi.first->seq[gi.second];
        const uint32_t t = it.type >= kItemSkin ? kTypeLRA : it.type;
        bucket[t].push_back(Pack(it, local));
        ibucket[t].push_back(it.index);
    }
    AppendTypedLevel(bucket, recs, true, ibucket, &rec_index);
}

// The coloured non-parallel group (team bodies in the coloured mode).  Within each run of
// one constraint type, in the reference's order (ProcessGroup applies the types one after
// the other), a constraint takes the lowest colour that no earlier constraint of the run
// holds on a vertex they share where either of them writes (an LRA reads its vertex 0);
// the constraints of one colour can then be projected in any order.  Colours run in order.
// Returns false when the group holds constraints other than volumes, bends, edges and LRAs.
bool ColourGroup(const std::vector<Item> &seq, uint32_t nv, std::vector<uint4> &recs, std::vector<uint32_t> &begin)
{
    constexpr uint32_t kBanks = 4;      // up to 256 colours per run
    std::vector<uint64_t> wmask((size_t)nv * kBanks, 0), rmask((size_t)nv * kBanks, 0);
    std::vector<uint32_t> touched;
    recs.clear();
    begin.clear();
    size_t k = 0;
    while (k < seq.size())
    {
        const uint32_t type = seq[k].type;
        if (type >= kItemSkin)
            return false;
        size_t e = k;
        while (e < seq.size() && seq[e].type == type)
            ++e;
        const bool lra = type == kTypeLRA;
        std::vector<std::vector<uint32_t>> colour;
        for (size_t i = k; i < e; ++i)
        {
            const Item &it = seq[i];
            uint32_t c = ~0u;
            for (uint32_t bank = 0; bank < kBanks && c == ~0u; ++bank)
            {
                uint64_t forbidden = 0;
                for (uint32_t j = 0; j < it.nv; ++j)
                {
                    const size_t m = (size_t)it.v[j] * kBanks + bank;
                    forbidden |= wmask[m];
                    if (!lra || j == 1)
                        forbidden |= rmask[m];
                }
                if (~forbidden)
                    c = bank * 64u + (uint32_t)__builtin_ctzll(~forbidden);
            }
            if (c == ~0u)
                return false;
            const uint64_t bit = 1ull << (c & 63u);
            for (uint32_t j = 0; j < it.nv; ++j)
            {
                const size_t m = (size_t)it.v[j] * kBanks + c / 64u;
                if (!lra || j == 1)
                    wmask[m] |= bit;
                else
                    rmask[m] |= bit;
                touched.push_back(it.v[j]);
            }
            if (colour.size() <= c)
                colour.resize(c + 1);
            colour[c].push_back((uint32_t)i);
        }
        for (const auto &members : colour)
        {
            begin.push_back((uint32_t)(recs.size() / 2));
            for (uint32_t i : members)
            {
                const Item &it = seq[i];
                float p0 = it.p0, p1 = it.p1;
                uint4 b;
                b.x = it.type;
                memcpy(&b.y, &p0, 4);
                memcpy(&b.z, &p1, 4);
                b.w = 0u;
                recs.push_back(make_uint4(it.v[0], it.nv > 1 ? it.v[1] : 0u, it.nv > 2 ? it.v[2] : 0u, it.nv > 3 ? it.v[3] : 0u));
                recs.push_back(b);
            }
        }
        for (uint32_t v : touched)
            for (uint32_t bank = 0; bank < kBanks; ++bank)
                wmask[(size_t)v * kBanks + bank] = rmask[(size_t)v * kBanks + bank] = 0;
        touched.clear();
        k = e;
    }
    begin.push_back((uint32_t)(recs.size() / 2));
    return true;
}

int BuildSchedule(const xpbd_settings_desc &d, xpbd_settings &s, int32_t schedule)
{
    if (schedule < XPBD_SCHEDULE_EXACT || schedule > XPBD_SCHEDULE_COLOURED)
        return Fail(XPBD_E_INVALID, "unknown schedule");
    const uint32_t nv = d.num_vertices;
    if (nv == 0 || d.vertices == nullptr)
        return Fail(XPBD_E_INVALID, "settings have no vertices");
    if ((d.num_rod_stretch_shear && !d.rod_stretch_shear) || (d.num_rod_bend_twist && !d.rod_bend_twist))
        return Fail(XPBD_E_INVALID, "rod constraints without their arrays");
    if (d.num_rod_stretch_shear > kRodIndexMask || d.num_rod_bend_twist > kRodIndexMask)
        return Fail(XPBD_E_CAPACITY, "too many rod constraints");
    const RodSSRec *rods = (const RodSSRec *)d.rod_stretch_shear;
    const RodBTRec *bts = (const RodBTRec *)d.rod_bend_twist;
    for (uint32_t k = 0; k < d.num_rod_stretch_shear; ++k)
        if (rods[k].v[0] >= nv || rods[k].v[1] >= nv || rods[k].v[0] == rods[k].v[1])
            return Fail(XPBD_E_INVALID, "rod stretch shear constraint vertex out of range or repeated");
    for (uint32_t k = 0; k < d.num_rod_bend_twist; ++k)
        if (bts[k].rod[0] >= d.num_rod_stretch_shear || bts[k].rod[1] >= d.num_rod_stretch_shear || bts[k].rod[0] == bts[k].rod[1])
            return Fail(XPBD_E_INVALID, "rod bend twist constraint rod out of range or repeated");
    if (d.num_skinned && (!d.skinned || (d.num_inv_bind_matrices && !d.inv_bind_matrices)
                          || (d.num_skinned_normals && !d.skinned_normals)))
        return Fail(XPBD_E_INVALID, "skinned constraints without their arrays");
    if (d.num_update_groups == 0 || d.update_groups == nullptr)
        return Fail(XPBD_E_INVALID, "settings have no update groups (SoftBodySharedSettings::Optimize was not run)");

    const EdgeRec *edges = (const EdgeRec *)d.edges;
    const BendRec *bends = (const BendRec *)d.dihedral_bends;
    const VolumeRec *vols = (const VolumeRec *)d.volumes;
    const LRARec *lras = (const LRARec *)d.lras;
    const GroupRec *groups = (const GroupRec *)d.update_groups;
    const SkinnedRec *skinned = (const SkinnedRec *)d.skinned;
    if (d.num_skinned)
    {
        std::vector<char> seen(nv, 0);
        for (uint32_t k = 0; k < d.num_skinned; ++k)
        {
            const SkinnedRec &sr = skinned[k];
            if (sr.vertex >= nv || seen[sr.vertex])
                return Fail(XPBD_E_INVALID, "skinned constraint vertex out of range or skinned twice");
            seen[sr.vertex] = 1;
            for (int j = 0; j < 4 && sr.w[j].weight != 0.0f; ++j)
                if (sr.w[j].inv_bind >= d.num_inv_bind_matrices)
                    return Fail(XPBD_E_INVALID, "skin weight refers to a missing inverse bind matrix");
            const uint32_t nf = sr.normal_info >> 24, first = sr.normal_info & 0xffffffu;
            if (nf && ((uint64_t)first + nf > d.num_skinned_normals))
                return Fail(XPBD_E_INVALID, "skinned constraint normal faces out of range");
        }
        for (uint32_t k = 0; k < d.num_skinned_normals; ++k)
            if (d.skinned_normals[k] >= d.num_faces)
                return Fail(XPBD_E_INVALID, "skinned constraint normal face out of range");
        s.skinned.assign(skinned, skinned + d.num_skinned);
        s.skinned_normals.assign(d.skinned_normals, d.skinned_normals + d.num_skinned_normals);
        s.inv_binds.assign((const InvBindRec *)d.inv_bind_matrices, (const InvBindRec *)d.inv_bind_matrices + d.num_inv_bind_matrices);
        s.face_recs.assign((const FaceRec *)d.faces, (const FaceRec *)d.faces + d.num_faces);
    }

    GroupRec prev{ 0, 0, 0, 0, 0, 0, 0 };
    for (uint32_t g = 0; g < d.num_update_groups; ++g)
    {
        const GroupRec &c = groups[g];
        if (c.edge_end < prev.edge_end || c.lra_end < prev.lra_end || c.bend_end < prev.bend_end || c.volume_end < prev.volume_end
            || c.skinned_end < prev.skinned_end || c.rod_ss_end < prev.rod_ss_end || c.rod_bt_end < prev.rod_bt_end)
            return Fail(XPBD_E_INVALID, "update group end indices are not monotone");
        prev = c;
    }
    if (prev.edge_end != d.num_edges || prev.lra_end != d.num_lras || prev.bend_end != d.num_dihedral_bends || prev.volume_end != d.num_volumes
        || prev.skinned_end != d.num_skinned || prev.rod_ss_end != d.num_rod_stretch_shear || prev.rod_bt_end != d.num_rod_bend_twist)
        return Fail(XPBD_E_INVALID, "the last update group does not end at the constraint counts");

    auto check = [nv](const uint32_t *v, int n) {
        for (int i = 0; i < n; ++i)
            if (v[i] >= nv)
                return false;
        return true;
    };

    // Per group: sequence in ProcessGroup order, then levels
    const uint32_t ng = d.num_update_groups;
    std::vector<GroupLevels> gl(ng);
    std::vector<uint32_t> last_write(nv, 0), last_read(nv, 0), rod_last(d.num_rod_stretch_shear, 0);
    std::vector<uint64_t> colour_w, colour_r;
    if (schedule != XPBD_SCHEDULE_EXACT)
    {
        colour_w.assign((size_t)nv * kColourBanks, 0);
        colour_r.assign((size_t)nv * kColourBanks, 0);
    }
    prev = GroupRec{ 0, 0, 0, 0, 0, 0, 0 };
    for (uint32_t g = 0; g < ng; ++g)
    {
        const GroupRec &c = groups[g];
        std::vector<Item> &seq = gl[g].seq;
        for (uint32_t k = prev.volume_end; k < c.volume_end; ++k)
        {
            if (!check(vols[k].v, 4)) return Fail(XPBD_E_INVALID, "volume constraint vertex out of range");
            seq.push_back({ kTypeVolume, { vols[k].v[0], vols[k].v[1], vols[k].v[2], vols[k].v[3] }, 4, vols[k].six_rest_volume, vols[k].compliance, k });
        }
        for (uint32_t k = prev.bend_end; k < c.bend_end; ++k)
        {
            if (!check(bends[k].v, 4)) return Fail(XPBD_E_INVALID, "bend constraint vertex out of range");
            seq.push_back({ kTypeBend, { bends[k].v[0], bends[k].v[1], bends[k].v[2], bends[k].v[3] }, 4, bends[k].compliance, bends[k].initial_angle, k });
        }
        // ProcessGroup order: volumes, bends, skinned, edges, rod stretch-shear, rod bend-twist,
        // LRAs (SoftBodyMotionProperties.cpp:1037-1054)
        for (uint32_t k = prev.skinned_end; k < c.skinned_end; ++k)
        {
            const SkinnedRec &sr = skinned[k];
            seq.push_back({ kItemSkin, { sr.vertex, 0, 0, 0 }, 1, sr.max_distance, sr.back_stop_distance, k });
        }
        for (uint32_t k = prev.edge_end; k < c.edge_end; ++k)
        {
            if (!check(edges[k].v, 2)) return Fail(XPBD_E_INVALID, "edge constraint vertex out of range");
            seq.push_back({ kTypeEdge, { edges[k].v[0], edges[k].v[1], 0, 0 }, 2, edges[k].rest, edges[k].compliance, k });
        }
        for (uint32_t k = prev.rod_ss_end; k < c.rod_ss_end; ++k)
            seq.push_back({ kItemRodSS, { rods[k].v[0], rods[k].v[1], 0, 0 }, 2, 0.0f, 0.0f, k, { k, 0 }, 1 });
        for (uint32_t k = prev.rod_bt_end; k < c.rod_bt_end; ++k)
            seq.push_back({ kItemRodBT, { 0, 0, 0, 0 }, 0, 0.0f, 0.0f, k, { bts[k].rod[0], bts[k].rod[1] }, 2 });
        for (uint32_t k = prev.lra_end; k < c.lra_end; ++k)
        {
            if (!check(lras[k].v, 2)) return Fail(XPBD_E_INVALID, "LRA constraint vertex out of range");
            if (lras[k].v[0] == lras[k].v[1])
                continue;       // an LRA of a vertex to itself never moves it (that encoding marks skin records)
            seq.push_back({ kTypeLRA, { lras[k].v[0], lras[k].v[1], 0, 0 }, 2, lras[k].max_distance, 0.0f, k });
        }
        const bool coloured = schedule == XPBD_SCHEDULE_COLOURED || (schedule == XPBD_SCHEDULE_COLOURED_TAIL && g + 1 == ng);
        if (!coloured || !ColourLevels(seq, colour_w, colour_r, gl[g].levels))
            gl[g].levels = LevelGroup(seq, last_write, last_read, rod_last);
        prev = c;
    }
    const uint32_t npar = ng - 1;
    const GroupLevels &tail = gl[npar];

    uint32_t max_par_levels = 0;
    for (uint32_t g = 0; g < npar; ++g)
        max_par_levels = std::max(max_par_levels, (uint32_t)gl[g].levels.size());

    using LevelList = std::vector<std::vector<std::pair<const GroupLevels *, uint32_t>>>;
    ProgramBuilder pb;
    const bool resident = nv <= kResidentMaxVertices;
    uint32_t window_capacity = 0;
    // One identity window with the merged levels of all parallel groups, one for the tail
    auto build_identity = [&](ProgramBuilder &out) {
        LevelList par(max_par_levels);
        for (uint32_t g = 0; g < npar; ++g)
            for (uint32_t L = 0; L < gl[g].levels.size(); ++L)
                for (uint32_t k : gl[g].levels[L])
                    par[L].push_back({ &gl[g], k });
        std::vector<uint32_t> none;
        if (npar > 0)
            out.AddWindow(par, nullptr, none);
        LevelList tl(tail.levels.size());
        for (uint32_t L = 0; L < tail.levels.size(); ++L)
            for (uint32_t k : tail.levels[L])
                tl[L].push_back({ &tail, k });
        out.AddWindow(tl, nullptr, none);
    };
    if (resident)
    {
        build_identity(pb);
        s.dev.num_par_windows = npar > 0 ? 1u : 0u;
        s.dev.num_tail_windows = 1;
    }
    else
    {
        // Gathered windows of at most kWindowCapacity vertices
        std::vector<uint32_t> local(nv, 0xffffffffu);
        auto vertices_of = [](const GroupLevels &G, const std::vector<uint32_t> &items, std::vector<uint32_t> &out, std::vector<uint32_t> &mark, uint32_t stamp) {
            for (uint32_t k : items)
                for (uint32_t j = 0; j < G.seq[k].nv; ++j)
                {
                    uint32_t v = G.seq[k].v[j];
                    if (mark[v] != stamp) { mark[v] = stamp; out.push_back(v); }
                }
        };
        std::vector<uint32_t> mark(nv, 0);
        uint32_t stamp = 0;
        // parallel groups: whole groups per window
        uint32_t g = 0;
        uint32_t npw = 0;
        while (g < npar)
        {
            std::vector<uint32_t> verts;
            ++stamp;
            uint32_t g_end = g;
            while (g_end < npar)
            {
                std::vector<uint32_t> gv;
                for (const auto &lv : gl[g_end].levels)
                    vertices_of(gl[g_end], lv, gv, mark, stamp);
                if (!verts.empty() && verts.size() + gv.size() > kWindowCapacity)
                {
                    for (uint32_t v : gv) mark[v] = 0;   // not in this window
                    break;
                }
                if (gv.size() > kWindowCapacity)
                    return Fail(XPBD_E_CAPACITY, "an update group touches more vertices than a window holds");
                verts.insert(verts.end(), gv.begin(), gv.end());
                ++g_end;
            }
            std::sort(verts.begin(), verts.end());
            for (uint32_t i = 0; i < verts.size(); ++i) local[verts[i]] = i;
            uint32_t nl = 0;
            for (uint32_t h = g; h < g_end; ++h) nl = std::max(nl, (uint32_t)gl[h].levels.size());
            LevelList par(nl);
            for (uint32_t h = g; h < g_end; ++h)
                for (uint32_t L = 0; L < gl[h].levels.size(); ++L)
                    for (uint32_t k : gl[h].levels[L])
                        par[L].push_back({ &gl[h], k });
            pb.AddWindow(par, local.data(), verts);
            window_capacity = std::max(window_capacity, (uint32_t)verts.size());
            ++npw;
            g = g_end;
        }
        s.dev.num_par_windows = npw;
        // tail: consecutive levels per window.  A coloured level may touch more vertices than a
        // window holds: its constraints are independent, so it is cut into consecutive levels
        // that fit.
        std::vector<uint32_t> fresh_mark(nv, 0);
        uint32_t fresh_stamp = 0;
        if (schedule != XPBD_SCHEDULE_EXACT && !tail.levels.empty())
        {
            std::vector<std::vector<uint32_t>> cut;
            for (const std::vector<uint32_t> &lv : tail.levels)
            {
                cut.emplace_back();
                uint32_t count = 0;
                ++fresh_stamp;
                for (uint32_t k : lv)
                {
                    uint32_t add = 0;
                    for (uint32_t j = 0; j < tail.seq[k].nv; ++j)
                        add += fresh_mark[tail.seq[k].v[j]] != fresh_stamp ? 1u : 0u;
                    if (count + add > kWindowCapacity)
                    {
                        cut.emplace_back();
                        count = 0;
                        ++fresh_stamp;
                        add = tail.seq[k].nv;
                    }
                    for (uint32_t j = 0; j < tail.seq[k].nv; ++j)
                        fresh_mark[tail.seq[k].v[j]] = fresh_stamp;
                    count += add;
                    cut.back().push_back(k);
                }
            }
            gl[npar].levels = std::move(cut);
        }
        uint32_t ntw = 0;
        uint32_t L = 0;
        const uint32_t ntl = (uint32_t)tail.levels.size();
        do
        {
            std::vector<uint32_t> verts;
            ++stamp;
            uint32_t L_end = L;
            while (L_end < ntl)
            {
                uint32_t sub = stamp;
                // count the new vertices of this level without committing them
                std::vector<uint32_t> fresh;
                ++fresh_stamp;
                for (uint32_t k : tail.levels[L_end])
                    for (uint32_t j = 0; j < tail.seq[k].nv; ++j)
                    {
                        uint32_t v = tail.seq[k].v[j];
                        if (mark[v] != sub && fresh_mark[v] != fresh_stamp)
                        {
                            fresh_mark[v] = fresh_stamp;
                            fresh.push_back(v);
                        }
                    }
                if (!verts.empty() && verts.size() + fresh.size() > kWindowCapacity)
                    break;
                if (fresh.size() > kWindowCapacity)
                    return Fail(XPBD_E_CAPACITY, "a level of the non-parallel group touches more vertices than a window holds");
                for (uint32_t v : fresh) { mark[v] = sub; verts.push_back(v); }
                ++L_end;
            }
            std::sort(verts.begin(), verts.end());
            for (uint32_t i = 0; i < verts.size(); ++i) local[verts[i]] = i;
            LevelList tl(L_end - L);
            for (uint32_t q = L; q < L_end; ++q)
                for (uint32_t k : tail.levels[q])
                    tl[q - L].push_back({ &tail, k });
            pb.AddWindow(tl, local.data(), verts);
            window_capacity = std::max(window_capacity, (uint32_t)verts.size());
            ++ntw;
            L = L_end;
        } while (L < ntl);
        s.dev.num_tail_windows = ntw;
        s.par_windows.assign(pb.windows.begin(), pb.windows.begin() + npw);
        std::vector<uint4> crecs;
        std::vector<uint32_t> cbegin;
        if (ColourGroup(tail.seq, nv, crecs, cbegin))
        {
            hipError_t e;
            if (s.host_schedule)
            {
                s.host_schedule->colour_recs = crecs;
                s.host_schedule->colour_begin = cbegin;
            }
            else if ((e = s.colour_recs.upload(crecs)) != hipSuccess || (e = s.colour_begin.upload(cbegin)) != hipSuccess)
                return Fail(XPBD_E_DEVICE, std::string("settings upload: ") + hipGetErrorString(e));
            s.num_colours = (uint32_t)cbegin.size() - 1u;
            s.colour_records = (uint32_t)(crecs.size() / 2);
            s.colourable = true;
        }
    }

    s.num_edges = d.num_edges;
    s.num_bends = d.num_dihedral_bends;
    s.num_volumes = d.num_volumes;
    s.num_lras = d.num_lras;
    // host copy of the identity program (bundles): every body whose vertex numbers
    // fit the records' 14 bits, streamed ones included (they join a bundle when
    // their positions fit in LDS at 12 bytes per vertex)
    if (nv < kMaxWindowVertices)
    {
        ProgramBuilder idp;
        const ProgramBuilder *src = &pb;
        if (!resident)
        {
            build_identity(idp);
            src = &idp;
        }
        s.host_recs = src->recs;
        s.host_rec_index = src->rec_index;
        s.host_level_begin.clear();
        for (const WindowDev &w : src->windows)
            for (uint32_t L = 0; L < w.nlev; ++L)
                s.host_level_begin.push_back(src->lvl[w.lbegin + L]);
        s.host_level_begin.push_back((uint32_t)src->recs.size());
    }

    s.stats[0] = npar;
    s.stats[1] = max_par_levels;
    s.stats[2] = (uint32_t)tail.levels.size();
    s.stats[3] = (uint32_t)tail.seq.size();
    s.stats[4] = (uint32_t)pb.recs.size();
    s.stats[5] = resident ? 1u : 0u;
    s.stats[6] = s.dev.num_par_windows;
    s.stats[7] = s.dev.num_tail_windows;

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
    if (s.host_schedule)
    {
        s.host_schedule->recs = pb.recs;
        s.host_schedule->rec_index = pb.rec_index;
        s.host_schedule->lvl = pb.lvl;
        s.host_schedule->windows = pb.windows;
        s.host_schedule->vlist = pb.vlist;
        s.dev.num_records = (uint32_t)pb.recs.size();
        s.dev.num_levels_total = (uint32_t)pb.lvl.size();
        s.dev.resident = resident ? 1u : 0u;
        s.dev.window_capacity = window_capacity;
        return XPBD_OK;
    }
    if (d.num_skinned)
    {
        const VertexRec *verts = (const VertexRec *)d.vertices;
        std::vector<SkinnedDev> sk(d.num_skinned);
        for (uint32_t k = 0; k < d.num_skinned; ++k)
        {
            const SkinnedRec &sr = skinned[k];
            SkinnedDev &o = sk[k];
            memcpy(o.bind, verts[sr.vertex].pos, sizeof(o.bind));
            o.vertex = sr.vertex;
            for (int j = 0; j < 4; ++j)
            {
                o.inv_bind[j] = sr.w[j].inv_bind;
                o.weight[j] = sr.w[j].weight;
            }
            o.normal_info = sr.normal_info;
            o.max_distance = sr.max_distance;
            o.pad[0] = o.pad[1] = 0;
        }
        std::vector<SkinInvBindDev> ib(d.num_inv_bind_matrices);
        for (uint32_t k = 0; k < d.num_inv_bind_matrices; ++k)
        {
            ib[k].joint = s.inv_binds[k].joint;
            ib[k].pad[0] = ib[k].pad[1] = ib[k].pad[2] = 0;
            memcpy(ib[k].m, s.inv_binds[k].m, sizeof(ib[k].m));
        }
        if ((e = s.d_skinned.upload(sk)) != hipSuccess || (e = s.d_inv_binds.upload(ib)) != hipSuccess
            || (e = s.d_skinned_normals.upload(s.skinned_normals)) != hipSuccess)
            return Fail(XPBD_E_DEVICE, std::string("settings upload: ") + hipGetErrorString(e));
    }
    if ((e = s.recs.upload(pb.recs)) != hipSuccess || (e = s.lvl.upload(pb.lvl)) != hipSuccess
        || (e = s.windows.upload(pb.windows)) != hipSuccess || (e = s.vlist.upload(pb.vlist)) != hipSuccess
        || (e = s.faces.upload(f3)) != hipSuccess || (e = s.vf_begin.upload(vf_begin)) != hipSuccess
        || (e = s.vf_list.upload(vf_list)) != hipSuccess)
        return Fail(XPBD_E_DEVICE, std::string("settings upload: ") + hipGetErrorString(e));
    s.num_vertices = nv;
    s.num_faces = d.num_faces;
    const VertexRec *vr = (const VertexRec *)d.vertices;
    s.vertices.assign(vr, vr + nv);
    s.uniform_mass = true;
    s.inv_mass = vr[0].inv_mass;
    for (uint32_t i = 1; i < nv && s.uniform_mass; ++i)
        s.uniform_mass = memcmp(&vr[i].inv_mass, &s.inv_mass, sizeof(float)) == 0;

    if (d.num_rod_stretch_shear)
    {
        std::vector<float4> rc(d.num_rod_stretch_shear), bc(2ull * d.num_rod_bend_twist);
        for (uint32_t k = 0; k < d.num_rod_stretch_shear; ++k)
            rc[k] = make_float4(rods[k].length, rods[k].inv_mass, rods[k].compliance, 0.0f);
        for (uint32_t k = 0; k < d.num_rod_bend_twist; ++k)
        {
            float r1, r2;
            memcpy(&r1, &bts[k].rod[0], 4);
            memcpy(&r2, &bts[k].rod[1], 4);
            bc[2 * k] = make_float4(bts[k].omega0[0], bts[k].omega0[1], bts[k].omega0[2], bts[k].omega0[3]);
            bc[2 * k + 1] = make_float4(r1, r2, bts[k].compliance, 0.0f);
        }
        if ((e = s.rod_const.upload(rc)) != hipSuccess || (bc.size() && (e = s.bt_const.upload(bc)) != hipSuccess))
            return Fail(XPBD_E_DEVICE, std::string("settings upload: ") + hipGetErrorString(e));
        s.rods.assign(rods, rods + d.num_rod_stretch_shear);
    }

    SettingsDev &sd = s.dev;
    sd.rod_const = s.rod_const.p;
    sd.bt_const = s.bt_const.p;
    sd.num_rods = d.num_rod_stretch_shear;
    sd.recs = s.recs.p;
    sd.lvl = s.lvl.p;
    sd.windows = s.windows.p;
    sd.vlist = s.vlist.p;
    sd.faces = s.faces.p;
    sd.vf_begin = s.vf_begin.p;
    sd.vf_list = s.vf_list.p;
    sd.num_vertices = nv;
    sd.num_faces = d.num_faces;
    sd.num_records = (uint32_t)pb.recs.size();
    sd.num_levels_total = (uint32_t)pb.lvl.size();
    sd.resident = resident ? 1u : 0u;
    sd.window_capacity = window_capacity;
    return XPBD_OK;
}

// ------------------------------------------------------------------ bundles
// Bodies whose positions fit in LDS are packed into bundles that share one
// workgroup.  Bodies are sorted by the length of their level sequence so a
// bundle holds bodies of similar depth, then packed greedily while the bundle
// fits the middle launch class.  The bundle's program merges the bodies' level
// sequences by index; inside a merged level the records are grouped by
// constraint type and each group is padded to whole wavefronts when that does
// not add a pass.
constexpr uint32_t kBundleMaxBodies = 64;

// Launch classes of the bundles.  A CU has 160 KB of LDS, so the classes hold
// 4, 2 and 1 workgroups per CU; their threads per workgroup keep the wavefronts
// of co-resident workgroups within what the kernel's registers allow (5 per
// SIMD at its ~86 VGPRs).  Bundles are packed to the middle class; single
// bodies too large for it take the largest.  A large update runs in two bundle
// launches plus the streamed bodies' launch, each on a hardware queue of its
// own; the smallest class only takes bundles whose programs fit in LDS with
// them (small scenes).
struct BundleClass { uint32_t lds; uint32_t block; double per_cu; };
constexpr int kNumBundleClasses = 3;
constexpr BundleClass kBundleClasses[kNumBundleClasses] = {
    { 40u * 1024u, 256u, 4.0 }, { 80u * 1024u, 512u, 2.0 }, { kMaxLdsBytes, 1024u, 1.0 } };

// Duration of one level pass (µs) of a workgroup of `block` threads, measured
// with per-workgroup timestamps on MI355X: the pass is a dependent chain of LDS
// reads, one projection and LDS writes, then a workgroup barrier.
double LevelPassMicroseconds(uint32_t block)
{
    return block >= 1024u ? 1.06 : block >= 512u ? 0.97 : 0.90;
}

// Level passes of a program for a workgroup of `block` threads: a level of n
// records takes ceil(n / block) passes.
double ProgramPasses(const std::vector<uint32_t> &lvl, uint32_t block)
{
    double p = 0.0;
    for (size_t k = 0; k + 1 < lvl.size(); ++k)
        p += (double)((lvl[k + 1] - lvl[k] + block - 1) / block);
    return p;
}

uint4 OffsetRecord(uint4 r, uint32_t voff, uint32_t skin_base)
{
    uint32_t type = (r.x >> 14) & 3u;
    if (type == kTypeLRA && (r.x & 0x3fffu) == (r.x >> 16))
    {
        // skin record: vertex 0 twice, y = the body vertex's skin entry
        const uint32_t v0 = (r.x & 0x3fffu) + voff;
        return make_uint4((v0 & 0x3fffu) | (type << 14) | (v0 << 16), r.y + skin_base, r.z, r.w);
    }
    uint32_t v0 = (r.x & 0x3fffu) + voff, v1 = (r.x >> 16) + voff;
    uint32_t v2 = (r.y & 0x3fffu) + voff, v3 = (r.y >> 16) + voff;
    uint4 o;
    o.x = (v0 & 0x3fffu) | (type << 14) | (v1 << 16);
    o.y = (v2 & 0xffffu) | (v3 << 16);
    o.z = r.z;
    o.w = r.w;
    return o;
}

void BuildBundleProgram(const xpbd_context *ctx, const BundleHost &bh, const std::vector<uint32_t> &voffs, uint32_t block,
                        std::vector<uint4> &recs, std::vector<uint32_t> &lvl)
{
    recs.clear();
    lvl.clear();
    size_t K = 0;
    for (uint32_t bi : bh.bodies)
        K = std::max(K, ctx->bodies[bi]->settings->host_level_begin.size() - 1);
    std::vector<uint4> bucket[4];
    for (size_t k = 0; k < K; ++k)
    {
        for (auto &b : bucket) b.clear();
        for (size_t j = 0; j < bh.bodies.size(); ++j)
        {
            const xpbd_settings &st = *ctx->bodies[bh.bodies[j]]->settings;
            if (k + 1 >= st.host_level_begin.size())
                continue;
            for (uint32_t r = st.host_level_begin[k]; r < st.host_level_begin[k + 1]; ++r)
            {
                if (st.host_recs[r].x == kNopRecordHost)
                    continue;
                if (((st.host_recs[r].x >> 14) & 3u) == kTypeBend && (st.host_recs[r].y & kSecondHalf))
                    continue;       // re-emitted with its first half
                uint4 rec = OffsetRecord(st.host_recs[r], voffs[j], ctx->bodies[bh.bodies[j]]->dev.skin_base);
                bucket[(rec.x >> 14) & 3u].push_back(rec);
            }
        }
        lvl.push_back((uint32_t)recs.size());
        AppendTypedLevel(bucket, recs, true, nullptr, nullptr, block);
    }
    lvl.push_back((uint32_t)recs.size());
}

// Estimated duration (µs per collision step) of a streamed body: its gathered
// windows cost about one level pass each for the gather and the scatter.
double StreamedMicroseconds(const Body &b)
{
    const SettingsDev &sd = b.settings->dev;
    return (double)(sd.num_levels_total + 2u * (sd.num_par_windows + sd.num_tail_windows)) * (double)b.dev.iterations
           * LevelPassMicroseconds(1024u);
}

// Teams.  A streamed body without rods, pressure, dynamic or sensor colliders runs on a
// team of workgroups when it has at least team_min_windows parallel-group windows, or
// when it asks for the coloured non-parallel group.  One member per window (the
// windows' levels then run concurrently), at most kMaxTeamBlocks workgroups in all, so
// that every member of every team holds a CU at once; each member takes a contiguous
// run of windows of about equal cost.
constexpr uint32_t kMaxTeamBlocks = 224;
constexpr uint32_t kMaxTeamSize = 224;

int BuildTeams(xpbd_context *ctx)
{
    ctx->teams.clear();
    ctx->team_blocks.clear();
    std::vector<uint32_t> keep, members;
    for (uint32_t i : ctx->streamed)
    {
        const Body &b = *ctx->bodies[i];
        const xpbd_settings &st = *b.settings;
        const bool eligible = !st.dev.resident && st.rods.empty() && !(b.dev.pressure > 0.0f && st.num_faces > 0)
                              && ctx->num_dynamic == 0 && ctx->num_sensors == 0;
        const bool coloured = b.tail_mode == XPBD_TAIL_COLOURED;
        if (coloured && !(eligible && st.colourable))
            return Fail(XPBD_E_UNSUPPORTED, "body " + std::to_string(i) + ": the coloured non-parallel group needs a streamed body "
                                            "(over 8192 vertices) without rods, pressure, skinned constraints in that group, "
                                            "dynamic or sensor colliders");
        if (eligible && (coloured || (ctx->team_min_windows > 0 && st.dev.num_par_windows >= ctx->team_min_windows)))
        {
            ctx->teams.push_back(TeamDev{});
            ctx->teams.back().body = i;
            // one member per window; a coloured body's colour passes take twice that by default
            const uint32_t want = ctx->team_size ? ctx->team_size : (coloured ? 2u : 1u) * st.dev.num_par_windows;
            members.push_back(std::max(1u, std::min(std::max(st.dev.num_par_windows, want), kMaxTeamSize)));
        }
        else
            keep.push_back(i);
    }
    ctx->streamed = keep;
    if (ctx->teams.empty())
        return XPBD_OK;
    uint32_t total = 0;
    for (uint32_t m : members) total += m;
    while (total > kMaxTeamBlocks)
    {
        uint32_t &largest = *std::max_element(members.begin(), members.end());
        if (largest == 1)
            return Fail(XPBD_E_CAPACITY, "more large bodies than teams can hold at once");
        --largest;
        --total;
    }
    const uint32_t nc = (uint32_t)ctx->colliders.size();
    std::vector<uint32_t> split;
    uint32_t capacity = 0;
    for (size_t k = 0; k < ctx->teams.size(); ++k)
    {
        TeamDev &t = ctx->teams[k];
        const Body &b = *ctx->bodies[t.body];
        const std::vector<WindowDev> &pw = b.settings->par_windows;
        t.size = members[k];
        capacity = std::max(capacity, b.settings->dev.window_capacity);
        // contiguous runs of windows, cost = levels + gather + scatter
        double all = 0.0;
        for (const WindowDev &w : pw) all += w.nlev + 2.0;
        const size_t first = split.size();
        split.push_back(0);
        double acc = 0.0;
        uint32_t w = 0;
        const uint32_t with_windows = std::max(1u, std::min(t.size, (uint32_t)pw.size()));
        for (uint32_t m = 1; m < t.size; ++m)
        {
            if (m >= with_windows)
            {
                split.push_back((uint32_t)pw.size());
                continue;
            }
            const double goal = all * m / with_windows;
            while (w < pw.size() && acc + 0.5 * (pw[w].nlev + 2.0) < goal)
                acc += pw[w++].nlev + 2.0;
            split.push_back(w);
        }
        split.push_back((uint32_t)pw.size());
        t.par_split = reinterpret_cast<const uint32_t *>(first);       // offset until uploaded
        for (uint32_t m = 0; m < t.size; ++m)
            ctx->team_blocks.push_back(make_uint2((uint32_t)k, m));
        if (b.tail_mode == XPBD_TAIL_COLOURED)
        {
            t.colour_recs = b.settings->colour_recs.p;
            t.colour_begin = b.settings->colour_begin.p;
            t.num_colours = b.settings->num_colours;
        }
    }
    const uint32_t no_leaves = team_lds_bytes(capacity, 0, nc);
    if (no_leaves > kMaxLdsBytes)
        return Fail(XPBD_E_CAPACITY, "a team member's window needs more LDS than a workgroup has");
    ctx->team_max_leaves = std::min(nc, (kMaxLdsBytes - no_leaves) / bundle_leaf_bytes());
    ctx->team_lds = team_lds_bytes(capacity, ctx->team_max_leaves, nc);
    HIP_TRY(ctx->d_team_split.upload(split));
    HIP_TRY(ctx->d_team_sync.alloc(64 * ctx->teams.size()));
    size_t red = 0;
    for (const TeamDev &t : ctx->teams) red += 32u * (t.size + 1u);
    HIP_TRY(ctx->d_team_red.alloc(red));
    HIP_TRY(ctx->d_team_leaves.alloc((size_t)std::max(1u, nc) * ctx->team_blocks.size() * bundle_leaf_bytes()));
    size_t member0 = 0;
    red = 0;
    for (size_t k = 0; k < ctx->teams.size(); ++k)
    {
        TeamDev &t = ctx->teams[k];
        t.par_split = ctx->d_team_split.p + reinterpret_cast<size_t>(t.par_split);
        t.sync = ctx->d_team_sync.p + 64 * k;
        t.red = ctx->d_team_red.p + red;
        red += 32u * (t.size + 1u);
        t.leaves = reinterpret_cast<Leaf *>(ctx->d_team_leaves.p + member0 * std::max(1u, nc) * bundle_leaf_bytes());
        member0 += t.size;
    }
    HIP_TRY(ctx->d_teams.upload(ctx->teams));
    HIP_TRY(ctx->d_team_blocks.upload(ctx->team_blocks));
    return XPBD_OK;
}

int BuildLayout(xpbd_context *ctx)
{
    const uint32_t nc = (uint32_t)ctx->colliders.size();
    // A body joins the bundles when its identity program exists and its positions
    // fit in one workgroup's LDS: 12 B per vertex when all its vertices share one
    // inverse mass, else 16 B.  Bundles never mix inverse-mass classes.
    auto uniform_of = [ctx](uint32_t i) { return ctx->bodies[i]->settings->uniform_mass; };
    auto mass_bits = [ctx](uint32_t i) {
        uint32_t u;
        memcpy(&u, &ctx->bodies[i]->settings->inv_mass, 4);
        return u;
    };
    std::vector<uint32_t> resident;
    ctx->streamed.clear();
    double critical = 0.0;      // longest estimated duration of a single workgroup (µs per step)
    for (uint32_t i = 0; i < ctx->bodies.size(); ++i)
    {
        const Body &b = *ctx->bodies[i];
        const bool fits = bundle_lds_bytes(b.nv, 0, 0, 1, b.settings->uniform_mass) <= kMaxLdsBytes;
        // rod states stay in global memory: bodies with rods run in the per-body kernel
        if (ctx->use_bundles && fits && !b.settings->host_level_begin.empty() && b.settings->rods.empty())
            resident.push_back(i);
        else
        {
            ctx->streamed.push_back(i);
            critical = std::max(critical, StreamedMicroseconds(b));
        }
    }
    if (int rc = BuildTeams(ctx))
        return rc;
    // same iteration count and inverse-mass class together, deepest level sequences first
    std::stable_sort(resident.begin(), resident.end(), [&](uint32_t a, uint32_t b) {
        const Body &A = *ctx->bodies[a], &B = *ctx->bodies[b];
        if (A.dev.iterations != B.dev.iterations)
            return A.dev.iterations < B.dev.iterations;
        if (uniform_of(a) != uniform_of(b))
            return uniform_of(a);
        if (uniform_of(a) && mass_bits(a) != mass_bits(b))
            return mass_bits(a) < mass_bits(b);
        return A.settings->host_level_begin.size() > B.settings->host_level_begin.size();
    });
    std::vector<std::unique_ptr<BundleHost>> bundles;
    const uint32_t pack_lds = kBundleClasses[1].lds;
    for (uint32_t bi : resident)
    {
        const Body &b = *ctx->bodies[bi];
        BundleHost *cur = bundles.empty() ? nullptr : bundles.back().get();
        const bool same_class = cur && uniform_of(cur->bodies[0]) == uniform_of(bi)
                                && (!uniform_of(bi) || mass_bits(cur->bodies[0]) == mass_bits(bi))
                                && ctx->bodies[cur->bodies[0]]->dev.iterations == b.dev.iterations;
        const bool fits = same_class && cur->bodies.size() < kBundleMaxBodies
                          && bundle_lds_bytes(cur->nv + b.nv, 0, 0, (uint32_t)cur->bodies.size() + 1u, uniform_of(bi)) <= pack_lds;
        if (!fits)
        {
            bundles.emplace_back(new BundleHost);
            cur = bundles.back().get();
            cur->uniform = uniform_of(bi);
            cur->inv_mass = b.settings->inv_mass;
        }
        cur->bodies.push_back(bi);
        cur->nv += b.nv;
    }

    // Programs and launch classes.  Each bundle goes to the class that spends the
    // fewest CU-microseconds on it (duration / workgroups per CU) without making
    // it longer than the longest workgroup the update has anyway.
    struct Plan
    {
        std::vector<uint32_t> voffs;
        std::vector<uint4> recs;
        std::vector<uint32_t> lvl;
        double passes[3];       // level passes per iteration at 256, 512, 1024 threads
        uint32_t plain_lds, staged_lds, iterations, foff;
        bool has_pressure, has_sleepers;
        int cls;
        double us;              // estimated µs per collision step in its class
    };
    std::vector<Plan> plans(bundles.size());
    const uint32_t blocks[3] = { 256u, 512u, 1024u };
    for (size_t k = 0; k < bundles.size(); ++k)
    {
        BundleHost &bh = *bundles[k];
        Plan &P = plans[k];
        const uint32_t nb = (uint32_t)bh.bodies.size();
        P.voffs.resize(nb);
        uint32_t voff = 0, foff = 0;
        P.has_pressure = P.has_sleepers = false;
        for (uint32_t j = 0; j < nb; ++j)
        {
            const Body &b = *ctx->bodies[bh.bodies[j]];
            P.voffs[j] = voff;
            voff += b.nv;
            if (b.dev.pressure > 0.0f && b.settings->num_faces > 0)
            {
                foff += b.settings->num_faces;
                P.has_pressure = true;
            }
            P.has_sleepers |= b.dev.allow_sleeping != 0;
        }
        if (voff >= kMaxWindowVertices)
            return Fail(XPBD_E_CAPACITY, "bundle exceeds the LDS window index range");
        bh.nv = voff;
        P.foff = foff;
        P.iterations = ctx->bodies[bh.bodies[0]]->dev.iterations;
        BuildBundleProgram(ctx, bh, P.voffs, 256u, P.recs, P.lvl);
        for (int q = 0; q < 3; ++q)
            P.passes[q] = ProgramPasses(P.lvl, blocks[q]);
        const uint32_t nlev = (uint32_t)P.lvl.size() - 1;
        P.plain_lds = bundle_lds_bytes(voff, 0, 0, nb, bh.uniform);
        P.staged_lds = bundle_lds_bytes(voff, (uint32_t)P.recs.size(), nlev + 1, nb, bh.uniform);
        if (P.plain_lds > kMaxLdsBytes)
            return Fail(XPBD_E_CAPACITY, "bundle needs more LDS than a workgroup has");
        // shortest duration any class can give this bundle
        double best = 1e300;
        for (int q = 0; q < 3; ++q)
            best = std::min(best, P.iterations * P.passes[q] * LevelPassMicroseconds(blocks[q]));
        critical = std::max(critical, best);
    }
    auto class_block = [&](int c) { return kBundleClasses[c].block; };
    auto class_us = [&](const Plan &P, int c) {
        const uint32_t blk = class_block(c);
        return P.iterations * P.passes[blk == 1024u ? 2 : blk == 512u ? 1 : 0] * LevelPassMicroseconds(blk);
    };
    for (Plan &P : plans)
    {
        // fewest CU-microseconds among the classes that keep the bundle within the
        // critical duration; the shortest class when none does
        P.cls = -1;
        int fastest = -1;
        double best_cost = 1e300, fastest_us = 1e300;
        for (int c = 0; c < kNumBundleClasses; ++c)
        {
            // the smallest class is for bundles whose records fit in LDS with them
            // (small programs): a fourth bundle launch would share a hardware queue
            if (P.plain_lds > kBundleClasses[c].lds || (c == 0 && P.staged_lds > kBundleClasses[0].lds))
                continue;
            const double us = class_us(P, c);
            if (us < fastest_us)
            {
                fastest = c;
                fastest_us = us;
            }
            const double cost = us / kBundleClasses[c].per_cu;
            if (us <= critical * 1.0001 && cost < best_cost)
            {
                P.cls = c;
                best_cost = cost;
            }
        }
        if (P.cls < 0)
            P.cls = fastest;
        P.us = class_us(P, P.cls);
    }

    HIP_TRY(ctx->d_elapsed.alloc(std::max<size_t>(1, bundles.size())));
    if (ctx->h_elapsed_n < bundles.size())
    {
        if (ctx->h_elapsed)
            HIP_TRY(hipHostFree(ctx->h_elapsed));
        ctx->h_elapsed = nullptr;
        ctx->h_elapsed_n = 0;
        HIP_TRY(hipHostMalloc((void **)&ctx->h_elapsed, sizeof(unsigned long long) * bundles.size(), hipHostMallocDefault));
        ctx->h_elapsed_n = bundles.size();
    }
    ctx->order_state = xpbd_context::kOrderMeasure;
    // one launch per class and LDS layout (staged records, 12 B positions): slot 4 class + 2 staged + uniform
    std::vector<std::unique_ptr<xpbd_context::BundleLaunch>> launches(4 * kNumBundleClasses);
    std::vector<std::vector<std::pair<double, BundleDev>>> launch_tables(4 * kNumBundleClasses);
    for (size_t k = 0; k < bundles.size(); ++k)
    {
        BundleHost &bh = *bundles[k];
        Plan &P = plans[k];
        const uint32_t nb = (uint32_t)bh.bodies.size();
        const uint32_t block = class_block(P.cls);
        if (block != 256u)
            BuildBundleProgram(ctx, bh, P.voffs, block, P.recs, P.lvl);
        std::vector<BundleBodyDev> table(nb);
        uint32_t foff = 0;
        for (uint32_t j = 0; j < nb; ++j)
        {
            const Body &b = *ctx->bodies[bh.bodies[j]];
            table[j] = { bh.bodies[j], P.voffs[j], b.nv, foff };
            if (b.dev.pressure > 0.0f && b.settings->num_faces > 0)
                foff += b.settings->num_faces;
        }
        const uint32_t voff = bh.nv;
        HIP_TRY(bh.recs.upload(P.recs));
        HIP_TRY(bh.lvl.upload(P.lvl));
        HIP_TRY(bh.table.upload(table));
        // bundle-contiguous vertex state, gathered from wherever each body's state lives now
        HIP_TRY(bh.x.alloc(voff)); HIP_TRY(bh.v.alloc(voff)); HIP_TRY(bh.prev.alloc(voff)); HIP_TRY(bh.plane.alloc(voff));
        HIP_TRY(bh.report.alloc(voff));
        HIP_TRY(bh.idx.alloc(voff)); HIP_TRY(bh.contact.alloc(voff));
        for (uint32_t j = 0; j < nb; ++j)
        {
            Body &b = *ctx->bodies[bh.bodies[j]];
            const size_t o = P.voffs[j], n = b.nv;
            HIP_TRY(hipMemcpy(bh.x.p + o, b.dev.x, n * sizeof(float4), hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(bh.v.p + o, b.dev.v, n * sizeof(float4), hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(bh.prev.p + o, b.dev.prev, n * sizeof(float4), hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(bh.plane.p + o, b.dev.plane, n * sizeof(float4), hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(bh.report.p + o, b.dev.report, n * sizeof(float4), hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(bh.idx.p + o, b.dev.idx, n * sizeof(int32_t), hipMemcpyDeviceToDevice));
            HIP_TRY(hipMemcpy(bh.contact.p + o, b.dev.contact, n * sizeof(int32_t), hipMemcpyDeviceToDevice));
        }
        HIP_TRY(bh.leaves_full.alloc((size_t)std::max(1u, nb * nc) * bundle_leaf_bytes()));
        HIP_TRY(bh.leaves.alloc((size_t)std::max(1u, nb * nc) * bundle_leaf_bytes()));
        HIP_TRY(bh.face_scratch.alloc(std::max(1u, P.foff)));
        BundleDev &d = bh.dev;
        d.recs = bh.recs.p; d.lvl = bh.lvl.p; d.nlev = (uint32_t)P.lvl.size() - 1; d.num_records = (uint32_t)P.recs.size();
        d.bodies = bh.table.p; d.num_bodies = nb; d.nv = voff;
        d.x = bh.x.p; d.v = bh.v.p; d.prev = bh.prev.p; d.plane = bh.plane.p; d.idx = bh.idx.p; d.contact = bh.contact.p;
        d.report = bh.report.p;
        d.leaves_full = bh.leaves_full.p; d.leaves = bh.leaves.p; d.face_scratch = bh.face_scratch.p;
        d.has_pressure = P.has_pressure ? 1u : 0u;
        d.has_sleepers = P.has_sleepers ? 1u : 0u;
        d.iterations = P.iterations;
        d.uniform = bh.uniform ? 1u : 0u;
        d.uniform_inv_mass = bh.uniform ? bh.inv_mass : 0.0f;
        d.elapsed = ctx->d_elapsed.p + k;
        d.skin = ctx->d_skin.p;
        // records staged in LDS when the bundle still fits the smallest class with them
        bh.staged = block == 256u && bundle_lds_bytes(voff, d.num_records, d.nlev + 1, nb, bh.uniform) <= kBundleClasses[0].lds;
        d.staged = bh.staged ? 1u : 0u;
        bh.lds = kBundleClasses[P.cls].lds;
        const int slot = 4 * P.cls + (bh.staged ? 2 : 0) + (bh.uniform ? 1 : 0);
        if (!launches[slot])
        {
            launches[slot].reset(new xpbd_context::BundleLaunch);
            launches[slot]->staged = bh.staged;
            launches[slot]->uniform = bh.uniform;
            launches[slot]->block = block;
            launches[slot]->lds = kBundleClasses[P.cls].lds;
        }
        launch_tables[slot].emplace_back(P.us, d);
        launches[slot]->cost = std::max(launches[slot]->cost, P.us);
        launches[slot]->work += P.us / kBundleClasses[P.cls].per_cu;
    }
    // bodies now live in bundle storage: repoint them and drop their own buffers
    for (auto &bp : bundles)
        for (uint32_t j = 0; j < bp->bodies.size(); ++j)
        {
            Body &b = *ctx->bodies[bp->bodies[j]];
            uint32_t o = 0;
            for (uint32_t k = 0; k < j; ++k) o += ctx->bodies[bp->bodies[k]]->nv;
            b.dev.x = bp->x.p + o; b.dev.v = bp->v.p + o; b.dev.prev = bp->prev.p + o; b.dev.plane = bp->plane.p + o;
            b.dev.idx = bp->idx.p + o; b.dev.contact = bp->contact.p + o; b.dev.report = bp->report.p + o;
            if (b.owns_storage)
            {
                (void)b.x.alloc(0); (void)b.v.alloc(0); (void)b.prev.alloc(0); (void)b.plane.alloc(0); (void)b.idx.alloc(0); (void)b.contact.alloc(0);
                (void)b.report.alloc(0);
                b.owns_storage = false;
            }
        }
    ctx->bundles.swap(bundles);      // old bundles (if any) are freed here, after the copies
    // Launch order, largest LDS first: a workgroup that needs a whole CU's LDS
    // starts before smaller workgroups take parts of every CU.  Inside a launch the
    // longest bundles come first (workgroups are dispatched in index order).
    ctx->bundle_launches.clear();
    for (int c = 4 * kNumBundleClasses - 1; c >= 0; --c)
    {
        if (!launches[c])
            continue;
        auto &tab = launch_tables[c];
        std::stable_sort(tab.begin(), tab.end(), [](const std::pair<double, BundleDev> &a, const std::pair<double, BundleDev> &b) {
            return a.first > b.first;
        });
        std::vector<BundleDev> &devs = launches[c]->host;
        for (auto &t : tab) devs.push_back(t.second);
        HIP_TRY(launches[c]->table.upload(devs));
        launches[c]->count = (uint32_t)devs.size();
        ctx->bundle_launches.push_back(std::move(launches[c]));
    }
    ctx->layout_dirty = false;
    ctx->layout_colliders = nc;
    ctx->tables_dirty = true;
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

int SyncTables(xpbd_context *ctx)
{
    if (!ctx->tables_dirty)
        return XPBD_OK;
    // keep the device-side body states: pull them before re-uploading the tables
    int rc = PullStates(ctx);
    if (rc != XPBD_OK)
        return rc;
    std::vector<BodyDev> devs;
    std::vector<BodyStateDev> states;
    for (auto &b : ctx->bodies)
    {
        devs.push_back(b->dev);
        states.push_back(b->state);
    }
    HIP_TRY(ctx->d_bodies.upload(devs));
    HIP_TRY(ctx->d_states.upload(states));
    ctx->tables_dirty = false;
    ctx->states_on_device = true;
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
    HIP_TRY(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ctx->elapsed_ready, hipEventDisableTiming));
    {
        const char *q = getenv("GPU_MAX_HW_QUEUES");
        const int queues = q && atoi(q) > 0 ? atoi(q) : 4;
        ctx->num_aux = std::max(0, std::min(xpbd_context::kMaxAux, queues - 1));
    }
    for (int a = 0; a < ctx->num_aux; ++a)
    {
        HIP_TRY(hipStreamCreateWithFlags(&ctx->aux[a], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ctx->join[a], hipEventDisableTiming));
    }
    {
        const char *e = getenv("XPBD_BUNDLES");
        ctx->use_bundles = !(e && e[0] == '0');
    }
    HIP_TRY(ctx->d_error.alloc(1));
    HIP_TRY(hipMemset(ctx->d_error.p, 0, sizeof(int32_t)));
    // the device's copy of the allocation-free rollout scratch (xpbd_rollout_forward_noalloc)
    HIP_TRY(rollout_scratch_setup(device));
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
    ctx->bundles.clear();
    ctx->bodies.clear();
    ctx->bundle_launches.clear();
    (void)hipEventDestroy(ctx->fork);
    (void)hipEventDestroy(ctx->elapsed_ready);
    if (ctx->h_elapsed)
        (void)hipHostFree(ctx->h_elapsed);
    for (int a = 0; a < ctx->num_aux; ++a)
    {
        (void)hipEventDestroy(ctx->join[a]);
        (void)hipStreamDestroy(ctx->aux[a]);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return XPBD_OK;
}

void *xpbd_context_stream(xpbd_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int xpbd_settings_create_scheduled(xpbd_context *ctx, const xpbd_settings_desc *desc, int32_t schedule, xpbd_settings **out_settings)
{
    if (!ctx || !desc || !out_settings)
        return Fail(XPBD_E_INVALID, "null argument");
    *out_settings = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<xpbd_settings> s(new xpbd_settings);
    s->ctx = ctx;
    int rc = BuildSchedule(*desc, *s, schedule);
    if (rc != XPBD_OK)
        return rc;
    *out_settings = s.release();
    return XPBD_OK;
}

int xpbd_settings_create(xpbd_context *ctx, const xpbd_settings_desc *desc, xpbd_settings **out_settings)
{
    return xpbd_settings_create_scheduled(ctx, desc, XPBD_SCHEDULE_EXACT, out_settings);
}

int xpbd_schedule_build_scheduled(const xpbd_settings_desc *desc, int32_t schedule, xpbd_settings **out)
{
    if (!desc || !out)
        return Fail(XPBD_E_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<xpbd_settings> s(new xpbd_settings);
    s->host_schedule.reset(new xpbd_settings::HostSchedule);
    int rc = BuildSchedule(*desc, *s, schedule);
    if (rc != XPBD_OK)
        return rc;
    *out = s.release();
    return XPBD_OK;
}

int xpbd_schedule_build(const xpbd_settings_desc *desc, xpbd_settings **out)
{
    return xpbd_schedule_build_scheduled(desc, XPBD_SCHEDULE_EXACT, out);
}

int xpbd_schedule_sizes(const xpbd_settings *s, uint32_t out[10])
{
    if (!s || !out || !s->host_schedule)
        return Fail(XPBD_E_INVALID, "not a host schedule");
    const auto &h = *s->host_schedule;
    out[0] = (uint32_t)h.recs.size();
    out[1] = (uint32_t)h.lvl.size();
    out[2] = (uint32_t)h.windows.size();
    out[3] = s->dev.num_par_windows;
    out[4] = (uint32_t)h.vlist.size();
    out[5] = (uint32_t)(h.colour_recs.size() / 2);
    out[6] = (uint32_t)h.colour_begin.size();
    out[7] = s->dev.resident;
    out[8] = (uint32_t)s->host_recs.size();
    out[9] = (uint32_t)s->host_level_begin.size();
    return XPBD_OK;
}

int xpbd_schedule_read(const xpbd_settings *s, uint32_t *recs, uint32_t *rec_index, uint32_t *lvl, uint32_t *windows,
                       uint32_t *vlist, uint32_t *colour_recs, uint32_t *colour_begin, uint32_t *identity_recs,
                       uint32_t *identity_rec_index, uint32_t *identity_level_begin)
{
    if (!s || !s->host_schedule)
        return Fail(XPBD_E_INVALID, "not a host schedule");
    const auto &h = *s->host_schedule;
    auto put = [](uint32_t *dst, const void *src, size_t bytes) { if (dst && bytes) memcpy(dst, src, bytes); };
    put(recs, h.recs.data(), h.recs.size() * sizeof(uint4));
    put(rec_index, h.rec_index.data(), h.rec_index.size() * 4);
    put(lvl, h.lvl.data(), h.lvl.size() * 4);
    put(windows, h.windows.data(), h.windows.size() * sizeof(WindowDev));
    put(vlist, h.vlist.data(), h.vlist.size() * 4);
    put(colour_recs, h.colour_recs.data(), h.colour_recs.size() * sizeof(uint4));
    put(colour_begin, h.colour_begin.data(), h.colour_begin.size() * 4);
    put(identity_recs, s->host_recs.data(), s->host_recs.size() * sizeof(uint4));
    put(identity_rec_index, s->host_rec_index.data(), s->host_rec_index.size() * 4);
    put(identity_level_begin, s->host_level_begin.data(), s->host_level_begin.size() * 4);
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

int xpbd_settings_stats(const xpbd_settings *s, uint32_t out[8])
{
    if (!s || !out)
        return Fail(XPBD_E_INVALID, "null argument");
    memcpy(out, s->stats, sizeof(s->stats));
    return XPBD_OK;
}

int xpbd_settings_team_stats(const xpbd_settings *s, uint32_t out[4])
{
    if (!s || !out)
        return Fail(XPBD_E_INVALID, "null argument");
    out[0] = s->dev.resident ? 0u : s->dev.num_par_windows;
    out[1] = s->dev.resident ? 0u : s->dev.num_tail_windows;
    out[2] = s->colourable ? 1u : 0u;
    out[3] = s->num_colours;
    return XPBD_OK;
}

int xpbd_context_set_team_size(xpbd_context *ctx, uint32_t workgroups)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    if (ctx->team_size != workgroups)
    {
        ctx->team_size = workgroups;
        ctx->layout_dirty = true;
    }
    return XPBD_OK;
}

int xpbd_context_set_team_min_windows(xpbd_context *ctx, uint32_t windows)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    if (ctx->team_min_windows != windows)
    {
        ctx->team_min_windows = windows;
        ctx->layout_dirty = true;
    }
    return XPBD_OK;
}

int xpbd_body_set_tail_mode(xpbd_context *ctx, int32_t body, int32_t mode)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    if (mode != XPBD_TAIL_EXACT && mode != XPBD_TAIL_COLOURED)
        return Fail(XPBD_E_INVALID, "tail mode must be XPBD_TAIL_EXACT or XPBD_TAIL_COLOURED");
    Body &b = *ctx->bodies[body];
    if (mode == XPBD_TAIL_COLOURED && (b.settings->dev.resident || !b.settings->colourable))
        return Fail(XPBD_E_UNSUPPORTED, "the coloured non-parallel group needs a streamed body (over 8192 vertices) whose "
                                        "non-parallel group holds only volumes, bends, edges and LRAs");
    if (b.tail_mode != mode)
    {
        b.tail_mode = mode;
        ctx->layout_dirty = true;
    }
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
    float no_contact = 0.0f;
    {
        const int32_t minus_one = -1;
        memcpy(&no_contact, &minus_one, sizeof(no_contact));
    }
    std::vector<float4> report(nv, make_float4(0.0f, 0.0f, 0.0f, no_contact));
    HIP_TRY(b->x.upload(x));
    HIP_TRY(b->v.upload(v));
    HIP_TRY(b->prev.upload(x));
    HIP_TRY(b->plane.upload(plane));
    HIP_TRY(b->idx.upload(idx));
    HIP_TRY(b->contact.upload(contact));
    HIP_TRY(b->report.upload(report));

    BodyDev &dv = b->dev;
    dv.s = s->dev;
    dv.x = b->x.p; dv.v = b->v.p; dv.prev = b->prev.p; dv.plane = b->plane.p; dv.idx = b->idx.p; dv.contact = b->contact.p;
    dv.report = b->report.p;
    if (!s->rods.empty())
    {
        // SoftBodyMotionProperties::Initialize (SoftBodyMotionProperties.cpp:83-95): rotation =
        // creation rotation (when made identity) * Bishop frame, angular velocity zero
        float rq[4];
        get_quaternion(rot, rq);
        const Q4 rotation_q = q4(rq[0], rq[1], rq[2], rq[3]);
        std::vector<float4> rs(2 * s->rods.size());
        for (size_t r = 0; r < s->rods.size(); ++r)
        {
            const float *bq = s->rods[r].bishop;
            const Q4 q = quat_mul(rotation_q, q4(bq[0], bq[1], bq[2], bq[3]));
            rs[2 * r] = make_float4(q.x, q.y, q.z, q.w);
            rs[2 * r + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        HIP_TRY(b->rods.upload(rs));
        dv.rods = b->rods.p;
    }
    if (!s->skinned.empty())
    {
        // SoftBodyMotionProperties::Initialize: skin states zero / NaN until SkinVertices
        const float nan = std::numeric_limits<float>::quiet_NaN();
        std::vector<float> radius(nv, 0.0f);
        for (const SkinnedRec &sr : s->skinned) radius[sr.vertex] = sr.back_stop_radius;
        b->skinned = true;
        dv.skin_base = (uint32_t)(ctx->skin_host.size() / 4);
        for (uint32_t i = 0; i < nv; ++i)
        {
            ctx->skin_host.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            ctx->skin_host.push_back(make_float4(nan, nan, nan, 0.0f));
            ctx->skin_host.push_back(make_float4(nan, nan, nan, 0.0f));
            ctx->skin_host.push_back(make_float4(1.0f, radius[i], 1.0f, 0.0f));
        }
        ctx->skin_dirty = true;
    }
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
    ctx->layout_dirty = true;
    *out_body = (int32_t)(ctx->bodies.size() - 1);
    return XPBD_OK;
}

namespace {
// A triangle tree's references stay inside its nodes and triangles, every node is reached
// once from the root (a tree, no cycles), and the depth is within the device walk's stack.
bool ValidateTree(const xpbd_triangle_tree &t, std::string &err)
{
    if (t.num_nodes >= (1u << 30) || t.num_triangles >= (1u << 24))
    {
        err = "too many nodes or triangles";
        return false;
    }
    std::vector<uint8_t> seen(t.num_nodes, 0);
    struct Item { uint32_t ref, depth; };
    std::vector<Item> todo{ { t.root, 0 } };
    while (!todo.empty())
    {
        const Item it = todo.back();
        todo.pop_back();
        const uint32_t kind = it.ref & 0xc0000000u;
        if (kind == XPBD_TREE_EMPTY)
            continue;
        if (kind == XPBD_TREE_LEAF)
        {
            const uint32_t first = it.ref & 0x00ffffffu, count = (it.ref >> 24) & 0x3fu;
            if (count == 0 || (uint64_t)first + count > t.num_triangles)
            {
                err = "leaf outside the triangle table";
                return false;
            }
            continue;
        }
        if (kind != XPBD_TREE_NODE || it.ref >= t.num_nodes)
        {
            err = "reference outside the node table";
            return false;
        }
        if (seen[it.ref])
        {
            err = "a node is reached twice";
            return false;
        }
        seen[it.ref] = 1;
        if (it.depth + 1 > XPBD_TREE_MAX_DEPTH)
        {
            err = "deeper than XPBD_TREE_MAX_DEPTH";
            return false;
        }
        for (int q = 0; q < 4; ++q)
            todo.push_back({ t.nodes[it.ref].child[q], it.depth + 1 });
    }
    return true;
}

// A tree reference moved into the pools, whose nodes and triangles start at the given offsets
uint32_t RebaseTreeRef(uint32_t ref, uint32_t node_base, uint32_t tri_base)
{
    const uint32_t kind = ref & 0xc0000000u;
    if (kind == XPBD_TREE_NODE)
        return ref + node_base;
    if (kind == XPBD_TREE_LEAF)
        return (ref & 0xff000000u) | ((ref & 0x00ffffffu) + tri_base);
    return ref;
}
} // namespace

int xpbd_set_colliders(xpbd_context *ctx, const xpbd_collider_desc *c, uint32_t n)
{
    if (!ctx || (n && !c))
        return Fail(XPBD_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<ColliderDev> v;
    v.reserve(n);
    std::vector<uint32_t> first(n);
    uint32_t num_dynamic = 0, num_sensors = 0;
    // convex hull tables: validated here, flattened into two pools, pointers set after upload
    std::vector<float4> hull_vec;
    std::vector<uint32_t> hull_idx;
    struct HullOffsets { uint32_t planes, points, face_first, face_count, index, num_faces, collider; };
    std::vector<HullOffsets> hull_offsets;
    // triangle trees: validated here, nodes and triangles flattened into two pools
    std::vector<TreeNodeDev> tree_nodes;
    std::vector<float> tree_tris;
    struct TreeOffsets { uint32_t nodes, tris, root, collider, flags; };
    std::vector<TreeOffsets> tree_offsets;
    // compound leaves' placement steps, one pool
    std::vector<PartOpDev> op_pool;
    std::vector<uint32_t> op_offset;        // per leaf
    // one leaf: the shape of a simple body, or a part of a compound one
    struct Shape { int32_t type; const float *params, *scale; const xpbd_convex_hull *hull; const xpbd_triangle_tree *tree;
                   uint32_t num_ops; const xpbd_part_op *ops; };
    auto add_leaf = [&](const ColliderDev &body, const Shape &sh, bool is_first, std::string &err) -> int {
        const uint32_t leaf = (uint32_t)v.size();
        if (sh.type < XPBD_SHAPE_PLANE || sh.type > XPBD_SHAPE_MESH)
        {
            err = "collider shape type not supported (plane, sphere, box, capsule, cylinder, tapered capsule, convex hull, "
                  "tapered cylinder, mesh only)";
            return XPBD_E_UNSUPPORTED;
        }
        const bool unit = sh.scale[0] == 0.0f && sh.scale[1] == 0.0f && sh.scale[2] == 0.0f;
        const float sc[3] = { unit ? 1.0f : sh.scale[0], unit ? 1.0f : sh.scale[1], unit ? 1.0f : sh.scale[2] };
        for (int k = 0; k < 3; ++k)
            if (!(std::fabs(sc[k]) >= 1.0e-6f) || !std::isfinite(sc[k]))
            {
                err = "collider scale components must be finite and at least 1e-6 in magnitude";
                return XPBD_E_INVALID;
            }
        if (sh.type == XPBD_SHAPE_MESH)
        {
            const xpbd_triangle_tree *t = sh.tree;
            if (!t || !t->triangles || (t->num_nodes && !t->nodes))
            {
                err = "mesh collider without tree data";
                return XPBD_E_INVALID;
            }
            std::string terr;
            if (!ValidateTree(*t, terr))
            {
                err = "mesh collider tree: " + terr;
                return XPBD_E_INVALID;
            }
            if (t->flags & ~XPBD_TREE_HEIGHT_FIELD)
            {
                err = "mesh collider tree: unknown flags";
                return XPBD_E_INVALID;
            }
            TreeOffsets o;
            o.nodes = (uint32_t)tree_nodes.size();
            o.tris = (uint32_t)tree_tris.size() / 9u;
            for (uint32_t k = 0; k < t->num_nodes; ++k)
            {
                TreeNodeDev d;
                static_assert(sizeof(TreeNodeDev) == sizeof(xpbd_tree_node), "tree node layouts");
                memcpy(&d, &t->nodes[k], sizeof(d));
                for (int q = 0; q < 4; ++q)
                    d.child[q] = RebaseTreeRef(d.child[q], o.nodes, o.tris);
                tree_nodes.push_back(d);
            }
            tree_tris.insert(tree_tris.end(), t->triangles, t->triangles + 9 * (size_t)t->num_triangles);
            o.root = RebaseTreeRef(t->root, o.nodes, o.tris);
            o.collider = leaf;
            o.flags = t->flags;
            tree_offsets.push_back(o);
        }
        if (sh.type == XPBD_SHAPE_CONVEX_HULL)
        {
            const xpbd_convex_hull *h = sh.hull;
            if (!h || !h->points || !h->planes || !h->face_first || !h->face_count || !h->vertex_index)
            {
                err = "convex hull collider without hull data";
                return XPBD_E_INVALID;
            }
            if (h->num_faces == 0 || h->num_points == 0)
            {
                err = "convex hull without faces or points";
                return XPBD_E_INVALID;
            }
            for (uint32_t f = 0; f < h->num_faces; ++f)
            {
                const uint64_t end = (uint64_t)h->face_first[f] + h->face_count[f];
                if (h->face_count[f] < 2 || end > h->num_vertex_index)
                {
                    err = "convex hull face outside the vertex index table";
                    return XPBD_E_INVALID;
                }
            }
            for (uint32_t k = 0; k < h->num_vertex_index; ++k)
                if (h->vertex_index[k] >= h->num_points)
                {
                    err = "convex hull vertex index out of range";
                    return XPBD_E_INVALID;
                }
            HullOffsets o;
            o.planes = (uint32_t)hull_vec.size();
            for (uint32_t f = 0; f < h->num_faces; ++f)
                hull_vec.push_back(make_float4(h->planes[4 * f], h->planes[4 * f + 1], h->planes[4 * f + 2], h->planes[4 * f + 3]));
            o.points = (uint32_t)hull_vec.size();
            for (uint32_t k = 0; k < h->num_points; ++k)
                hull_vec.push_back(make_float4(h->points[3 * k], h->points[3 * k + 1], h->points[3 * k + 2], 0.0f));
            o.face_first = (uint32_t)hull_idx.size();
            hull_idx.insert(hull_idx.end(), h->face_first, h->face_first + h->num_faces);
            o.face_count = (uint32_t)hull_idx.size();
            hull_idx.insert(hull_idx.end(), h->face_count, h->face_count + h->num_faces);
            o.index = (uint32_t)hull_idx.size();
            hull_idx.insert(hull_idx.end(), h->vertex_index, h->vertex_index + h->num_vertex_index);
            o.num_faces = h->num_faces;
            o.collider = leaf;
            hull_offsets.push_back(o);
        }
        if (sh.num_ops > XPBD_MAX_PART_OPS || (sh.num_ops && !sh.ops))
        {
            err = "collider part: more than XPBD_MAX_PART_OPS placement steps, or none given";
            return XPBD_E_INVALID;
        }
        op_offset.push_back((uint32_t)op_pool.size());
        for (uint32_t k = 0; k < sh.num_ops; ++k)
        {
            const xpbd_part_op &op = sh.ops[k];
            if (op.kind < XPBD_OP_CULL || op.kind > XPBD_OP_OFFSET)
            {
                err = "collider part: unknown placement step";
                return XPBD_E_INVALID;
            }
            PartOpDev d;
            static_assert(sizeof(PartOpDev) == sizeof(xpbd_part_op), "placement step layouts");
            memcpy(&d, &op, sizeof(d));
            op_pool.push_back(d);
        }
        ColliderDev d = body;
        d.type = sh.type;
        d.hull = nullptr;
        d.tree = nullptr;
        memcpy(d.scale, sc, sizeof(d.scale));
        memcpy(d.params, sh.params, sizeof(d.params));
        d.first = is_first ? 1 : 0;
        d.num_ops = sh.num_ops;
        d.ops = nullptr;
        v.push_back(d);
        return XPBD_OK;
    };
    for (uint32_t i = 0; i < n; ++i)
    {
        if (c[i].motion_type < 0 || c[i].motion_type > 2)
            return Fail(XPBD_E_INVALID, "collider motion type must be 0 (static), 1 (kinematic) or 2 (dynamic)");
        if (c[i].num_parts && !c[i].parts)
            return Fail(XPBD_E_INVALID, "compound collider without parts");
        // the body's own fields, shared by its leaves
        ColliderDev body{};
        body.motion = c[i].motion_type;
        body.dyn = -1;
        body.sensor = c[i].is_sensor ? (int32_t)num_sensors++ : -1;
        if (c[i].motion_type == 2)
        {
            body.dyn = (int32_t)num_dynamic++;
            body.inv_mass = c[i].inv_mass;
            memcpy(body.inv_inertia_diag, c[i].inv_inertia_diagonal, sizeof(body.inv_inertia_diag));
            memcpy(body.inertia_rot, c[i].inertia_rotation, sizeof(body.inertia_rot));
            memcpy(body.lin_vel, c[i].linear_velocity, sizeof(body.lin_vel));
            memcpy(body.ang_vel, c[i].angular_velocity, sizeof(body.ang_vel));
        }
        memcpy(body.pos, c[i].position, sizeof(body.pos));
        memcpy(body.rot, c[i].rotation, sizeof(body.rot));
        body.friction = c[i].friction;
        body.restitution = c[i].restitution;
        memcpy(body.bounds, c[i].world_bounds, sizeof(body.bounds));
        body.owner = (int32_t)i;
        first[i] = (uint32_t)v.size();
        std::string err;
        if (c[i].num_parts == 0)
        {
            const Shape sh{ c[i].type, c[i].params, c[i].scale, c[i].hull, c[i].tree, 0u, nullptr };
            if (int rc = add_leaf(body, sh, true, err))
                return Fail(rc, err);
        }
        else
            for (uint32_t k = 0; k < c[i].num_parts; ++k)
            {
                const xpbd_collider_part &pt = c[i].parts[k];
                const Shape sh{ pt.type, pt.params, pt.scale, pt.hull, pt.tree, pt.num_ops, pt.ops };
                if (int rc = add_leaf(body, sh, k == 0, err))
                    return Fail(rc, err);
            }
    }
    if (tree_nodes.size() >= (1u << 30) || tree_tris.size() / 9u >= (1u << 24))
        return Fail(XPBD_E_CAPACITY, "the mesh colliders hold more than 2^30 nodes or 2^24 triangles together");
    if (v.size() >= (1u << 24))
        return Fail(XPBD_E_CAPACITY, "more than 2^24 collider leaf shapes");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx->d_hull_vec.upload(hull_vec));
    HIP_TRY(ctx->d_hull_idx.upload(hull_idx));
    std::vector<HullDev> hulls(hull_offsets.size());
    for (size_t k = 0; k < hull_offsets.size(); ++k)
    {
        const HullOffsets &o = hull_offsets[k];
        hulls[k].planes = ctx->d_hull_vec.p + o.planes;
        hulls[k].points = ctx->d_hull_vec.p + o.points;
        hulls[k].face_first = ctx->d_hull_idx.p + o.face_first;
        hulls[k].face_count = ctx->d_hull_idx.p + o.face_count;
        hulls[k].index = ctx->d_hull_idx.p + o.index;
        hulls[k].num_faces = o.num_faces;
    }
    HIP_TRY(ctx->d_hulls.upload(hulls));
    for (size_t k = 0; k < hull_offsets.size(); ++k)
        v[hull_offsets[k].collider].hull = ctx->d_hulls.p + k;
    HIP_TRY(ctx->d_tree_nodes.upload(tree_nodes));
    HIP_TRY(ctx->d_tree_tris.upload(tree_tris));
    std::vector<TreeDev> trees(tree_offsets.size());
    for (size_t k = 0; k < tree_offsets.size(); ++k)
    {
        trees[k].nodes = ctx->d_tree_nodes.p;
        trees[k].triangles = ctx->d_tree_tris.p;
        trees[k].root = tree_offsets[k].root;
        trees[k].height_field = (tree_offsets[k].flags & XPBD_TREE_HEIGHT_FIELD) ? 1u : 0u;
    }
    HIP_TRY(ctx->d_trees.upload(trees));
    for (size_t k = 0; k < tree_offsets.size(); ++k)
        v[tree_offsets[k].collider].tree = ctx->d_trees.p + k;
    HIP_TRY(ctx->d_part_ops.upload(op_pool));
    for (size_t k = 0; k < v.size(); ++k)
        v[k].ops = v[k].num_ops ? ctx->d_part_ops.p + op_offset[k] : nullptr;
    if (v.size() != ctx->colliders.size() || (num_dynamic > 0) != (ctx->num_dynamic > 0) || (num_sensors > 0) != (ctx->num_sensors > 0))
        ctx->layout_dirty = true;
    ctx->colliders = v;
    ctx->collider_first = first;
    // per collider body: shape type, motion type, sensor flag, compound flag and the caller's
    // identifier; a contact settings table holds for the list it was given for
    ctx->collider_sig.resize(2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
    {
        ctx->collider_sig[2 * i] = (uint32_t)(c[i].num_parts ? 0 : c[i].type) | ((uint32_t)c[i].motion_type << 8)
                                   | (c[i].is_sensor ? 1u << 16 : 0u) | (std::min(c[i].num_parts, 0x7fffu) << 17);
        ctx->collider_sig[2 * i + 1] = c[i].user_id;
    }
    ctx->num_dynamic = num_dynamic;
    ctx->num_sensors = num_sensors;
    ctx->ext_colliders = false;
    for (const ColliderDev &d : v)
        ctx->ext_colliders |= d.type == XPBD_SHAPE_MESH || d.num_ops > 0
                              || (d.type == XPBD_SHAPE_CONVEX_HULL
                                  && (d.scale[0] - 1.0f) * (d.scale[0] - 1.0f) + (d.scale[1] - 1.0f) * (d.scale[1] - 1.0f)
                                         + (d.scale[2] - 1.0f) * (d.scale[2] - 1.0f) > 1.0e-8f);
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

namespace {
int PushSkin(xpbd_context *ctx);
} // namespace

int xpbd_update(xpbd_context *ctx, float dt, int32_t collision_steps, int32_t frames)
{
    if (!ctx)
        return Fail(XPBD_E_INVALID, "null context");
    if (collision_steps <= 0 || frames < 0 || !(dt > 0.0f))
        return Fail(XPBD_E_INVALID, "dt must be > 0, collision_steps > 0, frames >= 0");
    if (frames == 0 || ctx->bodies.empty())
        return XPBD_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t nc = (uint32_t)ctx->colliders.size();     // leaf shapes
    const uint32_t num_bodies = (uint32_t)ctx->collider_first.size();
    int rc;
    for (size_t i = 0; i < ctx->bodies.size(); ++i)
        if (ctx->bodies[i]->pair_n && (ctx->bodies[i]->pair_n != num_bodies || ctx->bodies[i]->pair_sig != ctx->collider_sig))
            return Fail(XPBD_E_INVALID, "body " + std::to_string(i) + " has contact settings for another collider list");
    // skin states go up before the layout, whose bundles point at them
    if ((rc = PushSkin(ctx)) != XPBD_OK)
        return rc;
    if (!ctx->skin_host.empty())
        ctx->skin_on_device = true;
    if (ctx->layout_dirty || ctx->layout_colliders != nc)
    {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if ((rc = PullStates(ctx)) != XPBD_OK)
            return rc;
        if ((rc = BuildLayout(ctx)) != XPBD_OK)
            return rc;
    }

    // Streamed bodies: launch classes (kernel variant x LDS size bin)
    struct Launch { bool resident, staged, rods; uint32_t lds, max_leaves; std::vector<uint32_t> bodies; };
    std::vector<Launch> launches;
    const uint32_t bins[] = { 24 * 1024, 40 * 1024, 64 * 1024, 96 * 1024, kMaxLdsBytes };
    std::vector<uint32_t> order = ctx->streamed;
    std::stable_sort(order.begin(), order.end(), [ctx](uint32_t a, uint32_t b) { return ctx->bodies[a]->nv > ctx->bodies[b]->nv; });
    for (uint32_t i : order)
    {
        Body &b = *ctx->bodies[i];
        const SettingsDev &sd = b.dev.s;
        const uint32_t nf = b.dev.pressure > 0.0f ? sd.num_faces : 0u;
        const bool resident = sd.resident != 0;
        const bool rods = b.dev.rods != nullptr;
        bool staged = false;
        // collider leaves: as many as the LDS left over holds, at most one per collider (a
        // body that overlaps more colliders in one step raises the capacity error)
        uint32_t max_leaves = nc;
        const uint32_t no_leaves = kernel_lds_bytes(resident, false, sd, 0, nc, nf);
        if (no_leaves <= kMaxLdsBytes)
            max_leaves = std::min(nc, (kMaxLdsBytes - no_leaves) / bundle_leaf_bytes());
        uint32_t need = kernel_lds_bytes(resident, false, sd, max_leaves, nc, nf);
        if (resident)
        {
            uint32_t need_staged = kernel_lds_bytes(true, true, sd, max_leaves, nc, nf);
            if (need_staged <= kMaxLdsBytes)
            {
                staged = true;
                need = need_staged;
            }
        }
        if (need > kMaxLdsBytes)
            return Fail(XPBD_E_CAPACITY, "body " + std::to_string(i) + " needs " + std::to_string(need)
                                             + " bytes of LDS; a workgroup has " + std::to_string(kMaxLdsBytes));
        if (b.dev.stage_records != (staged ? 1u : 0u))
        {
            b.dev.stage_records = staged ? 1u : 0u;
            ctx->tables_dirty = true;
        }
        uint32_t bin = 0;
        while (bins[bin] < need) ++bin;
        Launch *L = nullptr;
        for (Launch &l : launches)
            if (l.resident == resident && l.staged == staged && l.rods == rods && l.lds == bins[bin] && l.max_leaves == max_leaves)
                L = &l;
        if (!L)
        {
            launches.push_back({ resident, staged, rods, bins[bin], max_leaves, {} });
            L = &launches.back();
        }
        L->bodies.push_back(i);
    }
    // dynamic colliders: one row of states per body (the kernels clear a body's row at
    // the start of every call; zeroed here too, for bodies no launch runs)
    {
        const uint32_t nd = ctx->num_dynamic;
        const size_t need = nd ? ctx->bodies.size() * nd : 0;
        if (ctx->d_dyn.n != need || ctx->dyn_stride != nd)
        {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            HIP_TRY(ctx->d_dyn.alloc(need));
            if (need)
                HIP_TRY(hipMemset(ctx->d_dyn.p, 0, sizeof(DynStateDev) * need));
            for (size_t i = 0; i < ctx->bodies.size(); ++i)
            {
                ctx->bodies[i]->dev.dyn = nd ? ctx->d_dyn.p + i * nd : nullptr;
                ctx->bodies[i]->dev.num_dynamic = nd;
            }
            ctx->dyn_stride = nd;
            ctx->tables_dirty = true;
        }
    }
    // sensor colliders: one row of contact flags per body
    {
        const uint32_t ns = ctx->num_sensors;
        const size_t need = ns ? ctx->bodies.size() * ns : 0;
        if (ctx->d_sensor_hits.n != need || ctx->sensor_stride != ns)
        {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            HIP_TRY(ctx->d_sensor_hits.alloc(need));
            if (need)
                HIP_TRY(hipMemset(ctx->d_sensor_hits.p, 0, sizeof(int32_t) * need));
            for (size_t i = 0; i < ctx->bodies.size(); ++i)
            {
                ctx->bodies[i]->dev.sensor_hits = ns ? ctx->d_sensor_hits.p + i * ns : nullptr;
                ctx->bodies[i]->dev.num_sensors = ns;
            }
            ctx->sensor_stride = ns;
            ctx->tables_dirty = true;
        }
    }
    if ((rc = SyncTables(ctx)) != XPBD_OK)
        return rc;
    std::vector<uint32_t> &list = ctx->host_list;
    list.clear();
    for (const Launch &l : launches) list.insert(list.end(), l.bodies.begin(), l.bodies.end());
    if (!list.empty())
    {
        if (ctx->d_list.n != list.size())
            HIP_TRY(ctx->d_list.alloc(list.size()));
        HIP_TRY(hipMemcpyAsync(ctx->d_list.p, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice, ctx->stream));
    }

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profiling)
    {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, ctx->stream));
    }
    // Order each bundle launch by the durations the first update of this layout
    // measured (ordered on the context stream after the previous update's join,
    // so no running workgroup reads a table while it is rewritten).
    if (ctx->order_state == xpbd_context::kOrderPending && hipEventQuery(ctx->elapsed_ready) == hipSuccess)
    {
        for (auto &Lp : ctx->bundle_launches)
        {
            auto &L = *Lp;
            auto ticks = [&](const BundleDev &d) { return ctx->h_elapsed[d.elapsed - ctx->d_elapsed.p]; };
            std::stable_sort(L.host.begin(), L.host.end(), [&](const BundleDev &a, const BundleDev &b) { return ticks(a) > ticks(b); });
            HIP_TRY(hipMemcpyAsync(L.table.p, L.host.data(), sizeof(BundleDev) * L.host.size(), hipMemcpyHostToDevice, ctx->stream));
        }
        ctx->order_state = xpbd_context::kOrderMeasured;
    }

    // Independent groups of work run concurrently on the context stream and the
    // auxiliary streams, one per hardware queue.  Kernels of one stream run one
    // after the other, so each launch group gets a stream of its own while there
    // are enough: the streamed bodies on the context stream, then the bundle
    // launches, largest LDS first.  With more groups than streams a group goes to
    // the stream with the least estimated work queued.
    const int num_aux = ctx->num_aux;
    HIP_TRY(hipEventRecord(ctx->fork, ctx->stream));
    for (int a = 0; a < num_aux; ++a)
        HIP_TRY(hipStreamWaitEvent(ctx->aux[a], ctx->fork, 0));
    const float gx = ctx->gravity[0], gy = ctx->gravity[1], gz = ctx->gravity[2];
    hipStream_t streams[xpbd_context::kMaxAux + 1] = { ctx->stream };
    for (int a = 0; a < num_aux; ++a) streams[a + 1] = ctx->aux[a];
    double queued[xpbd_context::kMaxAux + 1] = {};
    auto least_queued = [&]() {
        int best = 0;
        for (int q = 1; q <= num_aux; ++q)
            if (queued[q] < queued[best])
                best = q;
        return best;
    };
    // teams first, on the context stream: their members need a CU each at once, which
    // they get before the other launches take the CUs
    if (!ctx->teams.empty())
    {
        HIP_TRY(hipMemsetAsync(ctx->d_team_sync.p, 0, sizeof(uint32_t) * ctx->d_team_sync.n, ctx->stream));
        HIP_TRY(launch_team(ctx->ext_colliders, ctx->d_bodies.p, ctx->d_states.p, ctx->d_teams.p, ctx->d_team_blocks.p, (uint32_t)ctx->team_blocks.size(),
                            ctx->d_colliders.p, nc, ctx->team_max_leaves, ctx->team_lds, gx, gy, gz, dt, collision_steps, frames,
                            ctx->d_error.p, ctx->stream));
        double longest = 0.0;
        for (const TeamDev &t : ctx->teams)
            longest = std::max(longest, StreamedMicroseconds(*ctx->bodies[t.body]));
        queued[0] += longest;
    }
    uint32_t offset = 0;
    for (const Launch &l : launches)
    {
        const int q = ctx->teams.empty() ? 0 : least_queued();
        HIP_TRY(launch_update(l.resident, l.staged, l.rods, ctx->ext_colliders, ctx->d_bodies.p, ctx->d_states.p, ctx->d_list.p + offset, (uint32_t)l.bodies.size(),
                              ctx->d_colliders.p, nc, l.max_leaves, l.lds, gx, gy, gz, dt, collision_steps, frames, ctx->d_error.p, streams[q]));
        offset += (uint32_t)l.bodies.size();
        double longest = 0.0;
        for (uint32_t bi : l.bodies)
            longest = std::max(longest, StreamedMicroseconds(*ctx->bodies[bi]));
        queued[q] += longest;
    }
    for (const auto &Lp : ctx->bundle_launches)
    {
        const auto &L = *Lp;
        const int best = least_queued();
        queued[best] += std::max(L.cost, L.work / 256.0);
        HIP_TRY(launch_bundles(L.staged, L.block, L.uniform, ctx->ext_colliders, L.table.p, L.count, ctx->d_bodies.p, ctx->d_states.p, ctx->d_colliders.p, nc,
                               L.lds, gx, gy, gz, dt, collision_steps, frames, streams[best]));
    }
    for (int a = 0; a < num_aux; ++a)
    {
        HIP_TRY(hipEventRecord(ctx->join[a], ctx->aux[a]));
        HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->join[a], 0));
    }
    if (ctx->profiling)
    {
        HIP_TRY(hipEventRecord(e1, ctx->stream));
        ctx->events.emplace_back(e0, e1);
    }
    if (ctx->order_state == xpbd_context::kOrderMeasure && !ctx->bundle_launches.empty())
    {
        HIP_TRY(hipMemcpyAsync(ctx->h_elapsed, ctx->d_elapsed.p, sizeof(unsigned long long) * ctx->bundles.size(),
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipEventRecord(ctx->elapsed_ready, ctx->stream));
        ctx->order_state = xpbd_context::kOrderPending;
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
        if (err & 4)
            return Fail(XPBD_E_DEVICE, "a team of workgroups waited too long at a barrier (its members did not all run at once); "
                                       "the update's results are invalid");
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
    auto get3 = [&](const float4 *src, float *dst, int comps) -> int {
        if (!dst)
            return XPBD_OK;
        HIP_TRY(hipMemcpy(tmp.data(), src, sizeof(float4) * nv, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nv; ++i)
        {
            dst[comps * i] = tmp[i].x; dst[comps * i + 1] = tmp[i].y; dst[comps * i + 2] = tmp[i].z;
            if (comps == 4) dst[comps * i + 3] = tmp[i].w;
        }
        return XPBD_OK;
    };
    if ((rc = get3(b.dev.x, pos, 3)) != XPBD_OK || (rc = get3(b.dev.prev, prev, 3)) != XPBD_OK || (rc = get3(b.dev.v, vel, 3)) != XPBD_OK
        || (rc = get3(b.dev.plane, plane, 4)) != XPBD_OK)
        return rc;
    if (shape_idx)
        HIP_TRY(hipMemcpy(shape_idx, b.dev.idx, sizeof(int32_t) * nv, hipMemcpyDeviceToHost));
    if (contact)
        HIP_TRY(hipMemcpy(contact, b.dev.contact, sizeof(int32_t) * nv, hipMemcpyDeviceToHost));
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
        HIP_TRY(hipMemcpy(tmp.data(), b.dev.x, sizeof(float4) * nv, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nv; ++i)
            tmp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], tmp[i].w);
        HIP_TRY(hipMemcpy(b.dev.x, tmp.data(), sizeof(float4) * nv, hipMemcpyHostToDevice));
    }
    if (vel)
    {
        for (uint32_t i = 0; i < nv; ++i)
            tmp[i] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f);
        HIP_TRY(hipMemcpy(b.dev.v, tmp.data(), sizeof(float4) * nv, hipMemcpyHostToDevice));
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

int xpbd_body_placement(xpbd_context *ctx, int32_t body, int32_t out[4])
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size() || !out)
        return Fail(XPBD_E_INVALID, "invalid body");
    if (ctx->layout_dirty || ctx->layout_colliders != (uint32_t)ctx->colliders.size())
        return Fail(XPBD_E_INVALID, "the layout is made by the next xpbd_update");
    out[0] = XPBD_PLACED_STREAMED; out[1] = 1024; out[2] = -1; out[3] = 1;
    for (size_t k = 0; k < ctx->teams.size(); ++k)
        if (ctx->teams[k].body == (uint32_t)body)
        {
            out[0] = XPBD_PLACED_TEAM; out[2] = (int32_t)k; out[3] = (int32_t)ctx->teams[k].size;
            return XPBD_OK;
        }
    for (const auto &Lp : ctx->bundle_launches)
        for (const BundleDev &d : Lp->host)
            for (size_t k = 0; k < ctx->bundles.size(); ++k)
                if (ctx->bundles[k]->dev.recs == d.recs)
                    for (uint32_t bi : ctx->bundles[k]->bodies)
                        if (bi == (uint32_t)body)
                        {
                            out[0] = XPBD_PLACED_BUNDLE; out[1] = (int32_t)Lp->block; out[2] = (int32_t)k;
                            out[3] = (int32_t)ctx->bundles[k]->bodies.size();
                            return XPBD_OK;
                        }
    return XPBD_OK;
}

namespace {
constexpr size_t kStateHeaderBytes = 4 * 12 + 48 + 4 + 1;   // velocities, force, torque, spheres, timer, allow sleeping
// header, per vertex position + velocity, per rod rotation (16 B) + angular velocity (12 B),
// per vertex of a skinned body the skin state (previous position, position, normal: 36 B), bounds
size_t StateBytes(uint32_t nv, size_t nrods, bool skinned)
{
    return kStateHeaderBytes + (size_t)nv * 24 + nrods * 28 + (skinned ? (size_t)nv * 36 : 0) + 4 * 12;
}
} // namespace

namespace {
// Brings the host copy of the skin states up to date after updates moved them.
// Skin states whose host copy is newer go to the device (allocated there on first use).
int PushSkin(xpbd_context *ctx)
{
    if (!ctx->skin_dirty)
        return XPBD_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->d_skin.n != ctx->skin_host.size())
    {
        HIP_TRY(ctx->d_skin.alloc(ctx->skin_host.size()));
        for (auto &b : ctx->bodies)
            if (b->skinned)
                b->dev.skin = ctx->d_skin.p;
        ctx->tables_dirty = true;
        ctx->layout_dirty = true;
    }
    HIP_TRY(hipMemcpy(ctx->d_skin.p, ctx->skin_host.data(), sizeof(float4) * ctx->skin_host.size(), hipMemcpyHostToDevice));
    ctx->skin_dirty = false;
    return XPBD_OK;
}

int PullSkin(xpbd_context *ctx)
{
    if (!ctx->skin_on_device || ctx->d_skin.n != ctx->skin_host.size())
        return XPBD_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(ctx->skin_host.data(), ctx->d_skin.p, sizeof(float4) * ctx->skin_host.size(), hipMemcpyDeviceToHost));
    ctx->skin_on_device = false;
    return XPBD_OK;
}
} // namespace

int xpbd_body_skin_vertices(xpbd_context *ctx, int32_t body, const float *joint_matrices, uint32_t num_joints,
                            int32_t hard_skin_all)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size() || (num_joints && !joint_matrices))
        return Fail(XPBD_E_INVALID, "invalid body or joint matrices");
    Body &b = *ctx->bodies[body];
    const xpbd_settings &st = *b.settings;
    if (!b.skinned)
        return Fail(XPBD_E_INVALID, "body has no skinned constraints");
    for (const InvBindRec &ib : st.inv_binds)
        if (ib.joint >= num_joints)
            return Fail(XPBD_E_INVALID, "inverse bind matrix refers to a missing joint");
    int rc = PushSkin(ctx);
    if (rc != XPBD_OK)
        return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    // the joint matrices are the call's only input from host memory
    if (b.d_joints.n < 16 * (size_t)num_joints)
    {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(b.d_joints.alloc(16 * (size_t)std::max(1u, num_joints)));
    }
    if (b.d_skin_m.n < 16 * st.inv_binds.size())
        HIP_TRY(b.d_skin_m.alloc(16 * std::max<size_t>(1, st.inv_binds.size())));
    if (num_joints)
    {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(b.d_joints.p, joint_matrices, sizeof(float) * 16 * num_joints, hipMemcpyHostToDevice));
    }
    HIP_TRY(launch_skin(b.d_joints.p, st.d_inv_binds.p, (uint32_t)st.inv_binds.size(), b.d_skin_m.p,
                        st.d_skinned.p, (uint32_t)st.skinned.size(), st.d_skinned_normals.p, st.faces.p,
                        ctx->d_skin.p + 4 * (size_t)b.dev.skin_base, b.nv, b.dev.x, b.dev.v, hard_skin_all, ctx->stream));
    ctx->skin_on_device = true;
    return XPBD_OK;
}

int xpbd_body_set_skin_options(xpbd_context *ctx, int32_t body, int32_t enable_skin_constraints, float max_distance_multiplier)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    Body &b = *ctx->bodies[body];
    if (!b.skinned)
        return Fail(XPBD_E_INVALID, "body has no skinned constraints");
    int rc = PullSkin(ctx);
    if (rc != XPBD_OK)
        return rc;
    float4 *E = ctx->skin_host.data() + 4 * (size_t)b.dev.skin_base;
    for (uint32_t i = 0; i < b.nv; ++i)
    {
        E[4 * (size_t)i + 3].x = max_distance_multiplier;
        E[4 * (size_t)i + 3].z = enable_skin_constraints ? 1.0f : 0.0f;
    }
    ctx->skin_dirty = true;
    return XPBD_OK;
}

int xpbd_body_save_state(xpbd_context *ctx, int32_t body, void *out, size_t capacity, size_t *size)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size() || !size)
        return Fail(XPBD_E_INVALID, "invalid body or size pointer");
    const Body &b = *ctx->bodies[body];
    const size_t nrods = b.settings->rods.size();
    *size = StateBytes(b.nv, nrods, b.skinned);
    if (!out)
        return XPBD_OK;
    if (capacity < *size)
        return Fail(XPBD_E_INVALID, "state buffer too small");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    if ((rc = PullStates(ctx)) != XPBD_OK)
        return rc;
    if (b.skinned && (rc = PullSkin(ctx)) != XPBD_OK)
        return rc;
    const BodyStateDev &s = b.state;
    std::vector<float4> x(b.nv), v(b.nv);
    if (b.nv)
    {
        HIP_TRY(hipMemcpy(x.data(), b.dev.x, sizeof(float4) * b.nv, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(v.data(), b.dev.v, sizeof(float4) * b.nv, hipMemcpyDeviceToHost));
    }
    unsigned char *p = (unsigned char *)out;
    auto put = [&p](const void *src, size_t n) { memcpy(p, src, n); p += n; };
    const float zero[12] = {};
    put(s.lin, 12); put(s.ang, 12);
    put(zero, 12); put(zero, 12);               // force, torque
    put(zero, 48);                              // sleep test spheres
    put(&s.sleep_timer, 4);
    const unsigned char allow = b.dev.allow_sleeping ? 1 : 0;
    put(&allow, 1);
    for (uint32_t i = 0; i < b.nv; ++i)
    {
        put(&x[i], 12);
        put(&v[i], 12);
    }
    if (nrods)
    {
        std::vector<float4> rs(2 * nrods);
        HIP_TRY(hipMemcpy(rs.data(), b.dev.rods, sizeof(float4) * rs.size(), hipMemcpyDeviceToHost));
        for (size_t r = 0; r < nrods; ++r)
        {
            put(&rs[2 * r], 16);
            put(&rs[2 * r + 1], 12);
        }
    }
    if (b.skinned)
    {
        // SkinState per vertex: previous position, position, normal
        const float4 *E = ctx->skin_host.data() + 4 * (size_t)b.dev.skin_base;
        for (uint32_t i = 0; i < b.nv; ++i)
            for (int k = 0; k < 3; ++k)
                put(&E[4 * i + k], 12);
    }
    put(s.bmin, 12); put(s.bmax, 12); put(s.pmin, 12); put(s.pmax, 12);
    return XPBD_OK;
}

int xpbd_body_restore_state(xpbd_context *ctx, int32_t body, const void *in, size_t size, const float *position,
                            const float *rotation)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size() || !in)
        return Fail(XPBD_E_INVALID, "invalid body or stream");
    Body &b = *ctx->bodies[body];
    const size_t nrods = b.settings->rods.size();
    if (size != StateBytes(b.nv, nrods, b.skinned))
        return Fail(XPBD_E_INVALID, "state stream size does not match the body");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    if ((rc = PullStates(ctx)) != XPBD_OK)
        return rc;
    if (b.skinned && (rc = PullSkin(ctx)) != XPBD_OK)
        return rc;
    const unsigned char *p = (const unsigned char *)in;
    auto get = [&p](void *dst, size_t n) { memcpy(dst, p, n); p += n; };
    BodyStateDev &s = b.state;
    get(s.lin, 12); get(s.ang, 12);
    p += 12 + 12 + 48;                          // force, torque, sleep test spheres
    get(&s.sleep_timer, 4);
    unsigned char allow = 0;
    get(&allow, 1);
    if ((b.dev.allow_sleeping != 0) != (allow != 0))
        ctx->layout_dirty = true;   // bundles record whether a body may sleep
    b.dev.allow_sleeping = allow ? 1 : 0;
    std::vector<float4> x(b.nv), v(b.nv);
    if (b.nv)
        HIP_TRY(hipMemcpy(x.data(), b.dev.x, sizeof(float4) * b.nv, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < b.nv; ++i)
    {
        float q[6];
        get(q, 24);
        x[i] = make_float4(q[0], q[1], q[2], x[i].w);
        v[i] = make_float4(q[3], q[4], q[5], 0.0f);
    }
    std::vector<float4> rs(2 * nrods);
    for (size_t r = 0; r < nrods; ++r)
    {
        float q[7];
        get(q, 28);
        rs[2 * r] = make_float4(q[0], q[1], q[2], q[3]);
        rs[2 * r + 1] = make_float4(q[4], q[5], q[6], 0.0f);
    }
    if (nrods)
        HIP_TRY(hipMemcpy(b.dev.rods, rs.data(), sizeof(float4) * rs.size(), hipMemcpyHostToDevice));
    if (b.skinned)
    {
        // skin states (the previous-position-valid flag is not part of the stream: it stays)
        float4 *E = ctx->skin_host.data() + 4 * (size_t)b.dev.skin_base;
        for (uint32_t i = 0; i < b.nv; ++i)
            for (int k = 0; k < 3; ++k)
            {
                float q[3];
                get(q, 12);
                E[4 * i + k] = make_float4(q[0], q[1], q[2], E[4 * i + k].w);
            }
        ctx->skin_dirty = true;
    }
    get(s.bmin, 12); get(s.bmax, 12); get(s.pmin, 12); get(s.pmax, 12);
    if (position) memcpy(s.pos, position, 12);
    if (rotation) memcpy(s.rot, rotation, 16);
    if (b.nv)
    {
        HIP_TRY(hipMemcpy(b.dev.x, x.data(), sizeof(float4) * b.nv, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b.dev.v, v.data(), sizeof(float4) * b.nv, hipMemcpyHostToDevice));
    }
    // the host copies of all body states are current (pulled above) and now hold the
    // restored one: they go up with the body table at the next update
    ctx->states_on_device = false;
    ctx->tables_dirty = true;
    return XPBD_OK;
}

int xpbd_body_contacts(xpbd_context *ctx, int32_t body, float *points, int32_t *collider)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    const Body &b = *ctx->bodies[body];
    std::vector<float4> rep(b.nv);
    if (b.nv)
        HIP_TRY(hipMemcpy(rep.data(), b.dev.report, sizeof(float4) * b.nv, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < b.nv; ++i)
    {
        int32_t c;
        memcpy(&c, &rep[i].w, sizeof(c));
        if (points) { points[3 * i] = rep[i].x; points[3 * i + 1] = rep[i].y; points[3 * i + 2] = rep[i].z; }
        if (collider) collider[i] = c;
    }
    return XPBD_OK;
}

int xpbd_body_sensor_contacts(xpbd_context *ctx, int32_t body, uint32_t num_colliders, int32_t *hits)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size() || !hits)
        return Fail(XPBD_E_INVALID, "invalid body or output");
    if (num_colliders != ctx->collider_first.size())
        return Fail(XPBD_E_INVALID, "num_colliders differs from the last xpbd_set_colliders");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    std::vector<int32_t> row(ctx->sensor_stride);
    if (ctx->sensor_stride && ctx->d_sensor_hits.n)
        HIP_TRY(hipMemcpy(row.data(), ctx->d_sensor_hits.p + (size_t)body * ctx->sensor_stride, sizeof(int32_t) * row.size(),
                          hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < num_colliders; ++c)
    {
        const int32_t k = ctx->colliders[ctx->collider_first[c]].sensor;
        hits[c] = k >= 0 && (uint32_t)k < row.size() ? row[k] : 0;
    }
    return XPBD_OK;
}

int xpbd_body_set_contact_settings(xpbd_context *ctx, int32_t body, uint32_t num_colliders, const float *settings)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    Body &b = *ctx->bodies[body];
    if (!settings)
    {
        b.pair_n = 0;
        b.dev.pair = nullptr;
        ctx->tables_dirty = true;
        return XPBD_OK;
    }
    if (num_colliders != ctx->collider_first.size())
        return Fail(XPBD_E_INVALID, "num_colliders differs from the last xpbd_set_colliders");
    for (uint32_t c = 0; c < num_colliders; ++c)
    {
        const float *q = settings + 4 * c;
        if (!(q[0] >= 0.0f && q[1] >= 0.0f && q[2] >= 0.0f) || (q[3] != 0.0f && q[3] != 1.0f))
            return Fail(XPBD_E_INVALID, "contact settings: scales must be >= 0 and accept 0 or 1");
    }
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    std::vector<float4> row(std::max(1u, num_colliders), make_float4(1.0f, 1.0f, 1.0f, 1.0f));
    for (uint32_t c = 0; c < num_colliders; ++c)
        row[c] = make_float4(settings[4 * c], settings[4 * c + 1], settings[4 * c + 2], settings[4 * c + 3]);
    HIP_TRY(b.pair.upload(row));
    b.pair_n = num_colliders;
    b.pair_sig = ctx->collider_sig;
    b.dev.pair = num_colliders ? b.pair.p : nullptr;
    ctx->tables_dirty = true;
    return XPBD_OK;
}

int xpbd_body_rod_states(xpbd_context *ctx, int32_t body, uint32_t num_rods, float *rotations, float *angular_velocities)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    const Body &b = *ctx->bodies[body];
    const size_t n = b.settings->rods.size();
    if (num_rods != n)
        return Fail(XPBD_E_INVALID, "num_rods differs from the body's rod stretch shear constraints");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    std::vector<float4> rs(2 * n);
    if (n)
        HIP_TRY(hipMemcpy(rs.data(), b.dev.rods, sizeof(float4) * rs.size(), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < n; ++r)
    {
        if (rotations)
            memcpy(rotations + 4 * r, &rs[2 * r], 16);
        if (angular_velocities)
            memcpy(angular_velocities + 3 * r, &rs[2 * r + 1], 12);
    }
    return XPBD_OK;
}

int xpbd_body_velocity_deltas(xpbd_context *ctx, int32_t body, uint32_t num_colliders, float *linear, float *angular,
                              int32_t *updated)
{
    if (!ctx || body < 0 || body >= (int32_t)ctx->bodies.size())
        return Fail(XPBD_E_INVALID, "invalid body");
    if (num_colliders != ctx->collider_first.size())
        return Fail(XPBD_E_INVALID, "num_colliders differs from the last xpbd_set_colliders");
    int rc = xpbd_synchronize(ctx);
    if (rc != XPBD_OK)
        return rc;
    std::vector<DynStateDev> row(ctx->dyn_stride);
    if (ctx->dyn_stride && ctx->d_dyn.n)
        HIP_TRY(hipMemcpy(row.data(), ctx->d_dyn.p + (size_t)body * ctx->dyn_stride, sizeof(DynStateDev) * row.size(),
                          hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < num_colliders; ++c)
    {
        const int32_t k = ctx->colliders[ctx->collider_first[c]].dyn;
        const bool have = k >= 0 && (uint32_t)k < row.size();
        for (int j = 0; j < 3; ++j)
        {
            if (linear) linear[3 * c + j] = have ? row[k].acc_lin[j] : 0.0f;
            if (angular) angular[3 * c + j] = have ? row[k].acc_ang[j] : 0.0f;
        }
        if (updated) updated[c] = have ? row[k].any_updated : 0;
    }
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

// ------------------------------------------------------------------ rollout
struct xpbd_rollout
{
    xpbd_context *ctx = nullptr;
    std::vector<xpbd_settings *> settings;
    DevBuf<uint4> recs;
    DevBuf<uint32_t> lvl;
    DevBuf<float4> x0;
    DevBuf<RolloutBodyDev> bodies;
    DevBuf<float4> ckpt_x, ckpt_v, seg_x, seg_v, vel, prev, rec_tape, adj_x, adj_v;
    DevBuf<double> vel64, prev64;
    RolloutDev dev{};
    uint32_t total_nv = 0, total_ne = 0;
    uint32_t lds_f32 = 0, lds_f64 = 0, lds_bwd = 0;
    bool has_forward = false;

    ~xpbd_rollout()
    {
        for (xpbd_settings *s : settings)
            xpbd_settings_release(s);
    }
};

namespace {
// The caller's stream as given: NULL is the default (null) stream, as in torch.
hipStream_t StreamOf(xpbd_rollout *, void *stream) { return (hipStream_t)stream; }
}

int xpbd_rollout_create(xpbd_context *ctx, const xpbd_rollout_desc *d, xpbd_rollout **out)
{
    if (!ctx || !d || !out || (d->num_bodies > 0 && !d->settings))
        return Fail(XPBD_E_INVALID, "null argument");
    *out = nullptr;
    if (d->num_bodies == 0 || d->num_iterations == 0 || d->frames == 0 || d->checkpoint_interval == 0 || !(d->dt > 0.0f))
        return Fail(XPBD_E_INVALID, "num_bodies, num_iterations, frames, checkpoint_interval and dt must be positive");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<xpbd_rollout> r(new xpbd_rollout);
    r->ctx = ctx;
    std::vector<uint4> recs;
    std::vector<uint32_t> lvl;
    std::vector<float4> x0;
    std::vector<RolloutBodyDev> bodies(d->num_bodies);
    for (uint32_t b = 0; b < d->num_bodies; ++b)
    {
        xpbd_settings *s = d->settings[b];
        if (!s || s->ctx != ctx)
            return Fail(XPBD_E_INVALID, "settings " + std::to_string(b) + " missing or from another context");
        if (s->host_level_begin.empty())
            return Fail(XPBD_E_CAPACITY, "rollout body " + std::to_string(b) + " does not fit in LDS");
        const uint32_t nv = s->num_vertices, ne = s->num_edges;
        const uint32_t nf = d->pressure > 0.0f ? s->num_faces : 0u;
        if (rollout_backward_lds_bytes(nv, ne, nf) > kMaxLdsBytes)
            return Fail(XPBD_E_CAPACITY, "rollout body " + std::to_string(b) + ": positions, edge adjoints and face terms exceed LDS");
        ++s->refcount;
        r->settings.push_back(s);
        RolloutBodyDev &B = bodies[b];
        B.nv = nv;
        B.num_edges = ne;
        B.voff = r->total_nv;
        B.eoff = r->total_ne;
        B.rbase = (uint32_t)recs.size();
        B.nlev = (uint32_t)s->host_level_begin.size() - 1;
        B.num_records = (uint32_t)s->host_recs.size();
        B.lvl = reinterpret_cast<const uint32_t *>((uintptr_t)lvl.size());     // index until uploaded
        B.num_faces = nf;
        B.pressure = d->pressure;
        B.faces = s->faces.p;           // the settings' device arrays (the rollout holds a reference)
        B.vf_begin = s->vf_begin.p;
        B.vf_list = s->vf_list.p;
        for (uint32_t L : s->host_level_begin)
            lvl.push_back(L);                       // relative to the body's first record
        B.edges_only = 1u;
        for (size_t k = 0; k < s->host_recs.size(); ++k)
        {
            uint4 q = s->host_recs[k];
            if (q.x != kNopRecordHost && ((q.x >> 14) & 3u) != kTypeEdge)
                B.edges_only = 0u;
            if (q.x != kNopRecordHost && ((q.x >> 14) & 3u) != kTypeVolume && ((q.x >> 14) & 3u) != kTypeBend)
            {
                // edges: their index in the body (for the compliance parameters) and a compliance
                // written by each float forward pass; LRAs: their index.  Volume and bend records
                // keep their vertices 2-3 and parameters.
                q.y = s->host_rec_index[k];
                q.w = 0u;
            }
            recs.push_back(q);
        }
        for (const VertexRec &v : s->vertices)
            x0.push_back(make_float4(v.pos[0], v.pos[1], v.pos[2], v.inv_mass));
        r->total_nv += nv;
        r->total_ne += ne;
        r->lds_bwd = std::max(r->lds_bwd, rollout_backward_lds_bytes(nv, ne, nf));
        r->lds_f32 = std::max(r->lds_f32, rollout_forward_lds_bytes(nv, nf, false));
        r->lds_f64 = std::max(r->lds_f64, rollout_forward_lds_bytes(nv, nf, true));
    }
    HIP_TRY(r->recs.upload(recs));
    HIP_TRY(r->lvl.upload(lvl));
    HIP_TRY(r->x0.upload(x0));
    for (RolloutBodyDev &B : bodies)
    {
        B.recs = r->recs.p + B.rbase;
        B.lvl = r->lvl.p + (uintptr_t)B.lvl;
        B.x0 = r->x0.p + B.voff;
    }
    HIP_TRY(r->bodies.upload(bodies));
    const uint32_t num_ckpt = (d->frames + d->checkpoint_interval - 1) / d->checkpoint_interval;
    const size_t nv = r->total_nv;
    const size_t seg_sub = (size_t)d->checkpoint_interval * d->num_iterations;
    HIP_TRY(r->ckpt_x.alloc(num_ckpt * nv)); HIP_TRY(r->ckpt_v.alloc(num_ckpt * nv));
    HIP_TRY(r->seg_x.alloc(seg_sub * nv)); HIP_TRY(r->seg_v.alloc(seg_sub * nv));
    HIP_TRY(r->vel.alloc(nv)); HIP_TRY(r->prev.alloc(nv));
    HIP_TRY(r->vel64.alloc(4 * nv)); HIP_TRY(r->prev64.alloc(4 * nv));
    HIP_TRY(r->rec_tape.alloc(4 * std::max<size_t>(recs.size(), 1)));
    HIP_TRY(r->adj_x.alloc(nv)); HIP_TRY(r->adj_v.alloc(nv));
    RolloutDev &R = r->dev;
    R.bodies = r->bodies.p;
    R.num_bodies = d->num_bodies;
    R.iterations = d->num_iterations;
    R.frames = d->frames;
    R.interval = d->checkpoint_interval;
    R.total_nv = r->total_nv;
    R.dt = d->dt;
    R.linear_damping = d->linear_damping;
    R.max_linear_velocity = d->max_linear_velocity;
    R.gravity[0] = ctx->gravity[0]; R.gravity[1] = ctx->gravity[1]; R.gravity[2] = ctx->gravity[2];
    R.gravity_factor = d->gravity_factor;
    R.ckpt_x = r->ckpt_x.p; R.ckpt_v = r->ckpt_v.p; R.seg_x = r->seg_x.p; R.seg_v = r->seg_v.p;
    R.vel = r->vel.p; R.prev = r->prev.p; R.vel64 = r->vel64.p; R.prev64 = r->prev64.p;
    R.rec_tape = r->rec_tape.p; R.adj_x = r->adj_x.p; R.adj_v = r->adj_v.p;
    *out = r.release();
    return XPBD_OK;
}

int xpbd_rollout_destroy(xpbd_rollout *r)
{
    if (!r)
        return Fail(XPBD_E_INVALID, "null rollout");
    (void)hipSetDevice(r->ctx->device);
    (void)hipDeviceSynchronize();       // work may be queued on any stream the caller passed
    delete r;
    return XPBD_OK;
}

int xpbd_rollout_sizes(xpbd_rollout *r, uint32_t *num_edges, uint32_t *num_vertices)
{
    if (!r)
        return Fail(XPBD_E_INVALID, "null rollout");
    if (num_edges) *num_edges = r->total_ne;
    if (num_vertices) *num_vertices = r->total_nv;
    return XPBD_OK;
}

int xpbd_rollout_forward(xpbd_rollout *r, const float *compliance, const float *velocity, float *positions, void *stream)
{
    if (!r || !compliance || !velocity || !positions)
        return Fail(XPBD_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(launch_rollout_forward_f32(r->dev, compliance, velocity, positions, r->lds_f32, StreamOf(r, stream)));
    r->has_forward = true;
    return XPBD_OK;
}

int xpbd_rollout_forward_noalloc(xpbd_rollout *r, const float *compliance, const float *velocity, float *positions, void *stream)
{
    if (!r || !compliance || !velocity || !positions)
        return Fail(XPBD_E_INVALID, "null argument");
    if (r->total_nv > rollout_noalloc_max_vertices())
        return Fail(XPBD_E_CAPACITY, "rollout has " + std::to_string(r->total_nv) + " vertices; the allocation-free forward pass holds at most "
                                         + std::to_string(rollout_noalloc_max_vertices()));
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(launch_rollout_forward_f32_noalloc(r->ctx->device, r->dev, compliance, velocity, positions, r->lds_f32, StreamOf(r, stream)));
    r->has_forward = true;
    return XPBD_OK;
}

int xpbd_rollout_forward_f64(xpbd_rollout *r, const double *compliance, const double *velocity, double *positions, void *stream)
{
    if (!r || !compliance || !velocity || !positions)
        return Fail(XPBD_E_INVALID, "null argument");
    if (r->lds_f64 > kMaxLdsBytes)
        return Fail(XPBD_E_CAPACITY, "double positions of the largest body exceed LDS");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(launch_rollout_forward_f64(r->dev, compliance, velocity, positions, r->lds_f64, StreamOf(r, stream)));
    return XPBD_OK;
}

int xpbd_rollout_backward(xpbd_rollout *r, const float *grad_positions, float *grad_compliance, float *grad_velocity, void *stream)
{
    if (!r || !grad_positions || !grad_compliance || !grad_velocity)
        return Fail(XPBD_E_INVALID, "null argument");
    if (!r->has_forward)
        return Fail(XPBD_E_INVALID, "backward needs a float forward pass first");
    HIP_TRY(hipSetDevice(r->ctx->device));
    HIP_TRY(launch_rollout_backward(r->dev, grad_positions, grad_compliance, grad_velocity, r->lds_bwd, StreamOf(r, stream)));
    return XPBD_OK;
}

} // extern "C"
