// Device-side data structures shared by the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbd {

// Packed constraint record (16 B), the device form of the reference's Edge /
// DihedralBend / Volume / LRA records (SoftBodySharedSettings.h:157-303):
//   a: bits 0-13 vertex 0, bits 14-15 type, bits 16-31 vertex 1
//   b: bits 0-15 vertex 2, bits 16-31 vertex 3
//   p0, p1: edge {rest, compliance}, bend {compliance, initialAngle},
//           volume {sixRestVolume, compliance}, LRA {maxDistance, -}
// Vertex numbers index the LDS window the record runs in (< 16384).
// Records of LRA type also carry the constraints with state beyond the vertices:
//   b bit 31 set: Cosserat rod stretch-shear, vertices 0 and 1, b & 0x3fffffff = rod
//   b bit 30 set: Cosserat rod bend-twist, b & 0x3fffffff = bend-twist constraint
//   vertex 0 == vertex 1: skinned constraint, b = skin entry
enum : uint32_t { kTypeVolume = 0, kTypeBend = 1, kTypeEdge = 2, kTypeLRA = 3 };
constexpr uint32_t kRodStretchShear = 0x80000000u, kRodBendTwist = 0x40000000u, kRodIndexMask = 0x3fffffffu;
constexpr uint32_t kMaxWindowVertices = 1u << 14;

// A window is a set of vertices held in LDS while a run of levels executes on
// it.  Resident bodies keep all their vertices in LDS for the whole launch and
// use identity windows (vcount == 0: record vertex numbers are body vertex
// indices).  Other bodies gather vlist[vbegin, vbegin+vcount) into LDS, run the
// levels, and scatter the vertices back.
struct WindowDev
{
    uint32_t vbegin, vcount;    // vertex list range (vcount 0: identity window)
    uint32_t lbegin, nlev;      // index of the first level offset in lvl, number of levels
};

struct SettingsDev
{
    const uint4 *recs;          // records ordered (window, level)
    const uint32_t *lvl;        // absolute record offset of each level start, nlev+1 per window
    const WindowDev *windows;   // parallel-group windows, then tail windows
    const uint32_t *vlist;      // gathered windows' vertex lists (body vertex indices)
    const uint32_t *faces;      // 3 vertex indices per face (pressure)
    const uint32_t *vf_begin;   // vertex -> face CSR, faces in increasing order (pressure)
    const uint32_t *vf_list;
    uint32_t num_vertices;
    uint32_t num_par_windows;   // windows [0, num_par_windows) hold the parallel groups
    uint32_t num_tail_windows;  // then the non-parallel group
    uint32_t num_faces;
    uint32_t num_records;
    uint32_t num_levels_total;  // entries of lvl
    uint32_t resident;          // 1: identity windows, positions stay in LDS
    uint32_t window_capacity;   // largest gathered window (vertices)
    // Cosserat rods (SoftBodySharedSettings.h:305-342): per rod {length, inverse mass,
    // compliance, 0}; per bend-twist constraint omega0, then {rod 1, rod 2 as bits, compliance, 0}
    const float4 *rod_const;
    const float4 *bt_const;
    uint32_t num_rods;
};

struct DynStateDev;

struct BodyDev
{
    SettingsDev s;
    float4 *x;                  // position xyz + inverse mass (body-local)
    float4 *v;                  // velocity
    float4 *prev;               // previous position
    float4 *plane;              // collision plane normal + constant
    int32_t *idx;               // colliding shape index (-1 none)
    int32_t *contact;           // has-contact flag
    float4 *report;             // last collision step's contact: point xyz (SoftBodyManifold::GetLocalContactPoint),
                                // w = collider index as int bits, -1 without contact
    uint32_t iterations;
    float linear_damping, max_linear_velocity, restitution, friction, pressure, gravity_factor, vertex_radius;
    int32_t update_position, allow_sleeping;
    uint32_t stage_records;     // 1: records + level table are copied into LDS once per launch
    DynStateDev *dyn;           // this body's row of dynamic-collider states (num_dynamic entries), or null
    uint32_t num_dynamic;
    const float4 *skin;         // skinned vertex states of the context (4 per entry), or null
    uint32_t skin_base;         // first entry of this body (entry = skin_base + body vertex index)
    int32_t *sensor_hits;       // this body's row: 1 where a sensor detected a vertex in the last step, or null
    uint32_t num_sensors;
    float4 *rods;               // rod states, 2 per rod: rotation, then the angular velocity (between
                                // steps) or the previous rotation (during a step); null without rods
    const float4 *pair;         // per collider: contact settings of the pair (see pair_settings), or null
};

// A large body run by a team of workgroups (xpbd_team.hip).  Member m takes the
// parallel-group windows [par_split[m], par_split[m + 1]) and a slice of the vertices.
struct TeamDev
{
    uint32_t body;              // index into the context's body table
    uint32_t size;              // members
    const uint32_t *par_split;  // size + 1 window offsets
    uint32_t *sync;             // team barrier counter (zeroed before every launch)
    float *red;                 // 32 floats per member (partial reductions), then 32 for member 0's result
    // coloured non-parallel group (null: the group runs in the reference's order on member 0):
    // 2 uint4 per record, {v0, v1, v2, v3} and {type, p0, p1, 0}
    const uint4 *colour_recs;
    const uint32_t *colour_begin;   // num_colours + 1 record offsets
    uint32_t num_colours;
    struct Leaf *leaves;        // per member: num_colliders leaves, used when the colliding list overflows LDS
};

// Skinned constraint on the device (SoftBodySharedSettings.h:225-285): the vertex's bind pose
// position, its vertex, up to four (inverse bind, weight) pairs, its normal faces
// {count << 24 | first} and its max distance.
struct SkinnedDev
{
    float bind[3];
    uint32_t vertex;
    uint32_t inv_bind[4];
    float weight[4];
    uint32_t normal_info;
    float max_distance;
    uint32_t pad[2];
};

// Inverse bind matrix (column major) and its joint.
struct SkinInvBindDev
{
    uint32_t joint;
    uint32_t pad[3];
    float m[16];
};

// Per-body dynamic state kept on the device between launches.
struct BodyStateDev
{
    float pos[3];
    float rot[4];
    float bmin[3], bmax[3];     // mLocalBounds
    float pmin[3], pmax[3];     // mLocalPredictedBounds
    float lin[3], ang[3];       // body linear / angular velocity (world)
    float sleep_timer;
    int32_t active;
};

// Convex hull data in device memory (a context-wide table, one entry per hull collider)
struct HullDev
{
    const float4 *planes;       // per face: normal, constant
    const float4 *points;       // xyz relative to the centre of mass
    const uint32_t *face_first; // per face: first entry in index
    const uint32_t *face_count;
    const uint32_t *index;      // face vertex indices, counter-clockwise
    uint32_t num_faces;
};

// A triangle tree (mesh or single triangle) as the reference walks it: four-way nodes with
// float child bounds and child references (XPBD_TREE_* encoding), triangles 9 floats each.
struct TreeNodeDev
{
    float min_x[4], min_y[4], min_z[4];
    float max_x[4], max_y[4], max_z[4];
    uint32_t child[4];
};

struct TreeDev
{
    const TreeNodeDev *nodes;
    const float *triangles;
    uint32_t root;
    uint32_t height_field;      // children with unscaled min y > max y are never visited
};

// One placement step of a compound's leaf (xpbd_part_op)
struct PartOpDev
{
    uint32_t kind;
    float v[6];
};

// One leaf shape of a collider body: a simple body has one, a compound or decorated body one per
// leaf shape (consecutive, in the order the reference collects them), each holding the body's
// pose, material and motion.
struct ColliderDev
{
    int32_t type;
    float params[8];
    float scale[3];             // the leaf shape's scale (ScaledShape)
    const TreeDev *tree;        // XPBD_SHAPE_MESH
    const HullDev *hull;        // XPBD_SHAPE_CONVEX_HULL
    float pos[3];
    float rot[4];
    float friction, restitution;
    float bounds[6];
    // motion (EMotionType: 0 static, 1 kinematic, 2 dynamic); dynamic bodies only:
    int32_t motion;
    int32_t dyn;                // rank among the dynamic colliders
    float inv_mass;
    float inv_inertia_diag[3];  // MotionProperties::GetInverseInertiaDiagonal
    float inertia_rot[4];       // MotionProperties::GetInertiaRotation
    float lin_vel[3], ang_vel[3];   // world space
    int32_t sensor;             // rank among the sensor colliders, -1 for a colliding body
    int32_t owner;              // index of the collider body in the caller's list
    int32_t first;              // 1 for the body's first leaf
    uint32_t num_ops;           // placement steps from the body's centre of mass to the leaf
    const PartOpDev *ops;
};

// State of one dynamic collider as one soft body sees it during a collision step
// (the reference's CollidingShape velocity fields, SoftBodyMotionProperties.h:190-212),
// plus what the step changed, summed over the steps of an update call.
struct DynStateDev
{
    float inv_inertia[9];       // mInvInertia, soft-body local, column major 3x3
    float inv_mass;
    float lin[3], ang[3];       // mLinearVelocity / mAngularVelocity (soft-body local)
    float orig_lin[3], orig_ang[3];
    float com[3];               // mCenterOfMassTransform's translation (soft-body local)
    float acc_lin[3], acc_ang[3];   // world-space velocity changes, summed over the call's steps
    int32_t updated;            // mUpdateVelocities of the current step
    int32_t any_updated;        // some step of the call set mUpdateVelocities
};

} // namespace xpbd

namespace xpbd {

// A bundle is a set of LDS-resident bodies that share one workgroup.  Each
// body's level sequence (merged parallel-group levels, then the non-parallel
// group's levels) is merged with the others' by index, so level k of every
// body runs in one barrier-separated pass; within a level the records are
// grouped by constraint type and padded to whole wavefronts.
struct BundleBodyDev
{
    uint32_t body;              // index into the context's body table
    uint32_t voff;              // first vertex of the body in the bundle
    uint32_t nv;
    uint32_t face_off;          // first face slot in the bundle's face scratch (pressure)
};

struct BundleDev
{
    const uint4 *recs;          // merged level program (vertex numbers are bundle-local)
    const uint32_t *lvl;        // nlev + 1 absolute record offsets
    uint32_t nlev;
    uint32_t num_records;
    const BundleBodyDev *bodies;
    uint32_t num_bodies;
    uint32_t nv;                // vertices of all bodies of the bundle
    float4 *x, *v, *prev, *plane, *report;  // bundle-contiguous vertex state (bodies point into these)
    int32_t *idx, *contact;
    void *leaves_full;          // num_bodies x num_colliders Leaf: per-step candidate leaves
    void *leaves;               // num_bodies x num_colliders Leaf: compacted colliding list
    float4 *face_scratch;       // pressure scratch, one slot per face of the bundle's pressure bodies
    uint32_t has_pressure;
    uint32_t has_sleepers;
    uint32_t iterations;
    uint32_t staged;            // records + level offsets copied into LDS once per launch (256-thread launches)
    uint32_t uniform;           // 12 B LDS positions: every vertex has the inverse mass uniform_inv_mass
    float uniform_inv_mass;
    unsigned long long *elapsed;    // where the workgroup stores its duration (realtime-clock ticks), or null
    const float4 *skin;         // skinned vertex states of the context (the program's skin records index it)
};

} // namespace xpbd

namespace xpbd {

// Differentiable rollout: one workgroup per body.  Records of the body's
// level program carry {a: v0 | type | v1, b: edge index in the body, z: rest
// length}; the compliance comes from the parameter array.
struct RolloutBodyDev
{
    const uint4 *recs;
    const uint32_t *lvl;        // nlev + 1 absolute offsets
    uint32_t nlev, num_records;
    uint32_t nv, num_edges;
    uint32_t voff, eoff;        // first vertex / edge in the concatenated arrays
    uint32_t rbase;             // first record slot in the record tape
    const float4 *x0;           // initial positions + inverse mass
    // pressure (num_faces 0: none): faces (3 vertices each) and vertex -> face lists in face order
    const uint32_t *faces, *vf_begin, *vf_list;
    uint32_t num_faces;
    float pressure;
    uint32_t edges_only;        // 1: every record is an edge (the level passes skip the type dispatch)
};

struct RolloutDev
{
    const RolloutBodyDev *bodies;
    uint32_t num_bodies, iterations, frames, interval, total_nv;
    float dt, linear_damping, max_linear_velocity;
    float gravity[3];           // world gravity (bodies have identity rotation, so also body-local)
    float gravity_factor;
    float4 *ckpt_x, *ckpt_v;    // (x + inverse mass, v) at frames 0, interval, 2 interval, ...
    float4 *seg_x, *seg_v;      // state at every sub-step start of the segment being reversed
    float4 *vel, *prev;         // float forward state
    double *vel64, *prev64;     // double forward state (4 per vertex)
    float4 *rec_tape;           // 4 per record: pre-projection positions (+ inverse mass) of its vertices
    float4 *adj_x, *adj_v;      // adjoint state
};

} // namespace xpbd

