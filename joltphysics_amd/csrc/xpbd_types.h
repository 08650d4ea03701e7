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
// Vertex numbers are indices into the LDS window of the body (< 16384).
enum : uint32_t { kTypeVolume = 0, kTypeBend = 1, kTypeEdge = 2, kTypeLRA = 3 };
constexpr uint32_t kMaxWindowVertices = 1u << 14;

struct SettingsDev
{
    const uint4 *recs;          // records ordered (group, level, sequence)
    const uint32_t *lvl;        // absolute record offset of each level start, nlev+1 per group
    const uint2 *groups;        // {index of first lvl entry, num levels}; parallel groups then the tail
    const uint32_t *faces;      // 3 vertex indices per face (pressure)
    const uint32_t *vf_begin;   // vertex -> face CSR, faces in increasing order (pressure)
    const uint32_t *vf_list;
    uint32_t num_vertices;
    uint32_t num_parallel;      // number of parallel groups (tail group index = num_parallel)
    uint32_t num_faces;
    uint32_t tail_one_wave;     // 1 when every tail level has <= 64 constraints
};

struct BodyDev
{
    SettingsDev s;
    float4 *x;                  // position xyz + inverse mass (body-local)
    float4 *v;                  // velocity
    float4 *prev;               // previous position
    float4 *plane;              // collision plane normal + constant
    int32_t *idx;               // colliding shape index (-1 none)
    int32_t *contact;           // has-contact flag
    uint32_t iterations;
    float linear_damping, max_linear_velocity, restitution, friction, pressure, gravity_factor, vertex_radius;
    int32_t update_position, allow_sleeping;
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

struct ColliderDev
{
    int32_t type;
    float params[5];
    float pos[3];
    float rot[4];
    float friction, restitution;
    float bounds[6];
};

} // namespace xpbd
