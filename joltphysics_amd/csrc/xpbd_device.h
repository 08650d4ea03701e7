// Device functions shared by the update kernels: constraint projections,
// collider leaves and the vertex-vs-shape plane search, all in the reference's
// float operation order (jolt_math.h).
#pragma once

#include <cfloat>

#include "jolt_math.h"
#include "jolt_xpbd.h"
#include "xpbd_types.h"

namespace xpbd {

constexpr int kWave = 64;

// Collider resolved into the soft body's local space for one collision step
// (the reference's LeafShape + CollidingShape, SoftBodyMotionProperties.h).
struct Leaf
{
    M44 t;          // leaf transform (soft-body local)
    M44 inv;        // InversedRotationTranslation of t (box, capsule)
    float pn[3];    // plane shape: transformed plane normal | tapered capsule {bottom radius, sin, tan alpha} |
                    // tapered cylinder {bottom radius} | scaled hull, triangle tree: scale
    float pc;       // plane shape: constant | tapered capsule: sign of the y scale | scaled hull: -1 for an
                    // inside-out scale, else 1 | triangle tree: normal sign
    float a, b, c;  // sphere {radius} | box half extent | capsule, cylinder {half height, radius} |
                    // tapered capsule {top centre, bottom centre, top radius} (scaled) |
                    // tapered cylinder {top, bottom, top radius} (scaled) | scaled hull: 1 / scale
    int32_t type;
    float friction, restitution;
    union
    {
        const HullDev *hull;    // convex hull tables (device memory)
        const TreeDev *tree;    // triangle tree (device memory)
    };
    int32_t dyn;            // dynamic collider: rank into the body's DynStateDev row; -1 otherwise
    int32_t collider;       // index of the collider in the context's list
    float inv_mass_scale1;  // SoftBodyContactSettings::mInvMassScale1 of the pair (CollidingShape::mSoftBodyInvMassScale)
};

// Per-pair contact settings {mInvMassScale1, mInvMassScale2, mInvInertiaScale2, accept}
// (SoftBodyContactListener.h:19-28); without a table every pair is accepted with unit scales.
__device__ __forceinline__ float4 pair_settings(const float4 *row, uint32_t collider)
{
    return row ? row[collider] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
}
// DetermineCollidingShapes with a contact listener (SoftBodyMotionProperties.cpp:160-171): a
// rejected pair is skipped, and so is a colliding pair without any interaction
__device__ __forceinline__ bool pair_skipped(float4 pair, bool sensor, bool dynamic)
{
    return pair.w == 0.0f || (!sensor && pair.x == 0.0f && (!dynamic || pair.y == 0.0f));
}

__device__ __forceinline__ V3 xyz(float4 p) { return v3(p.x, p.y, p.z); }
__device__ __forceinline__ float4 f4(V3 p, float w) { return make_float4(p.x, p.y, p.z, w); }

// Positions in LDS.  A float4* holds position + inverse mass per vertex (16 B);
// PositionsF3 holds three floats per vertex (12 B) for bodies whose vertices all
// share one inverse mass.  pget / pset give both the same float4 interface.
struct PositionsF3
{
    float *x;
    float w;
};

__device__ __forceinline__ float4 pget(float4 *X, uint32_t i) { return X[i]; }
__device__ __forceinline__ void pset(float4 *X, uint32_t i, float4 p) { X[i] = p; }
__device__ __forceinline__ float4 pget(const PositionsF3 &X, uint32_t i)
{
    const float *q = X.x + 3u * i;
    return make_float4(q[0], q[1], q[2], X.w);
}
__device__ __forceinline__ void pset(const PositionsF3 &X, uint32_t i, float4 p)
{
    float *q = X.x + 3u * i;
    q[0] = p.x;
    q[1] = p.y;
    q[2] = p.z;
}

// ---------------------------------------------------------------- constraints
// ApplyEdgeConstraints body (SoftBodyMotionProperties.cpp:583-600)
template <class P>
__device__ __forceinline__ void project_edge(const P &X, uint32_t i0, uint32_t i1, float rest, float compliance, float inv_dt_sq)
{
    float4 p0 = pget(X, i0), p1 = pget(X, i1);
    V3 x0 = xyz(p0), x1 = xyz(p1);
    float w0 = p0.w, w1 = p1.w;
    V3 delta = x1 - x0;
    float length = len(delta);
    float denom = length * ((w0 + w1) + compliance * inv_dt_sq);
    if (denom < 1.0e-12f)
        return;
    V3 correction = (delta * (length - rest)) / denom;
    pset(X, i0, f4(x0 + w0 * correction, w0));
    pset(X, i1, f4(x1 - w1 * correction, w1));
}

// ApplyDihedralBendConstraints body (SoftBodyMotionProperties.cpp:375-455)
template <class P>
__device__ __forceinline__ void project_bend(const P &X, uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, float compliance, float initial_angle, float inv_dt_sq)
{
    float4 p0 = pget(X, i0), p1 = pget(X, i1), p2 = pget(X, i2), p3 = pget(X, i3);
    V3 x0 = xyz(p0), x1 = xyz(p1), x2 = xyz(p2), x3 = xyz(p3);
    V3 e = x1 - x0;
    float e_len = len(e);
    if (e_len < 1.0e-6f)
        return;
    V3 x1x2 = x2 - x1;
    V3 x1x3 = x3 - x1;
    V3 n1 = cross(x2 - x0, x1x2);
    V3 n2 = cross(x1x3, x3 - x0);
    float n1_len_sq = len_sq(n1);
    float n2_len_sq = len_sq(n2);
    float n1n2 = n1_len_sq * n2_len_sq;
    if (n1n2 < 1.0e-24f)
        return;
    float sign = sign_of(dot(cross(n2, n1), e));
    float d = dot(n1, n2) / sqrtf(n1n2);
    float c = sign * acos_approx(d) - initial_angle;
    if (c > kPi)
        c -= 2.0f * kPi;
    else if (c < -kPi)
        c += 2.0f * kPi;
    n1 = n1 / n1_len_sq;
    n2 = n2 / n2_len_sq;
    V3 d0c = (dot(x1x2, e) * n1 + dot(x1x3, e) * n2) / e_len;
    V3 d2c = e_len * n1;
    V3 d3c = e_len * n2;
    V3 d1c = (neg(d0c) - d2c) - d3c;
    float w0 = p0.w, w1 = p1.w, w2 = p2.w, w3 = p3.w;
    float denom = (((w0 * len_sq(d0c) + w1 * len_sq(d1c)) + w2 * len_sq(d2c)) + w3 * len_sq(d3c)) + compliance * inv_dt_sq;
    if (denom < 1.0e-12f)
        return;
    float minus_lambda = c / denom;
    pset(X, i0, f4(x0 - (minus_lambda * w0) * d0c, w0));
    pset(X, i1, f4(x1 - (minus_lambda * w1) * d1c, w1));
    pset(X, i2, f4(x2 - (minus_lambda * w2) * d2c, w2));
    pset(X, i3, f4(x3 - (minus_lambda * w3) * d3c, w3));
}

// The same projection shared by two adjacent lanes (lanes 2k, 2k + 1 hold the two
// halves of one record; `second` tells them apart).  Every value is computed with the
// reference's operations in the reference's order, the two normals, the two
// normalisations, two of the three divisions of d0 and the squared lengths of the
// gradients on one lane each, exchanged with the partner lane by a DPP quad
// permutation; each lane then moves two of the four vertices.  Both lanes take every
// early return together, since they decide on the same values.
__device__ __forceinline__ float swap_pair(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xb1, 0xf, 0xf, true));   // quad_perm [1,0,3,2]
}
__device__ __forceinline__ V3 swap_pair(V3 v) { return v3(swap_pair(v.x), swap_pair(v.y), swap_pair(v.z)); }
__device__ __forceinline__ V3 select(bool c, V3 a, V3 b) { return v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }

template <class P>
__device__ __forceinline__ void project_bend_pair(const P &X, uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, float compliance,
                                                  float initial_angle, float inv_dt_sq, bool second)
{
    const float4 p0 = pget(X, i0), p1 = pget(X, i1), p2 = pget(X, i2), p3 = pget(X, i3);
    const V3 x0 = xyz(p0), x1 = xyz(p1), x2 = xyz(p2), x3 = xyz(p3);
    const V3 e = x1 - x0;
    const float e_len = len(e);
    if (e_len < 1.0e-6f)
        return;
    const V3 x1x2 = x2 - x1;
    const V3 x1x3 = x3 - x1;
    // this lane's normal: n1 = (x2 - x0) x x1x2 on the first, n2 = x1x3 x (x3 - x0) on the second
    const V3 nm = cross(select(second, x1x3, x2 - x0), select(second, x3 - x0, x1x2));
    const float nm_len_sq = len_sq(nm);
    const V3 no = swap_pair(nm);
    const float no_len_sq = swap_pair(nm_len_sq);
    V3 n1 = select(second, no, nm), n2 = select(second, nm, no);
    const float n1_len_sq = second ? no_len_sq : nm_len_sq, n2_len_sq = second ? nm_len_sq : no_len_sq;
    const float n1n2 = n1_len_sq * n2_len_sq;
    if (n1n2 < 1.0e-24f)
        return;
    const float sign = sign_of(dot(cross(n2, n1), e));
    const float d = dot(n1, n2) / sqrtf(n1n2);
    float c = sign * acos_approx(d) - initial_angle;
    if (c > kPi)
        c -= 2.0f * kPi;
    else if (c < -kPi)
        c += 2.0f * kPi;
    // n1 / |n1|^2 on the first lane, n2 / |n2|^2 on the second
    const V3 um = nm / nm_len_sq;
    const V3 uo = swap_pair(um);
    n1 = select(second, uo, um);
    n2 = select(second, um, uo);
    // dot(x1x2, e) on the first lane, dot(x1x3, e) on the second
    const float dm = dot(select(second, x1x3, x1x2), e);
    const float dother = swap_pair(dm);
    const V3 t = (second ? dother : dm) * n1 + (second ? dm : dother) * n2;
    // d0 = t / |e|: x on the first lane, z on the second, y on both
    const float qm = (second ? t.z : t.x) / e_len;
    const float qy = t.y / e_len;
    const float qo = swap_pair(qm);
    const V3 d0c = v3(second ? qo : qm, qy, second ? qm : qo);
    const V3 d2c = e_len * n1;
    const V3 d3c = e_len * n2;
    const V3 d1c = (neg(d0c) - d2c) - d3c;
    const float w0 = p0.w, w1 = p1.w, w2 = p2.w, w3 = p3.w;
    // |d0|^2, |d2|^2 on the first lane, |d1|^2, |d3|^2 on the second
    const V3 da = select(second, d1c, d0c), db = select(second, d3c, d2c);
    const float la = len_sq(da), lb = len_sq(db);
    const float lao = swap_pair(la), lbo = swap_pair(lb);
    const float l0 = second ? lao : la, l1 = second ? la : lao, l2 = second ? lbo : lb, l3 = second ? lb : lbo;
    const float denom = (((w0 * l0 + w1 * l1) + w2 * l2) + w3 * l3) + compliance * inv_dt_sq;
    if (denom < 1.0e-12f)
        return;
    const float minus_lambda = c / denom;
    // the first lane moves x0 and x2, the second x1 and x3
    const float wa = second ? w1 : w0, wb = second ? w3 : w2;
    pset(X, second ? i1 : i0, f4(select(second, x1, x0) - (minus_lambda * wa) * da, wa));
    pset(X, second ? i3 : i2, f4(select(second, x3, x2) - (minus_lambda * wb) * db, wb));
}

// ApplyVolumeConstraints body (SoftBodyMotionProperties.cpp:467-506)
template <class P>
__device__ __forceinline__ void project_volume(const P &X, uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4, float six_rest_volume, float compliance, float inv_dt_sq)
{
    float4 p1 = pget(X, i1), p2 = pget(X, i2), p3 = pget(X, i3), p4 = pget(X, i4);
    V3 x1 = xyz(p1), x2 = xyz(p2), x3 = xyz(p3), x4 = xyz(p4);
    V3 x1x2 = x2 - x1;
    V3 x1x3 = x3 - x1;
    V3 x1x4 = x4 - x1;
    float c = fabsf(dot(cross(x1x2, x1x3), x1x4)) - six_rest_volume;
    V3 d1c = cross(x4 - x2, x3 - x2);
    V3 d2c = cross(x1x3, x1x4);
    V3 d3c = cross(x1x4, x1x2);
    V3 d4c = cross(x1x2, x1x3);
    float w1 = p1.w, w2 = p2.w, w3 = p3.w, w4 = p4.w;
    float denom = (((w1 * len_sq(d1c) + w2 * len_sq(d2c)) + w3 * len_sq(d3c)) + w4 * len_sq(d4c)) + compliance * inv_dt_sq;
    if (denom < 1.0e-12f)
        return;
    float minus_lambda = c / denom;
    pset(X, i1, f4(x1 - (minus_lambda * w1) * d1c, w1));
    pset(X, i2, f4(x2 - (minus_lambda * w2) * d2c, w2));
    pset(X, i3, f4(x3 - (minus_lambda * w3) * d3c, w3));
    pset(X, i4, f4(x4 - (minus_lambda * w4) * d4c, w4));
}

// ApplyLRAConstraints body (SoftBodyMotionProperties.cpp:685-697): only vertex 1 moves
template <class P>
__device__ __forceinline__ void project_lra(const P &X, uint32_t i0, uint32_t i1, float max_distance)
{
    V3 x0 = xyz(pget(X, i0));
    float4 p1 = pget(X, i1);
    V3 delta = xyz(p1) - x0;
    float delta_len_sq = len_sq(delta);
    if (delta_len_sq > max_distance * max_distance)
        pset(X, i1, f4(x0 + (delta * max_distance) / sqrtf(delta_len_sq), p1.w));
}

// ApplySkinConstraints body for one skinned vertex (SoftBodyMotionProperties.cpp:509-571).
// The vertex's skin state is entry g of `skin`, four float4: previous and current
// skinned position, normal, and {max distance multiplier, back stop radius, skin
// constraints enabled, previous position valid}.  `factor` = (iteration + 1) /
// iterations, the reference's mNextIteration / mNumIterations.
struct RecordCtx
{
    const float4 *skin;
    uint32_t base;              // added to the record's entry index
    float factor;
    // Cosserat rods of the body (null without rods): states (2 float4 per rod) and the
    // settings' rod and bend-twist constants
    float4 *rods;
    const float4 *rod_const;
    const float4 *bt_const;
};

template <class P>
__device__ __forceinline__ void project_skin(const P &X, uint32_t i, uint32_t g, float max_distance_setting,
                                             float back_stop_distance, const RecordCtx &sk)
{
    const float4 *e = sk.skin + 4u * (size_t)g;
    const float4 q = e[3];
    if (q.z == 0.0f)
        return;                 // mEnableSkinConstraints false
    const float factor = q.w != 0.0f ? sk.factor : 1.0f;
    const float prev_factor = 1.0f - factor;
    const float max_distance = max_distance_setting * q.x;
    const V3 skin_pos = prev_factor * xyz(e[0]) + factor * xyz(e[1]);
    const float4 p = pget(X, i);
    V3 x = xyz(p);
    if (max_distance > 0.0f)
    {
        const float radius = q.y;
        if (back_stop_distance < max_distance)
        {
            const V3 normal = xyz(e[2]);
            const V3 center = skin_pos - normal * (back_stop_distance + radius);
            const V3 delta = x - center;
            const float delta_len_sq = len_sq(delta);
            if (delta_len_sq < radius * radius)
            {
                const float delta_len = sqrtf(delta_len_sq);
                x = delta_len > 0.0f ? center + delta * (radius / delta_len) : center + normal * radius;
            }
        }
        if (max_distance < FLT_MAX)
        {
            const V3 delta = x - skin_pos;
            const float delta_len_sq = len_sq(delta);
            const float max_distance_sq = max_distance * max_distance;
            if (delta_len_sq > max_distance_sq)
                x = skin_pos + delta * sqrtf(max_distance_sq / delta_len_sq);
        }
    }
    else
        x = skin_pos;
    pset(X, i, f4(x, p.w));
}

// ApplyRodStretchShearConstraints body for rod r between vertices i0, i1
// (SoftBodyMotionProperties.cpp:603-635; Kugelstadt and Schoemer 2016, eq. 37).
template <class P>
__device__ __forceinline__ void project_rod_stretch_shear(const P &X, uint32_t i0, uint32_t i1, uint32_t r, float inv_dt_sq,
                                                          const RecordCtx &c)
{
    const float4 k = c.rod_const[r];                // length, inverse mass, compliance
    const float4 p0 = pget(X, i0), p1 = pget(X, i1);
    const float denom = ((p0.w + p1.w) + (4.0f * k.y) * (k.x * k.x)) + k.z * inv_dt_sq;
    if (denom < 1.0e-12f)
        return;
    const float4 s = c.rods[2u * r];
    Q4 rotation = q4(s.x, s.y, s.z, s.w);
    const V3 x0 = xyz(p0), x1 = xyz(p1);
    const V3 delta = ((x1 - x0) - rotate_axis_z(rotation) * k.x) / denom;
    pset(X, i0, f4(x0 + p0.w * delta, p0.w));
    pset(X, i1, f4(x1 - p1.w * delta, p1.w));
    // q * e3_bar = [-qy, qx, -qw, qz]
    const Q4 q_e3_bar = q4(-rotation.y, rotation.x, -rotation.w, rotation.z);
    rotation = rotation + mul_imaginary(delta, q_e3_bar) * ((2.0f * k.y) * k.x);
    const Q4 n = normalized(rotation);
    c.rods[2u * r] = make_float4(n.x, n.y, n.z, n.w);
}

// ApplyRodBendTwistConstraints body for bend-twist constraint b
// (SoftBodyMotionProperties.cpp:637-677; Kugelstadt and Schoemer 2016, eq. 40).
__device__ __forceinline__ void project_rod_bend_twist(uint32_t b, float inv_dt_sq, const RecordCtx &c)
{
    const float4 o = c.bt_const[2u * b], kb = c.bt_const[2u * b + 1u];
    const uint32_t r1 = __float_as_uint(kb.x), r2 = __float_as_uint(kb.y);
    const float denom = (c.rod_const[r1].y + c.rod_const[r2].y) + kb.z * inv_dt_sq;
    if (denom < 1.0e-12f)
        return;
    const float4 s1 = c.rods[2u * r1], s2 = c.rods[2u * r2];
    Q4 rotation1 = q4(s1.x, s1.y, s1.z, s1.w), rotation2 = q4(s2.x, s2.y, s2.z, s2.w);
    const Q4 omega = quat_mul(conjugated(rotation1), rotation2);
    const Q4 omega0 = q4(o.x, o.y, o.z, o.w);
    const Q4 minus = omega - omega0, plus = omega + omega0;
    // the shorter of the two rotations
    Q4 delta_omega = dot4(plus, plus) < dot4(minus, minus) ? plus : minus;
    delta_omega = delta_omega / denom;
    delta_omega.w = 0.0f;       // the Darboux vector's real part does not vanish
    const Q4 delta_rod2 = quat_mul(rotation1 * c.rod_const[r2].y, delta_omega);
    rotation1 = rotation1 + quat_mul(rotation2 * c.rod_const[r1].y, delta_omega);
    rotation2 = rotation2 - delta_rod2;
    const Q4 n1 = normalized(rotation1), n2 = normalized(rotation2);
    c.rods[2u * r1] = make_float4(n1.x, n1.y, n1.z, n1.w);
    c.rods[2u * r2] = make_float4(n2.x, n2.y, n2.z, n2.w);
}

// Padding record: skipped (levels pad each constraint type to whole wavefronts)
constexpr uint32_t kNopRecord = 0xffffffffu;

// A record of LRA type whose two vertices are equal is a skinned constraint: vertex 0,
// y = its skin entry (before `base`), z = max distance, w = back stop distance.  Bits 31
// and 30 of y mark rod stretch-shear and bend-twist records (their states live in
// global memory, so their levels end in sync_global()).
template <bool RODS, class P>
__device__ __forceinline__ void project_record(const P &X, uint4 r, float inv_dt_sq, const RecordCtx &sk)
{
    if (r.x == kNopRecord)
        return;
    uint32_t type = (r.x >> 14) & 3u;
    uint32_t i0 = r.x & 0x3fffu, i1 = r.x >> 16, i2 = r.y & 0x3fffu, i3 = r.y >> 16;
    float p0 = __uint_as_float(r.z), p1 = __uint_as_float(r.w);
    switch (type)
    {
    case kTypeEdge: project_edge(X, i0, i1, p0, p1, inv_dt_sq); break;
    case kTypeBend: project_bend_pair(X, i0, i1, i2, i3, p0, p1, inv_dt_sq, ((r.y >> 15) & 1u) != 0); break;
    case kTypeVolume: project_volume(X, i0, i1, i2, i3, p0, p1, inv_dt_sq); break;
    default:
        if (RODS && (r.y & kRodStretchShear))
            project_rod_stretch_shear(X, i0, i1, r.y & kRodIndexMask, inv_dt_sq, sk);
        else if (RODS && (r.y & kRodBendTwist))
            project_rod_bend_twist(r.y & kRodIndexMask, inv_dt_sq, sk);
        else if (i0 == i1)
            project_skin(X, i0, r.y + sk.base, p0, p1, sk);
        else
            project_lra(X, i0, i1, p0);
        break;
    }
}

// ---------------------------------------------------------------- level programs
// A level program is a sequence of levels; the records of one level write
// disjoint vertices and run concurrently, one record per thread, with a
// workgroup barrier between levels.
typedef uint32_t RecordVec __attribute__((ext_vector_type(4)));

// s_waitcnt immediate (gfx9 encoding): vmcnt 0, expcnt 7, lgkmcnt 15.
constexpr int kWaitVmcntZero = 0x0f70;

// Workgroup barrier for data that the wavefronts of a workgroup hand to each
// other through global memory (the level programs exchange data through LDS
// only): a release fence, the barrier and an acquire fence, all at workgroup
// scope on the global address space, which the compiler's memory model orders
// (every wavefront of a workgroup runs on one CU and shares its vector L1, so
// the fences need no cache maintenance outside threadgroup-split mode, which
// these kernels do not use).  Each wavefront also waits until its own global
// stores have completed before the barrier, so the hand-off does not depend on
// the order in which the CU's memory pipeline completes a store and a later
// load of another wavefront.
__device__ __forceinline__ void sync_global()
{
    __builtin_amdgcn_s_waitcnt(kWaitVmcntZero);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "global");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "global");
}

// Record load from global memory through an address-space-1 pointer, so the
// load is a global (not flat) load: only vmcnt counts it and it can stay in
// flight across the LDS-only barriers between levels.
__device__ __forceinline__ RecordVec load_record(const uint4 *recs, uint32_t i)
{
    return *(const __attribute__((address_space(1))) RecordVec *)(const void *)(recs + i);
}

// Level offsets are uniform: scalar loads through the constant address space.
__device__ __forceinline__ uint32_t load_level_offset(const uint32_t *lvl, uint32_t i)
{
    return *(const __attribute__((address_space(4))) uint32_t *)(const void *)(lvl + i);
}

// Values every lane of the workgroup holds alike, moved to scalar registers
// (the compiler cannot always prove it for values read through pointers).
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = (uint64_t)(uintptr_t)p;
    const uint64_t lo = uniform((uint32_t)v), hi = uniform((uint32_t)(v >> 32));
    return (T *)(uintptr_t)((hi << 32) | lo);
}

__device__ __forceinline__ uint4 as_uint4(RecordVec r) { return make_uint4(r.x, r.y, r.z, r.w); }

// Marks the registers of a prefetched record as defined here, so that no use
// of the record is scheduled above the barrier of the previous level (which
// would make that level wait for a load meant for a later one).
__device__ __forceinline__ void pin_record(RecordVec &r) { asm volatile("" : "+v"(r)); }

struct NoSkip
{
    __device__ __forceinline__ bool operator()(const uint4 &) const { return false; }
};

// Level program with records and offsets in global memory, forward or in
// reverse level order.  Thread t's record of each level is prefetched three
// levels ahead into a ring of three registers (the loop is unrolled three
// times so the ring never moves); records beyond the first BLOCK of a level are
// loaded when needed.  `handle(record, index)` processes one record (no-op records
// included) and its index; a level's records are always assigned to threads the same way,
// whatever the direction, so per-record data written by a thread in one pass is
// read back by the same thread in another.
template <int BLOCK, bool REVERSE, class Handle>
__device__ __forceinline__ void run_program_global(const uint4 *recs, const uint32_t *lvl, uint32_t nlev, uint32_t tid,
                                                   Handle handle, bool gsync = false)
{
    recs = uniform_ptr(recs);
    lvl = uniform_ptr(lvl);
    nlev = uniform(nlev);
    if (nlev == 0)
        return;
    const uint32_t first = load_level_offset(lvl, 0), last = load_level_offset(lvl, nlev);
    if (first == last)
    {
        __syncthreads();
        return;
    }
    const uint32_t max_index = last - 1;
    // Drain the vector-memory counter once here: with global stores of an
    // earlier phase still pending, the compiler treats the counter as out of
    // order and would wait for every load at every level.
    __builtin_amdgcn_s_waitcnt(kWaitVmcntZero);
    // level of step k (k >= nlev: an empty level past the end)
    auto level_of = [&](uint32_t k) -> uint32_t { return k < nlev ? (REVERSE ? nlev - 1 - k : k) : nlev; };
    auto fetch = [&](uint32_t k) -> RecordVec {
        const uint32_t lo = load_level_offset(lvl, level_of(k));
        const uint32_t i = lo + tid;
        return load_record(recs, i < max_index ? i : max_index);    // validity is checked at use
    };
    RecordVec r0 = fetch(0);
    asm volatile("" ::: "memory");
    RecordVec r1 = fetch(1);
    asm volatile("" ::: "memory");
    RecordVec r2 = fetch(2);
    auto level = [&](uint32_t k, RecordVec &r) {
        const uint32_t L = level_of(k);
        const uint32_t lo = load_level_offset(lvl, L);
        const uint32_t hi = L < nlev ? load_level_offset(lvl, L + 1) : lo;
        pin_record(r);
        if (lo + tid < hi)
            handle(as_uint4(r), lo + tid);
        for (uint32_t i = lo + tid + BLOCK; i < hi; i += BLOCK)
            handle(as_uint4(load_record(recs, i)), i);
        r = fetch(k + 3);
        if (gsync)
            sync_global();
        else
            __syncthreads();
    };
    for (uint32_t k = 0; k < nlev; k += 3)
    {
        level(k, r0);
        level(k + 1, r1);
        level(k + 2, r2);
    }
}

// RODS: the program may hold rod records (compiled in only where bodies with rods run)
template <int BLOCK, bool RODS = false, class P, class Skip>
__device__ __forceinline__ void run_levels_global(const uint4 *recs, const uint32_t *lvl, uint32_t nlev, const P &X,
                                                  float inv_dt_sq, uint32_t tid, Skip skip, const RecordCtx &sk,
                                                  bool gsync = false)
{
    run_program_global<BLOCK, false>(recs, lvl, nlev, tid, [&](const uint4 &q, uint32_t) {
        if (!skip(q))
            project_record<RODS>(X, q, inv_dt_sq, sk);
    }, gsync);
}

// Level program with records and offsets in LDS: the first record of the next
// level is read before the barrier (the barrier's LDS wait covers it).
template <int BLOCK, bool RODS = false, class P, class Skip>
__device__ __forceinline__ void run_levels_lds(const uint4 *recs, const uint32_t *lvl, uint32_t nlev, const P &X,
                                               float inv_dt_sq, uint32_t tid, Skip skip, const RecordCtx &sk,
                                               bool gsync = false)
{
    if (nlev == 0)
        return;
    const uint4 nop = make_uint4(kNopRecord, 0u, 0u, 0u);
    uint32_t lo = lvl[0], hi = lvl[1];
    uint4 r = lo + tid < hi ? recs[lo + tid] : nop;
    for (uint32_t L = 0; L < nlev; ++L)
    {
        const uint32_t next_hi = lvl[L + 2 <= nlev ? L + 2 : nlev];
        if (lo + tid < hi && !skip(r))
            project_record<RODS>(X, r, inv_dt_sq, sk);
        for (uint32_t i = lo + tid + BLOCK; i < hi; i += BLOCK)
        {
            const uint4 q = recs[i];
            if (!skip(q))
                project_record<RODS>(X, q, inv_dt_sq, sk);
        }
        const uint4 rn = hi + tid < next_hi ? recs[hi + tid] : nop;
        if (gsync)
            sync_global();
        else
            __syncthreads();
        r = rn;
        lo = hi;
        hi = next_hi;
    }
}

// ---------------------------------------------------------------- collision
// Shape::CollideSoftBodyVertices for plane / sphere / box / capsule / cylinder /
// tapered capsule / convex hull with the argmax rule of
// CollideSoftBodyVertexIterator::UpdatePenetration (CollideSoftBodyVertexIterator.h:83-97):
// a leaf replaces the stored plane unless the stored penetration is >= the new one.
// Vec3::NormalizedOr (Vec3.inl:978-1000): zero_value when |v|^2 <= FLT_MIN
__device__ __forceinline__ V3 normalized_or(V3 v, V3 zero_value)
{
    const float l2 = len_sq(v);
    return l2 <= FLT_MIN ? zero_value : v / sqrtf(l2);
}

// ClosestPoint::GetClosestPointOnLine (ClosestPoint.h:129-151) with
// GetBaryCentricCoordinates for a line (ClosestPoint.h:15-41): closest point to
// the origin on segment (a, b); a degenerate segment picks the nearer end.
__device__ __forceinline__ V3 closest_point_on_line(V3 a, V3 b)
{
    const V3 ab = b - a;
    const float denominator = len_sq(ab);
    if (denominator < FLT_EPSILON * FLT_EPSILON)
        return len_sq(a) < len_sq(b) ? a : b;
    const float v = -dot(a, ab) / denominator;
    const float u = 1.0f - v;
    if (v <= 0.0f)
        return a;
    if (u <= 0.0f)
        return b;
    return u * a + v * b;
}

// ConvexHullShape::CollideSoftBodyVertices for one vertex, unscaled
// (ConvexHullShape.cpp:1060-1164): the most facing plane, then, outside the
// hull, the closest point on that face's edges that the vertex lies beyond.
__device__ __forceinline__ void collide_hull(const Leaf &L, uint32_t l, V3 x, float &largest, int32_t &idx, Pl &plane)
{
    const HullDev &H = *L.hull;
    const V3 local = mul_point(L.inv, x);
    float max_distance = -FLT_MAX;
    V3 max_plane_normal = v3(0.0f, 0.0f, 0.0f);
    uint32_t max_plane_idx = 0;
    for (uint32_t f = 0; f < H.num_faces; ++f)
    {
        const float4 pf = H.planes[f];
        Pl p; p.n = xyz(pf); p.c = pf.w;
        const float distance = signed_distance(p, local);
        if (distance > max_distance)
        {
            max_distance = distance;
            max_plane_normal = p.n;
            max_plane_idx = f;
        }
    }
    V3 closest_point = -max_distance * max_plane_normal;
    const bool is_outside = max_distance > 0.0f;
    if (is_outside)
    {
        float closest_point_dist_sq = FLT_MAX;
        const uint32_t *vi = H.index + H.face_first[max_plane_idx];
        const uint32_t count = H.face_count[max_plane_idx];
        for (uint32_t k = 0; k < count; ++k)
        {
            const V3 p1 = xyz(H.points[vi[k]]);
            const V3 p2 = xyz(H.points[vi[k + 1 == count ? 0 : k + 1]]);
            const V3 edge_normal = cross(p2 - p1, max_plane_normal);
            if (dot(edge_normal, local - p1) > 0.0f)
            {
                const V3 closest = closest_point_on_line(p1 - local, p2 - local);
                const float distance_sq = len_sq(closest);
                if (distance_sq < closest_point_dist_sq)
                {
                    closest_point_dist_sq = distance_sq;
                    closest_point = closest;
                }
            }
        }
    }
    V3 normal = neg(closest_point);
    const float normal_length = len(normal);
    float penetration = normal_length;
    if (is_outside)
        penetration = -penetration;
    else
        normal = neg(normal);
    if (!(largest >= penetration))
    {
        largest = penetration;
        normal = normal_length > 1.0e-12f ? normal / normal_length : max_plane_normal;
        plane = plane_transformed(plane_from_point_normal(local + closest_point, normal), L.t);
        idx = (int32_t)l;
    }
}

// ConvexHullShape::CollideSoftBodyVertices for one vertex, scaled (ConvexHullShape.cpp:1060-
// 1164, the `is_not_scaled` false branch): each plane rebuilt from the scaled normal and the
// scaled first point of its face; edges from scaled points, tested with the inside-out sign.
// L.pn: scale, (L.a, L.b, L.c): 1 / scale, L.pc: -1 for an inside-out scale, else 1.
__device__ __noinline__ void collide_hull_scaled(const Leaf &L, uint32_t l, V3 x, float &largest, int32_t &idx, Pl &plane)
{
    const HullDev &H = *L.hull;
    const V3 scale = v3(L.pn[0], L.pn[1], L.pn[2]), inv_scale = v3(L.a, L.b, L.c);
    const float scale_flip = L.pc;
    const V3 local = mul_point(L.inv, x);
    float max_distance = -FLT_MAX;
    V3 max_plane_normal = v3(0.0f, 0.0f, 0.0f);
    uint32_t max_plane_idx = 0;
    for (uint32_t f = 0; f < H.num_faces; ++f)
    {
        const V3 n = mulv(inv_scale, xyz(H.planes[f]));
        const V3 plane_normal = n / len(n);                                         // Vec3::Normalized
        const V3 plane_point = mulv(scale, xyz(H.points[H.index[H.face_first[f]]]));
        const float distance = dot(plane_normal, local - plane_point);
        if (distance > max_distance)
        {
            max_distance = distance;
            max_plane_normal = plane_normal;
            max_plane_idx = f;
        }
    }
    V3 closest_point = -max_distance * max_plane_normal;
    const bool is_outside = max_distance > 0.0f;
    if (is_outside)
    {
        float closest_point_dist_sq = FLT_MAX;
        const uint32_t *vi = H.index + H.face_first[max_plane_idx];
        const uint32_t count = H.face_count[max_plane_idx];
        for (uint32_t k = 0; k < count; ++k)
        {
            const V3 p1 = mulv(scale, xyz(H.points[vi[k]]));
            const V3 p2 = mulv(scale, xyz(H.points[vi[k + 1 == count ? 0 : k + 1]]));
            const V3 edge_normal = cross(p2 - p1, max_plane_normal);
            if (scale_flip * dot(edge_normal, local - p1) > 0.0f)
            {
                const V3 closest = closest_point_on_line(p1 - local, p2 - local);
                const float distance_sq = len_sq(closest);
                if (distance_sq < closest_point_dist_sq)
                {
                    closest_point_dist_sq = distance_sq;
                    closest_point = closest;
                }
            }
        }
    }
    V3 normal = neg(closest_point);
    const float normal_length = len(normal);
    float penetration = normal_length;
    if (is_outside)
        penetration = -penetration;
    else
        normal = neg(normal);
    if (!(largest >= penetration))
    {
        largest = penetration;
        normal = normal_length > 1.0e-12f ? normal / normal_length : max_plane_normal;
        plane = plane_transformed(plane_from_point_normal(local + closest_point, normal), L.t);
        idx = (int32_t)l;
    }
}

// ClosestPoint::GetClosestPointOnTriangle (ClosestPoint.h:157-318, MustIncludeC false): the
// point of triangle (a, b, c) closest to the origin and the feature it lies on (bit 0: a,
// bit 1: b, bit 2: c; 7: the interior).  Without FMA, CrossPrecise is Cross.
__device__ __forceinline__ float clamp01(float v)
{
    const float lo = v > 0.0f ? v : 0.0f;              // max(v, 0)
    return lo < 1.0f ? lo : 1.0f;                      // min(., 1)
}

__device__ __forceinline__ V3 closest_point_on_triangle(V3 in_a, V3 in_b, V3 in_c, uint32_t &out_set)
{
    const V3 ba = in_a - in_b, bc = in_c - in_b;
    const bool swap_ac = dot(bc, bc) < dot(ba, ba);     // the shorter of ab, cb becomes ab
    const V3 a = swap_ac ? in_c : in_a;
    const V3 c = swap_ac ? in_a : in_c;
    const V3 ab = in_b - a;
    const V3 ac = c - a;
    const V3 n = cross(ab, ac);
    const float n_len_sq = len_sq(n);
    if (n_len_sq < 1.0e-10f)
    {
        // degenerate: vertices, then edges ac, bc, ab
        uint32_t closest_set = 4u;
        V3 closest_point = in_c;
        float best_dist_sq = len_sq(in_c);
        const float a_len_sq = len_sq(in_a);
        if (a_len_sq < best_dist_sq)
        {
            closest_set = 1u; closest_point = in_a; best_dist_sq = a_len_sq;
        }
        const float b_len_sq = len_sq(in_b);
        if (b_len_sq < best_dist_sq)
        {
            closest_set = 2u; closest_point = in_b; best_dist_sq = b_len_sq;
        }
        const float ac_len_sq = len_sq(ac);
        if (ac_len_sq > FLT_EPSILON * FLT_EPSILON)
        {
            const float v = clamp01(-dot(a, ac) / ac_len_sq);
            const V3 q = a + v * ac;
            const float dist_sq = len_sq(q);
            if (dist_sq < best_dist_sq)
            {
                closest_set = 5u; closest_point = q; best_dist_sq = dist_sq;
            }
        }
        const V3 bc2 = in_c - in_b;
        const float bc_len_sq = len_sq(bc2);
        if (bc_len_sq > FLT_EPSILON * FLT_EPSILON)
        {
            const float v = clamp01(-dot(in_b, bc2) / bc_len_sq);
            const V3 q = in_b + v * bc2;
            const float dist_sq = len_sq(q);
            if (dist_sq < best_dist_sq)
            {
                closest_set = 6u; closest_point = q; best_dist_sq = dist_sq;
            }
        }
        const V3 ab2 = in_b - in_a;
        const float ab_len_sq = len_sq(ab2);
        if (ab_len_sq > FLT_EPSILON * FLT_EPSILON)
        {
            const float v = clamp01(-dot(in_a, ab2) / ab_len_sq);
            const V3 q = in_a + v * ab2;
            const float dist_sq = len_sq(q);
            if (dist_sq < best_dist_sq)
            {
                closest_set = 3u; closest_point = q; best_dist_sq = dist_sq;
            }
        }
        out_set = closest_set;
        return closest_point;
    }
    const V3 ap = neg(a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f)
    {
        out_set = swap_ac ? 4u : 1u;
        return a;
    }
    const V3 bp = neg(in_b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3)
    {
        out_set = 2u;
        return in_b;
    }
    if (d1 * d4 <= d3 * d2 && d1 >= 0.0f && d3 <= 0.0f)
    {
        const float v = d1 / (d1 - d3);
        out_set = swap_ac ? 6u : 3u;
        return a + v * ab;
    }
    const V3 cp = neg(c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6)
    {
        out_set = swap_ac ? 1u : 4u;
        return c;
    }
    if (d5 * d2 <= d1 * d6 && d2 >= 0.0f && d6 <= 0.0f)
    {
        const float w = d2 / (d2 - d6);
        out_set = 5u;
        return a + w * ac;
    }
    const float d4_d3 = d4 - d3, d5_d6 = d5 - d6;
    if (d3 * d6 <= d5 * d4 && d4_d3 >= 0.0f && d5_d6 >= 0.0f)
    {
        const float w = d4_d3 / (d4_d3 + d5_d6);
        out_set = swap_ac ? 3u : 6u;
        return in_b + w * (c - in_b);
    }
    out_set = 7u;
    return (n * dot((a + in_b) + c, n)) / (3.0f * n_len_sq);
}

// Vec4::sSort4Reverse (Vec4.inl:480-502): the compare-exchange network that orders four
// values from high to low, carrying their identifiers; ties keep the network's order.
__device__ __forceinline__ void sort4_reverse(float (&v)[4], uint32_t (&id)[4])
{
    auto exchange = [&](int i, int j) {         // larger of (i, j) to i
        if (v[j] > v[i])
        {
            const float tv = v[i]; v[i] = v[j]; v[j] = tv;
            const uint32_t ti = id[i]; id[i] = id[j]; id[j] = ti;
        }
    };
    exchange(0, 2); exchange(1, 3);
    exchange(0, 1); exchange(2, 3);
    exchange(1, 2);
}

constexpr int kTreeStack = 3 * XPBD_TREE_MAX_DEPTH + 4;

// CollideSoftBodyVerticesVsTriangles (CollideSoftBodyVerticesVsTriangles.h) over a triangle
// tree for one vertex, walked as MeshShape::CollideSoftBodyVertices walks its tree
// (MeshShape.cpp:833-885, NodeCodecQuadTreeHalfFloat.h:245-318): from the root, a node's
// four children get the squared distance of their scaled bounds to the vertex
// (AABox4Scale, AABox4DistanceSqToPoint), are sorted from far to near (SortReverseAndStore)
// and the ones nearer than the closest triangle so far are pushed, nearest on top; a leaf's
// triangles are tested in order, each scaled, and replace the closest one when strictly
// nearer; popped entries are skipped unless still nearer than the closest triangle.  Then
// FinishVertex: an interior closest point gives the triangle's plane (penetration below the
// 0.1 m triangle thickness), an edge or vertex the plane through the closest point facing
// the vertex, back-facing edges ignored.  L.pn: scale, L.pc: normal sign (-1 inside out).
// A height field's hierarchy walks the same way (HeightFieldShape.cpp:1757-2093, 2242-2293).
__device__ __noinline__ void collide_tree(const Leaf &L, uint32_t l, V3 x, float &largest, int32_t &idx, Pl &plane)
{
    const TreeDev &T = *L.tree;
    const V3 scale = v3(L.pn[0], L.pn[1], L.pn[2]);
    const V3 local = mul_point(L.inv, x);
    float closest_dist_sq = FLT_MAX;
    V3 cv0 = v3(0.0f, 0.0f, 0.0f), cv1 = cv0, cv2 = cv0, closest = cv0;
    uint32_t closest_set = 0;
    uint32_t stack[kTreeStack];
    float dist_stack[kTreeStack];
    int top = 0;
    stack[0] = T.root;
    dist_stack[0] = 0.0f;
    do
    {
        const uint32_t ref = stack[top];
        const uint32_t kind = ref & 0xc0000000u;
        if (kind == XPBD_TREE_NODE)
        {
            const TreeNodeDev &N = T.nodes[ref];
            float d[4];
            uint32_t id[4];
            for (int k = 0; k < 4; ++k)
            {
                const float sx0 = scale.x * N.min_x[k], sx1 = scale.x * N.max_x[k];
                const float sy0 = scale.y * N.min_y[k], sy1 = scale.y * N.max_y[k];
                const float sz0 = scale.z * N.min_z[k], sz1 = scale.z * N.max_z[k];
                const float cx = smin(smax(local.x, smin(sx0, sx1)), smax(sx0, sx1));
                const float cy = smin(smax(local.y, smin(sy0, sy1)), smax(sy0, sy1));
                const float cz = smin(smax(local.z, smin(sz0, sz1)), smax(sz0, sz1));
                const float dx = cx - local.x, dy = cy - local.y, dz = cz - local.z;
                d[k] = (dx * dx + dy * dy) + dz * dz;
                // a height field's block without collision (HeightFieldShape.cpp:2268-2269)
                if (T.height_field && !(N.min_y[k] <= N.max_y[k]))
                    d[k] = FLT_MAX;
                id[k] = N.child[k];
            }
            sort4_reverse(d, id);
            int n = 0;
            for (int k = 0; k < 4; ++k)
                n += d[k] < closest_dist_sq ? 1 : 0;
            // the n nearest (the last n after the sort) go on the stack, farthest first
            for (int j = 0; j < n; ++j)
            {
                stack[top + j] = id[4 - n + j];
                dist_stack[top + j] = d[4 - n + j];
            }
            top += n;
        }
        else if (kind == XPBD_TREE_LEAF)
        {
            const uint32_t first = ref & 0x00ffffffu, count = (ref >> 24) & 0x3fu;
            for (uint32_t t = first; t < first + count; ++t)
            {
                const float *tv = T.triangles + 9u * t;
                const V3 v0 = mulv(scale, v3(tv[0], tv[1], tv[2]));
                const V3 v1 = mulv(scale, v3(tv[3], tv[4], tv[5]));
                const V3 v2 = mulv(scale, v3(tv[6], tv[7], tv[8]));
                uint32_t set;
                const V3 cp = closest_point_on_triangle(v0 - local, v1 - local, v2 - local, set);
                const float dist_sq = len_sq(cp);
                if (dist_sq < closest_dist_sq)
                {
                    cv0 = v0; cv1 = v1; cv2 = v2;
                    closest = cp;
                    closest_dist_sq = dist_sq;
                    closest_set = set;
                }
            }
        }
        // the next entry still nearer than the closest triangle
        do
            --top;
        while (top >= 0 && !(dist_stack[top] < closest_dist_sq));
    } while (top >= 0);

    if (!(closest_dist_sq < FLT_MAX))
        return;
    const V3 w0 = mul_point(L.t, cv0), w1 = mul_point(L.t, cv1), w2 = mul_point(L.t, cv2);
    const V3 triangle_normal = L.pc * normalized_or(cross(w1 - w0, w2 - w0), v3(0.0f, 1.0f, 0.0f));
    if (closest_set == 7u)
    {
        const float penetration = dot(triangle_normal, w0 - x);
        if (penetration < 0.1f && !(largest >= penetration))
        {
            largest = penetration;
            plane = plane_from_point_normal(w0, triangle_normal);
            idx = (int32_t)l;
        }
    }
    else
    {
        const V3 closest_point = mul_point(L.t, local + closest);
        const V3 normal = x - closest_point;
        if (dot(normal, triangle_normal) > 0.0f)
        {
            const float normal_length = len(normal);
            const float penetration = -normal_length;
            if (!(largest >= penetration))
            {
                largest = penetration;
                plane = plane_from_point_normal(closest_point, normal_length > 0.0f ? normal / normal_length : triangle_normal);
                idx = (int32_t)l;
            }
        }
    }
}

// EXT: the kernel variant that also handles triangle trees and scaled convex hulls (their
// walks cost registers that the other variants keep for the level loop).
template <bool EXT>
__device__ __forceinline__ void collide_vertex(const Leaf *leaves, uint32_t num_leaves, V3 x, float &largest, int32_t &idx, Pl &plane)
{
    for (uint32_t l = 0; l < num_leaves; ++l)
    {
        const Leaf &L = leaves[l];
        switch (L.type)
        {
        case 0: // PlaneShape.cpp:243-256
            {
                Pl p; p.n = v3(L.pn[0], L.pn[1], L.pn[2]); p.c = L.pc;
                float pen = -signed_distance(p, x);
                if (!(largest >= pen)) { largest = pen; plane = p; idx = (int32_t)l; }
            }
            break;
        case 1: // SphereShape.cpp:279-301
            {
                V3 center = col3(L.t, 3);
                float radius = L.a;
                V3 delta = x - center;
                float distance = len(delta);
                float pen = radius - distance;
                if (!(largest >= pen))
                {
                    largest = pen;
                    V3 normal = distance > 0.0f ? delta / distance : v3(0.0f, 1.0f, 0.0f);
                    V3 point = center + radius * normal;
                    plane = plane_from_point_normal(point, normal);
                    idx = (int32_t)l;
                }
            }
            break;
        case 2: // BoxShape.cpp:234-283
            {
                V3 he = v3(L.a, L.b, L.c);
                V3 local = mul_point(L.inv, x);
                V3 clamped = vmax(vmin(local, he), neg(he));
                if (clamped.x == local.x && clamped.y == local.y && clamped.z == local.z)
                {
                    V3 delta = he - vabs(local);
                    int index = lowest_index(delta);
                    float pen = comp(delta, index);
                    if (!(largest >= pen))
                    {
                        largest = pen;
                        V3 axis = v3(index == 0 ? 1.0f : 0.0f, index == 1 ? 1.0f : 0.0f, index == 2 ? 1.0f : 0.0f);
                        V3 normal = mulv(vsign(local), axis);
                        V3 point = mulv(normal, he);
                        plane = plane_transformed(plane_from_point_normal(point, normal), L.t);
                        idx = (int32_t)l;
                    }
                }
                else
                {
                    V3 normal = local - clamped;
                    float normal_length = len(normal);
                    float pen = -normal_length;
                    if (!(largest >= pen))
                    {
                        largest = pen;
                        normal = normal / normal_length;
                        plane = plane_transformed(plane_from_point_normal(clamped, normal), L.t);
                        idx = (int32_t)l;
                    }
                }
            }
            break;
        case 3: // CapsuleShape.cpp:324-375
            {
                float hh = L.a, radius = L.b;
                V3 local = mul_point(L.inv, x);
                if (fabsf(local.y) <= hh)
                {
                    V3 normal = v3(local.x, 0.0f, local.z);
                    float normal_length = len(normal);
                    float pen = radius - normal_length;
                    if (!(largest >= pen))
                    {
                        largest = pen;
                        normal = normal_length > 0.0f ? normal / normal_length : v3(1.0f, 0.0f, 0.0f);
                        V3 point = radius * normal;
                        plane = plane_transformed(plane_from_point_normal(point, normal), L.t);
                        idx = (int32_t)l;
                    }
                }
                else
                {
                    V3 center = v3(0.0f, sign_of(local.y) * hh, 0.0f);
                    V3 delta = local - center;
                    float distance = len(delta);
                    float pen = radius - distance;
                    if (!(largest >= pen))
                    {
                        largest = pen;
                        V3 normal = delta / distance;
                        V3 point = center + radius * normal;
                        plane = plane_transformed(plane_from_point_normal(point, normal), L.t);
                        idx = (int32_t)l;
                    }
                }
            }
            break;
        case 4: // CylinderShape.cpp:317-366 (scale 1)
            {
                const float half_height = L.a, radius = L.b;
                const V3 local = mul_point(L.inv, x);
                const V3 side_normal = v3(local.x, 0.0f, local.z);
                const float side_normal_length = len(side_normal);
                const float side_penetration = radius - side_normal_length;
                const float top_penetration = half_height - fabsf(local.y);
                V3 point, normal;
                if (side_penetration < 0.0f && top_penetration < 0.0f)
                {
                    point = side_normal * (radius / side_normal_length) + v3(0.0f, half_height * sign_of(local.y), 0.0f);
                    normal = normalized_or(local - point, v3(0.0f, 1.0f, 0.0f));
                }
                else if (side_penetration < top_penetration)
                {
                    normal = side_normal_length > 0.0f ? side_normal / side_normal_length : v3(1.0f, 0.0f, 0.0f);
                    point = radius * normal;
                }
                else
                {
                    normal = v3(0.0f, sign_of(local.y), 0.0f);
                    point = half_height * normal;
                }
                const Pl p = plane_from_point_normal(point, normal);
                const float pen = -signed_distance(p, local);
                if (!(largest >= pen))
                {
                    largest = pen;
                    plane = plane_transformed(p, L.t);
                    idx = (int32_t)l;
                }
            }
            break;
        case 5: // TaperedCapsuleShape.cpp:304-371 (centres and radii scaled in make_leaf)
            {
                const V3 top_center = v3(0.0f, L.a, 0.0f), bottom_center = v3(0.0f, L.b, 0.0f);
                const float top_radius = L.c, bottom_radius = L.pn[0], sin_alpha = L.pn[1], tan_alpha = L.pn[2];
                const V3 scale_y_flip = v3(1.0f, L.pc, 1.0f);
                const V3 local = mulv(scale_y_flip, mul_point(L.inv, x));
                V3 position, normal;
                const V3 to_top = local - top_center;
                const float to_top_len = len(to_top);
                if (to_top.y >= sin_alpha * to_top_len)
                {
                    normal = to_top_len != 0.0f ? to_top / to_top_len : v3(0.0f, 1.0f, 0.0f);
                    position = top_center + top_radius * normal;
                }
                else
                {
                    const V3 to_bottom = local - bottom_center;
                    const float to_bottom_len = len(to_bottom);
                    if (to_bottom.y <= sin_alpha * to_bottom_len)
                        normal = to_bottom_len != 0.0f ? to_bottom / to_bottom_len : neg(v3(0.0f, 1.0f, 0.0f));
                    else
                    {
                        normal = normalized_or(v3(local.x, 0.0f, local.z), v3(1.0f, 0.0f, 0.0f));
                        normal.y = tan_alpha;
                        normal = normalized_or(normal, v3(1.0f, 0.0f, 0.0f));
                    }
                    position = bottom_center + bottom_radius * normal;
                }
                Pl p = plane_from_point_normal(position, normal);
                const float pen = -signed_distance(p, local);
                if (!(largest >= pen))
                {
                    largest = pen;
                    p.n = mulv(scale_y_flip, p.n);          // a flipped capsule flips the normal's y
                    plane = plane_transformed(p, L.t);
                    idx = (int32_t)l;
                }
            }
            break;
        case 6:
            collide_hull(L, l, x, largest, idx, plane);
            break;
        case 7: // TaperedCylinderShape.cpp:390-506 (top, bottom and radii scaled in make_leaf)
            {
                const float top = L.a, bottom = L.b, top_radius = L.c, bottom_radius = L.pn[0];
                const V3 top_3d = v3(0.0f, top, 0.0f), bottom_3d = v3(0.0f, bottom, 0.0f);
                const V3 local = mul_point(L.inv, x);
                const V3 normal_xz = normalized_or(mulv(v3(1.0f, 0.0f, 1.0f), local), v3(1.0f, 0.0f, 0.0f));
                const float tan_alpha = (bottom_radius - top_radius) / (top - bottom);
                const V3 sn = v3(normal_xz.x, tan_alpha, normal_xz.z);
                const V3 side_normal = sn / len(sn);
                const V3 side_support_top = normal_xz * top_radius + top_3d;
                const float side_penetration = dot(side_support_top - local, side_normal);
                const float top_penetration = top - local.y;
                const float bottom_penetration = local.y - bottom;
                const float min_top_bottom_penetration = smin(top_penetration, bottom_penetration);
                V3 point, normal;
                if (side_penetration < 0.0f || min_top_bottom_penetration < 0.0f)
                {
                    const V3 side_support_bottom = normal_xz * bottom_radius + bottom_3d;
                    const V3 bottom_to_top = side_support_top - side_support_bottom;
                    const float fraction = dot(local - side_support_bottom, bottom_to_top);
                    const float distance_to_axis = dot(normal_xz, local);
                    const bool inside_top_radius = distance_to_axis <= top_radius;
                    const bool inside_bottom_radius = distance_to_axis <= bottom_radius;
                    if (fraction >= len_sq(bottom_to_top) && !inside_top_radius)
                    {
                        point = side_support_top;
                        normal = normalized_or(local - point, v3(0.0f, 1.0f, 0.0f));
                    }
                    else if (fraction < 0.0f && !inside_bottom_radius)
                    {
                        point = side_support_bottom;
                        normal = normalized_or(local - point, v3(0.0f, 1.0f, 0.0f));
                    }
                    else if (top_penetration < 0.0f && inside_top_radius)
                    {
                        point = top_3d;
                        normal = v3(0.0f, 1.0f, 0.0f);
                    }
                    else if (bottom_penetration < 0.0f && inside_bottom_radius)
                    {
                        point = bottom_3d;
                        normal = v3(0.0f, -1.0f, 0.0f);
                    }
                    else
                    {
                        point = side_support_top;
                        normal = side_normal;
                    }
                }
                else if (side_penetration < min_top_bottom_penetration)
                {
                    point = side_support_top;
                    normal = side_normal;
                }
                else if (top_penetration < bottom_penetration)
                {
                    point = top_3d;
                    normal = v3(0.0f, 1.0f, 0.0f);
                }
                else
                {
                    point = bottom_3d;
                    normal = v3(0.0f, -1.0f, 0.0f);
                }
                const Pl p = plane_from_point_normal(point, normal);
                const float pen = -signed_distance(p, local);
                if (!(largest >= pen))
                {
                    largest = pen;
                    plane = plane_transformed(p, L.t);
                    idx = (int32_t)l;
                }
            }
            break;
        case 8:
            if constexpr (EXT)
                collide_tree(L, l, x, largest, idx, plane);
            break;
        case 9:
            if constexpr (EXT)
                collide_hull_scaled(L, l, x, largest, idx, plane);
            break;
        }
    }
}

// ---------------------------------------------------------------- dynamic colliders
__device__ __forceinline__ V3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(float *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// The inverse inertia as a Mat44 with zero translation, so that M * v adds the
// zero column as Mat44 * Vec3 does (Mat44.inl:295-301).
__device__ __forceinline__ M44 inv_inertia_matrix(const DynStateDev &D)
{
    M44 m;
    for (int i = 0; i < 3; ++i)
    {
        m.c[i][0] = D.inv_inertia[3 * i]; m.c[i][1] = D.inv_inertia[3 * i + 1]; m.c[i][2] = D.inv_inertia[3 * i + 2];
        m.c[i][3] = 0.0f;
    }
    m.c[3][0] = 0.0f; m.c[3][1] = 0.0f; m.c[3][2] = 0.0f; m.c[3][3] = 1.0f;
    return m;
}

// The CollidingShape fields of a dynamic body (SoftBodyMotionProperties.cpp:224-233, no
// contact listener: every scale 1): inverse mass, inverse inertia for the relative
// rotation (MotionProperties::GetInverseInertiaForRotation, MotionProperties.inl:66-82,
// all angular DOFs allowed) and the velocities in soft-body space.
__device__ __forceinline__ void init_dynamic(const ColliderDev &c, const M44 &inv_com, const M44 &rel, DynStateDev &D,
                                             float inv_mass_scale, float inv_inertia_scale)
{
    // rotation = rel.Multiply3x3(sRotation(mInertiaRotation)) (Mat44.inl:428-447)
    const M44 ir = rotation_translation(c.inertia_rot[0], c.inertia_rot[1], c.inertia_rot[2], c.inertia_rot[3], v3(0.0f, 0.0f, 0.0f));
    V3 rc[3];
    for (int i = 0; i < 3; ++i)
        rc[i] = (col3(rel, 0) * ir.c[i][0] + col3(rel, 1) * ir.c[i][1]) + col3(rel, 2) * ir.c[i][2];
    // columns of rotation scaled by the inverse inertia diagonal
    const V3 sc[3] = { c.inv_inertia_diag[0] * rc[0], c.inv_inertia_diag[1] * rc[1], c.inv_inertia_diag[2] * rc[2] };
    // rotation.Multiply3x3RightTransposed(scaled) (Mat44.inl:495-507)
    for (int j = 0; j < 3; ++j)
    {
        const V3 col = (rc[0] * comp(sc[0], j) + rc[1] * comp(sc[1], j)) + rc[2] * comp(sc[2], j);
        // times mInvInertiaScale2 (SoftBodyMotionProperties.cpp:215)
        D.inv_inertia[3 * j] = col.x * inv_inertia_scale; D.inv_inertia[3 * j + 1] = col.y * inv_inertia_scale;
        D.inv_inertia[3 * j + 2] = col.z * inv_inertia_scale;
    }
    D.inv_mass = inv_mass_scale * c.inv_mass;
    const V3 lin = mul3x3(inv_com, ld3(c.lin_vel)), ang = mul3x3(inv_com, ld3(c.ang_vel));
    st3(D.lin, lin); st3(D.orig_lin, lin);
    st3(D.ang, ang); st3(D.orig_ang, ang);
    st3(D.com, col3(rel, 3));
    D.updated = 0;
}

// The dynamic-body branch of ApplyCollisionConstraintsAndUpdateVelocities
// (SoftBodyMotionProperties.cpp:752-797) for one vertex in contact: friction and
// restitution impulse shared between the vertex and the body, whose velocities the
// next vertex in contact sees.  x: the vertex after the push out, v: its XPBD velocity.
__device__ __forceinline__ void dynamic_contact(DynStateDev &D, V3 com_t, V3 x, V3 &v, V3 prev_v, V3 n, float projected_distance,
                                                float w, float friction, float restitution, float dt, float restitution_threshold,
                                                float inv_mass_scale1)
{
    const V3 r2 = x - com_t;
    V3 lin = ld3(D.lin), ang = ld3(D.ang);
    const V3 v2 = lin + cross(ang, r2);
    const V3 relative_velocity = v - v2;
    const V3 v_normal = n * dot(n, relative_velocity);
    const V3 v_tangential = relative_velocity - v_normal;
    const float v_tangential_length = len(v_tangential);
    const float vertex_inv_mass = inv_mass_scale1 * w;
    const M44 I = inv_inertia_matrix(D);
    const V3 r2_cross_n = cross(r2, n);
    const float w2 = D.inv_mass + dot(r2_cross_n, mul_point(I, r2_cross_n));
    const float w1_plus_w2 = vertex_inv_mass + w2;
    if (w1_plus_w2 > 0.0f)
    {
        V3 dv;
        if (v_tangential_length > 0.0f)
        {
            const float f = friction * projected_distance / (v_tangential_length * dt);
            dv = v_tangential * (f < 1.0f ? f : 1.0f);
        }
        else
            dv = v3(0.0f, 0.0f, 0.0f);
        dv = dv + v_normal;
        const float prev_v_normal = dot(prev_v - v2, n);
        if (prev_v_normal < restitution_threshold)
            dv = dv + (restitution * prev_v_normal) * n;
        const V3 p = dv / w1_plus_w2;
        v = v - p * vertex_inv_mass;
        lin = lin + p * D.inv_mass;
        ang = ang + mul_point(I, cross(r2, p));
        st3(D.lin, lin);
        st3(D.ang, ang);
        D.updated = 1;
    }
}

// UpdateRigidBodyVelocities for one soft body (SoftBodyMotionProperties.cpp:923-935):
// the velocity changes of the step in world space, added to the call's sums.
__device__ __forceinline__ void finish_dynamic(DynStateDev &D, const M44 &com)
{
    if (D.updated)
    {
        st3(D.acc_lin, ld3(D.acc_lin) + mul3x3(com, ld3(D.lin) - ld3(D.orig_lin)));
        st3(D.acc_ang, ld3(D.acc_ang) + mul3x3(com, ld3(D.ang) - ld3(D.orig_ang)));
        D.any_updated = 1;
        D.updated = 0;
    }
}

// What the contact listener's SoftBodyManifold reports for a vertex at the end of a
// collision step (SoftBodyManifold.h:24-40, read before UpdateSoftBodyState resets the
// collision state): the local contact point x - SignedDistance(x) * n and the collider.
__device__ __forceinline__ float4 contact_report(V3 x, float4 pl, bool has_contact, int32_t collider)
{
    if (!has_contact)
        return make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    Pl plane; plane.n = v3(pl.x, pl.y, pl.z); plane.c = pl.w;
    return f4(x - signed_distance(plane, x) * plane.n, __int_as_float(collider));
}

// DetermineSensorCollisions (SoftBodyMotionProperties.cpp:267-289) for one vertex: the
// largest penetration of the sensor's leaf over the vertices decides its contact.
template <bool EXT>
__device__ __forceinline__ float sensor_penetration(const Leaf &leaf, float4 p)
{
    if (!(p.w > 0.0f))
        return -FLT_MAX;
    float largest = -FLT_MAX;
    int32_t idx = -1;
    Pl plane;
    plane.n = v3(0.0f, 0.0f, 0.0f);
    plane.c = 0.0f;
    collide_vertex<EXT>(&leaf, 1, xyz(p), largest, idx, plane);
    return largest;
}

__device__ __forceinline__ void clear_dynamic(DynStateDev &D)
{
    st3(D.acc_lin, v3(0.0f, 0.0f, 0.0f));
    st3(D.acc_ang, v3(0.0f, 0.0f, 0.0f));
    D.any_updated = 0;
    D.updated = 0;
}

// ScaleHelpers::IsNotScaled and IsInsideOut (ScaleHelpers.h:18-33)
__device__ __forceinline__ bool not_scaled(V3 scale)
{
    return len_sq(scale - v3(1.0f, 1.0f, 1.0f)) <= 1.0e-8f;                       // Vec3::IsClose
}
__device__ __forceinline__ bool inside_out(V3 scale)
{
    return (((scale.x < 0.0f) ? 1 : 0) + ((scale.y < 0.0f) ? 1 : 0) + ((scale.z < 0.0f) ? 1 : 0)) % 2 == 1;
}

// AABox4VsBox (AABox4.h:86-185) for one box b[0..5] (min xyz, max xyz) against the soft body's
// local bounds [lmin, lmax] seen from a compound at rotation q and position p: the oriented box
// OrientedBox(Mat44::sInverseRotationTranslation(q, p), bounds) (CompoundShapeVisitors.h:229,
// Mat44.inl:156-161, OrientedBox.h:27), epsilon 1e-6.  The reference evaluates every axis and
// ANDs the results; stopping at the first separating axis gives the same answer.
__device__ __forceinline__ bool box_overlaps(Q4 q, V3 p, V3 lmin, V3 lmax, const float *b)
{
    const M44 m = rotation_translation(-q.x, -q.y, -q.z, q.w, v3(0.0f, 0.0f, 0.0f));    // sRotation(Conjugated)
    const V3 center = 0.5f * (lmin + lmax), he = 0.5f * (lmax - lmin);                // AABox::GetCenter / GetExtent
    const V3 tr = neg(mul3x3(m, p)) + mul3x3(m, center);                              // PreTranslated(center)
    // R(r, c) = m.c[c][r]; ar[i][j] = |R(j, i)| + epsilon
    float ar[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            ar[i][j] = fabsf(m.c[i][j]) + 1.0e-6f;
    float t[3], a[3];
    for (int i = 0; i < 3; ++i)
    {
        t[i] = comp(tr, i) - 0.5f * (b[i] + b[3 + i]);
        a[i] = 0.5f * (b[3 + i] - b[i]);
    }
    // L = A0, A1, A2
    for (int i = 0; i < 3; ++i)
        if (!(fabsf(t[i]) <= a[i] + ((he.x * ar[0][i] + he.y * ar[1][i]) + he.z * ar[2][i])))
            return false;
    // L = B0, B1, B2
    for (int i = 0; i < 3; ++i)
    {
        const float ra = (a[0] * ar[i][0] + a[1] * ar[i][1]) + a[2] * ar[i][2];
        const float d = (t[0] * m.c[i][0] + t[1] * m.c[i][1]) + t[2] * m.c[i][2];
        if (!(fabsf(d) <= ra + comp(he, i)))
            return false;
    }
    // L = Ai x Bj: ra over the two other axes of A, rb over the two other axes of B
    #define XPBD_CROSS_AXIS(ra, rb, d) if (!(fabsf(d) <= (ra) + (rb))) return false;
    XPBD_CROSS_AXIS(a[1] * ar[0][2] + a[2] * ar[0][1], he.y * ar[2][0] + he.z * ar[1][0], t[2] * m.c[0][1] - t[1] * m.c[0][2])
    XPBD_CROSS_AXIS(a[1] * ar[1][2] + a[2] * ar[1][1], he.x * ar[2][0] + he.z * ar[0][0], t[2] * m.c[1][1] - t[1] * m.c[1][2])
    XPBD_CROSS_AXIS(a[1] * ar[2][2] + a[2] * ar[2][1], he.x * ar[1][0] + he.y * ar[0][0], t[2] * m.c[2][1] - t[1] * m.c[2][2])
    XPBD_CROSS_AXIS(a[0] * ar[0][2] + a[2] * ar[0][0], he.y * ar[2][1] + he.z * ar[1][1], t[0] * m.c[0][2] - t[2] * m.c[0][0])
    XPBD_CROSS_AXIS(a[0] * ar[1][2] + a[2] * ar[1][0], he.x * ar[2][1] + he.z * ar[0][1], t[0] * m.c[1][2] - t[2] * m.c[1][0])
    XPBD_CROSS_AXIS(a[0] * ar[2][2] + a[2] * ar[2][0], he.x * ar[1][1] + he.y * ar[0][1], t[0] * m.c[2][2] - t[2] * m.c[2][0])
    XPBD_CROSS_AXIS(a[0] * ar[0][1] + a[1] * ar[0][0], he.y * ar[2][2] + he.z * ar[1][2], t[1] * m.c[0][0] - t[0] * m.c[0][1])
    XPBD_CROSS_AXIS(a[0] * ar[1][1] + a[1] * ar[1][0], he.x * ar[2][2] + he.z * ar[0][2], t[1] * m.c[1][0] - t[0] * m.c[1][1])
    XPBD_CROSS_AXIS(a[0] * ar[2][1] + a[1] * ar[2][0], he.x * ar[1][2] + he.y * ar[0][2], t[1] * m.c[2][0] - t[0] * m.c[2][1])
    #undef XPBD_CROSS_AXIS
    return true;
}

// The leaf's rotation q and position p relative to the soft body: the collider body's centre of
// mass transform com = (inverse soft body transform * body transform), split into quaternion and
// translation as SoftBodyMotionProperties.cpp:174-188 passes it to CollectTransformedShapes, then
// the part's placement steps (CompoundShapeVisitors.h:258-272, RotatedTranslatedShape.cpp:169-176,
// OffsetCenterOfMassShape.cpp:135-142).  With `cull`, false when a compound's bounds test leaves
// the part out (the soft body's local bounds [lmin, lmax]).
template <bool EXT>
__device__ __forceinline__ bool part_pose(const ColliderDev &c, const M44 &rel, Q4 &q, V3 &p, bool cull, V3 lmin, V3 lmax)
{
    float qf[4];
    get_quaternion(rel, qf);
    q = q4(qf[0], qf[1], qf[2], qf[3]);
    p = col3(rel, 3);
    if constexpr (EXT)
        for (uint32_t k = 0; k < c.num_ops; ++k)
        {
            const PartOpDev &op = c.ops[k];
            switch (op.kind)
            {
            case XPBD_OP_CULL:
                if (cull && !box_overlaps(q, p, lmin, lmax, op.v))
                    return false;
                break;
            case XPBD_OP_TRANSLATE:
                p = p + quat_rotate(q, v3(op.v[0], op.v[1], op.v[2]));
                break;
            case XPBD_OP_ROTATE:
                q = quat_mul(q, q4(op.v[0], op.v[1], op.v[2], op.v[3]));
                break;
            default:    // XPBD_OP_OFFSET
                p = p - quat_rotate(q, v3(op.v[0], op.v[1], op.v[2]));
                break;
            }
        }
    return true;
}

__device__ __forceinline__ M44 collider_relative(const ColliderDev &c, const M44 &inv_com)
{
    return mul(inv_com, rotation_translation(c.rot[0], c.rot[1], c.rot[2], c.rot[3], v3(c.pos[0], c.pos[1], c.pos[2])));
}

// Whether the soft body collects leaf c of an overlapping body this step: always for a simple
// body; for a compound's leaf, when every bounds test on its way passes.
template <bool EXT>
__device__ __forceinline__ bool part_collected(const ColliderDev &c, const M44 &inv_com, V3 lmin, V3 lmax)
{
    if constexpr (EXT)
    {
        if (c.num_ops == 0)
            return true;
        Q4 q;
        V3 p;
        return part_pose<EXT>(c, collider_relative(c, inv_com), q, p, true, lmin, lmax);
    }
    return true;
}

// Leaf of collider c in the soft body's local space (SoftBodyMotionProperties.cpp:174-186:
// com = (inverse soft body transform * collider transform), then the leaf
// transform is rebuilt from com's quaternion and translation, after a compound's placement
// steps).  A dynamic collider also initialises its entry of the body's DynStateDev row.
template <bool EXT>
__device__ __forceinline__ void make_leaf(const ColliderDev &c, const M44 &inv_com, float body_friction, float body_restitution, Leaf &out,
                                          DynStateDev *dyn_row, float4 pair = make_float4(1.0f, 1.0f, 1.0f, 1.0f))
{
    const M44 rel = collider_relative(c, inv_com);
    Q4 q;
    V3 pos;
    part_pose<EXT>(c, rel, q, pos, false, v3(0.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.0f));
    M44 t = rotation_translation(q.x, q.y, q.z, q.w, pos);
    out.t = t;
    out.type = c.type;
    out.dyn = -1;
    out.collider = -1;
    out.inv_mass_scale1 = pair.x;
    if (c.motion == 2 && dyn_row)
    {
        out.dyn = c.dyn;
        init_dynamic(c, inv_com, rel, dyn_row[c.dyn], pair.y, pair.z);
    }
    out.friction = sqrtf(body_friction * c.friction);                                 // ContactConstraintManager.h:554
    out.restitution = body_restitution > c.restitution ? body_restitution : c.restitution; // ContactConstraintManager.h:555
    // the shape's scale (ScaledShape.cpp:142-145), each shape scaling as its
    // CollideSoftBodyVertices does
    const V3 scale = v3(c.scale[0], c.scale[1], c.scale[2]);
    const V3 abs_scale = vabs(scale);                                               // Vec3::Abs
    switch (c.type)
    {
    case 0:
        {
            // Plane::Scaled (Plane.h:49-54) then GetTransformed (PlaneShape.cpp:246)
            const V3 n = v3(c.params[0], c.params[1], c.params[2]);
            const V3 sn = v3(n.x / scale.x, n.y / scale.y, n.z / scale.z);
            const float snl = len(sn);
            Pl p; p.n = sn / snl; p.c = c.params[3] / snl;
            const Pl tp = plane_transformed(p, t);
            out.pn[0] = tp.n.x; out.pn[1] = tp.n.y; out.pn[2] = tp.n.z; out.pc = tp.c;
        }
        break;
    case 1:
        out.a = abs_scale.x * c.params[0];                                          // SphereShape.cpp:53-59
        break;
    case 2:
        out.inv = inversed_rotation_translation(t);
        out.a = abs_scale.x * c.params[0]; out.b = abs_scale.y * c.params[1]; out.c = abs_scale.z * c.params[2];   // BoxShape.cpp:237
        break;
    case 3:
        {
            out.inv = inversed_rotation_translation(t);
            const float s = fabsf(scale.x);                                         // CapsuleShape.cpp:331-333
            out.a = s * c.params[0]; out.b = s * c.params[1];
        }
        break;
    case 4:
        out.inv = inversed_rotation_translation(t);
        out.a = abs_scale.y * c.params[0]; out.b = abs_scale.x * c.params[1];      // CylinderShape.cpp:324-326
        break;
    case 5:
        // TaperedCapsuleShape.cpp:311-318: centres by |scale.y|, radii by |scale.x|, y flip
        out.inv = inversed_rotation_translation(t);
        out.a = abs_scale.y * c.params[0]; out.b = abs_scale.y * c.params[1];
        out.c = abs_scale.x * c.params[2]; out.pn[0] = abs_scale.x * c.params[3];
        out.pn[1] = c.params[4]; out.pn[2] = c.params[5];
        out.pc = sign_of(scale.y);
        break;
    case 7:
        {
            // TaperedCylinderShape::GetScaled (TaperedCylinderShape.cpp:181-199)
            out.inv = inversed_rotation_translation(t);
            float top = scale.y * c.params[0], bottom = scale.y * c.params[1];
            float top_radius = abs_scale.x * c.params[2], bottom_radius = abs_scale.x * c.params[3];
            if (bottom > top)
            {
                const float ft = top; top = bottom; bottom = ft;
                const float fr = top_radius; top_radius = bottom_radius; bottom_radius = fr;
            }
            out.a = top; out.b = bottom; out.c = top_radius; out.pn[0] = bottom_radius;
        }
        break;
    case 8:
        // CollideSoftBodyVerticesVsTriangles (scale, normal sign of an inside-out scale)
        out.inv = inversed_rotation_translation(t);
        out.tree = c.tree;
        out.pn[0] = scale.x; out.pn[1] = scale.y; out.pn[2] = scale.z;
        out.pc = inside_out(scale) ? -1.0f : 1.0f;
        break;
    default:
        out.inv = inversed_rotation_translation(t);
        out.hull = c.hull;
        if (!not_scaled(scale))
        {
            // the scaled branch of ConvexHullShape.cpp:1063-1066, 1091-1106, 1126-1134
            out.type = 9;
            out.pn[0] = scale.x; out.pn[1] = scale.y; out.pn[2] = scale.z;
            out.a = 1.0f / scale.x; out.b = 1.0f / scale.y; out.c = 1.0f / scale.z;   // Vec3::Reciprocal
            out.pc = inside_out(scale) ? -1.0f : 1.0f;
        }
        break;
    }
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) { float u = __shfl_xor(v, o, kWave); v = u < v ? u : v; }
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) { float u = __shfl_xor(v, o, kWave); v = u > v ? u : v; }
    return v;
}
__device__ __forceinline__ float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

} // namespace xpbd
