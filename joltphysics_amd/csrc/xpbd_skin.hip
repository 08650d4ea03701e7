// SoftBodyMotionProperties::SkinVertices on the device (SoftBodyMotionProperties.cpp:1139-1226).
//
// Three stream-ordered launches, each one thread per item:
//   1. skin matrices: joint matrix x inverse bind matrix (Mat44 * Mat44), per inverse bind;
//   2. skinned positions: the previous skinned position keeps the current one, the current
//      one becomes the weighted sum of skin matrix x bind pose position, per skinned constraint;
//   3. skinned normals from the new positions of the constraint's faces (Vec3::NormalizedOr of
//      each face normal, summed, normalised), hard skinning (all skinned vertices, or the
//      kinematic ones when skin constraints are disabled) and the previous-position-valid flag.
// The skin states are the context's device array (4 float4 per vertex: previous and current
// skinned position, normal, {max distance multiplier, back stop radius, enabled, previous
// position valid}); nothing goes through host memory but the joint matrices.  Arithmetic in
// the reference's operation order (jolt_math.h), compiled with -ffp-contract=off.

#include <cfloat>

#include "xpbd_device.h"

namespace xpbd {

__global__ __launch_bounds__(256) void skin_matrices_kernel(const float *__restrict__ joints, const SkinInvBindDev *__restrict__ inv_binds,
                                                            uint32_t n, float *__restrict__ skin)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const SkinInvBindDev ib = inv_binds[k];
    M44 j, m;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
        {
            j.c[c][r] = joints[16u * ib.joint + 4u * c + r];
            m.c[c][r] = ib.m[4 * c + r];
        }
    const M44 s = mul(j, m);
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            skin[16u * k + 4u * c + r] = s.c[c][r];
}

__global__ __launch_bounds__(256) void skin_positions_kernel(const SkinnedDev *__restrict__ recs, uint32_t n, const float *__restrict__ skin,
                                                             float4 *__restrict__ states)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const SkinnedDev sr = recs[k];
    const V3 bind = v3(sr.bind[0], sr.bind[1], sr.bind[2]);
    V3 pos = v3(0.0f, 0.0f, 0.0f);
    for (int w = 0; w < 4; ++w)
    {
        if (sr.weight[w] == 0.0f)
            break;
        const float *s = skin + 16u * sr.inv_bind[w];
        M44 m;
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r)
                m.c[c][r] = s[4 * c + r];
        pos = pos + sr.weight[w] * mul_point(m, bind);
    }
    float4 *e = states + 4u * sr.vertex;
    const float4 cur = e[1];
    e[0] = make_float4(cur.x, cur.y, cur.z, 0.0f);
    e[1] = make_float4(pos.x, pos.y, pos.z, 0.0f);
}

__global__ __launch_bounds__(256) void skin_normals_kernel(const SkinnedDev *__restrict__ recs, uint32_t n, const uint32_t *__restrict__ normal_faces,
                                                           const uint32_t *__restrict__ faces, float4 *__restrict__ states, uint32_t nv,
                                                           float4 *__restrict__ x, float4 *__restrict__ v, int32_t hard_skin_all)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    // mSkinStatePreviousPositionValid for every vertex of the body
    for (uint32_t i = k; i < nv; i += gridDim.x * blockDim.x)
        states[4u * i + 3u].w = 1.0f;
    if (k >= n)
        return;
    const SkinnedDev sr = recs[k];
    V3 normal = v3(0.0f, 0.0f, 0.0f);
    const uint32_t nf = sr.normal_info >> 24;
    if (nf > 0)
    {
        const uint32_t *f = normal_faces + (sr.normal_info & 0xffffffu);
        for (uint32_t q = 0; q < nf; ++q)
        {
            const uint32_t *face = faces + 3u * f[q];
            const V3 v0 = xyz(states[4u * face[0] + 1u]), v1 = xyz(states[4u * face[1] + 1u]), v2 = xyz(states[4u * face[2] + 1u]);
            const V3 c = cross(v1 - v0, v2 - v0);
            const float l2 = len_sq(c);
            normal = normal + (l2 <= FLT_MIN ? v3(0.0f, 0.0f, 0.0f) : c / sqrtf(l2));     // NormalizedOr(zero)
        }
        const float l2 = len_sq(normal);
        normal = l2 <= FLT_MIN ? v3(0.0f, 0.0f, 0.0f) : normal / sqrtf(l2);
    }
    float4 *e = states + 4u * sr.vertex;
    e[2] = make_float4(normal.x, normal.y, normal.z, 0.0f);
    // hard skinning (SoftBodyMotionProperties.cpp:1195-1226): every skinned vertex on request,
    // the kinematic ones (max distance 0) when the skin constraints are disabled
    const bool enabled = states[4u * recs[0].vertex + 3u].z != 0.0f;
    const float4 cur = e[1];
    if (hard_skin_all)
    {
        e[0] = make_float4(cur.x, cur.y, cur.z, 0.0f);
        x[sr.vertex] = make_float4(cur.x, cur.y, cur.z, x[sr.vertex].w);
        v[sr.vertex] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    else if (!enabled && sr.max_distance == 0.0f)
        x[sr.vertex] = make_float4(cur.x, cur.y, cur.z, x[sr.vertex].w);
}

hipError_t launch_skin(const float *joints, const SkinInvBindDev *inv_binds, uint32_t num_inv_binds, float *skin,
                       const SkinnedDev *recs, uint32_t num_recs, const uint32_t *normal_faces, const uint32_t *faces,
                       float4 *states, uint32_t nv, float4 *x, float4 *v, int32_t hard_skin_all, hipStream_t stream)
{
    if (num_inv_binds)
        hipLaunchKernelGGL(skin_matrices_kernel, dim3((num_inv_binds + 255) / 256), dim3(256), 0, stream, joints, inv_binds,
                           num_inv_binds, skin);
    if (num_recs)
        hipLaunchKernelGGL(skin_positions_kernel, dim3((num_recs + 255) / 256), dim3(256), 0, stream, recs, num_recs, skin, states);
    const uint32_t items = num_recs > nv ? num_recs : nv;
    hipLaunchKernelGGL(skin_normals_kernel, dim3((items + 255) / 256), dim3(256), 0, stream, recs, num_recs, normal_faces, faces,
                       states, nv, x, v, hard_skin_all);
    return hipGetLastError();
}

} // namespace xpbd
