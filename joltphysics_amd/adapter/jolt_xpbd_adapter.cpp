// Jolt Physics side of the MI355X soft-body update (see jolt_xpbd_adapter.h).  Uses only
// Jolt's public API; everything it hands the device goes through include/jolt_xpbd.h.

#include "jolt_xpbd_adapter.h"
#include "jolt_xpbd_shapes.h"

#include <Jolt/Core/StreamWrapper.h>
#include <Jolt/Physics/Body/BodyLock.h>
#include <Jolt/Physics/Collision/CollisionCollectorImpl.h>
#include <Jolt/Physics/SoftBody/SoftBodyMotionProperties.h>
#include <Jolt/Physics/StateRecorderImpl.h>

#include <cstring>
#include <sstream>

#include "jolt_xpbd.h"

namespace JoltXpbd {

namespace {

// SoftBodySharedSettings::SaveBinaryState writes each array as a uint32 count and the raw
// records (StreamOut.h:33-50), in the order vertices, faces, edges, dihedral bends, volumes,
// skinned constraints, skinned normals, LRAs, update groups (SoftBodySharedSettings.cpp:
// 1203-1213).  The update groups and the skinned normals are private: read them from there.
struct BinaryArrays
{
    std::vector<uint32_t> skinned_normals;
    std::vector<uint32_t> update_groups;        // 7 x uint32 per group
};

bool ReadBinaryArrays(const std::string &data, BinaryArrays &out)
{
    const size_t sizes[] = { 28, 16, 16, 24, 24, 52, 4, 12, 28 };
    size_t at = 0;
    for (int a = 0; a < 9; ++a)
    {
        uint32_t n;
        if (at + 4 > data.size())
            return false;
        memcpy(&n, data.data() + at, 4);
        at += 4;
        const size_t bytes = size_t(n) * sizes[a];
        if (at + bytes > data.size())
            return false;
        if (a == 6)
            out.skinned_normals.assign((const uint32_t *)(data.data() + at), (const uint32_t *)(data.data() + at) + n);
        if (a == 8)
            out.update_groups.assign((const uint32_t *)(data.data() + at), (const uint32_t *)(data.data() + at) + 7 * size_t(n));
        at += bytes;
    }
    return true;
}

// Broadphase hits in query order
class BodyIDCollector : public JPH::CollideShapeBodyCollector
{
public:
    std::vector<JPH::BodyID> mHits;
    void AddHit(const JPH::BodyID &inBodyID) override { mHits.push_back(inBodyID); }
};

} // namespace

SoftBodyUpdater::SoftBodyUpdater(JPH::PhysicsSystem &system, int device) : mSystem(system)
{
    if (xpbd_context_create(device, &mContext) != XPBD_OK)
    {
        mContext = nullptr;
        Fail(xpbd_last_error());
    }
}

SoftBodyUpdater::~SoftBodyUpdater()
{
    for (auto &s : mSettings)
        xpbd_settings_release(s.second);
    if (mContext)
        xpbd_context_destroy(mContext);
}

bool SoftBodyUpdater::Fail(const std::string &msg)
{
    mError = msg;
    return false;
}

xpbd_settings *SoftBodyUpdater::Settings(const JPH::SoftBodySharedSettings &s)
{
    auto it = mSettings.find(&s);
    if (it != mSettings.end())
        return it->second;
    std::stringstream ss;
    {
        JPH::StreamOutWrapper out(ss);
        s.SaveBinaryState(out);
    }
    BinaryArrays arrays;
    if (!ReadBinaryArrays(ss.str(), arrays))
    {
        Fail("cannot read the settings' update groups");
        return nullptr;
    }
    xpbd_settings_desc d = {};
    d.vertices = s.mVertices.data();                      d.num_vertices = (uint32_t)s.mVertices.size();
    d.faces = s.mFaces.data();                            d.num_faces = (uint32_t)s.mFaces.size();
    d.edges = s.mEdgeConstraints.data();                  d.num_edges = (uint32_t)s.mEdgeConstraints.size();
    d.dihedral_bends = s.mDihedralBendConstraints.data(); d.num_dihedral_bends = (uint32_t)s.mDihedralBendConstraints.size();
    d.volumes = s.mVolumeConstraints.data();              d.num_volumes = (uint32_t)s.mVolumeConstraints.size();
    d.lras = s.mLRAConstraints.data();                    d.num_lras = (uint32_t)s.mLRAConstraints.size();
    d.update_groups = arrays.update_groups.data();        d.num_update_groups = (uint32_t)(arrays.update_groups.size() / 7);
    d.num_skinned = (uint32_t)s.mSkinnedConstraints.size();
    d.skinned = s.mSkinnedConstraints.data();
    d.skinned_normals = arrays.skinned_normals.data();    d.num_skinned_normals = (uint32_t)arrays.skinned_normals.size();
    d.inv_bind_matrices = s.mInvBindMatrices.data();      d.num_inv_bind_matrices = (uint32_t)s.mInvBindMatrices.size();
    d.num_rod_stretch_shear = (uint32_t)s.mRodStretchShearConstraints.size();
    d.num_rod_bend_twist = (uint32_t)s.mRodBendTwistConstraints.size();
    d.rod_stretch_shear = s.mRodStretchShearConstraints.data();
    d.rod_bend_twist = s.mRodBendTwistConstraints.data();
    xpbd_settings *out = nullptr;
    if (xpbd_settings_create(mContext, &d, &out) != XPBD_OK)
    {
        Fail(xpbd_last_error());
        return nullptr;
    }
    mSettings[&s] = out;
    return out;
}

bool SoftBodyUpdater::Add(JPH::BodyID id)
{
    if (!mContext)
        return false;
    Managed m;
    m.id = id;
    {
        JPH::BodyLockRead lock(mSystem.GetBodyLockInterface(), id);
        if (!lock.Succeeded() || !lock.GetBody().IsSoftBody())
            return Fail("not a soft body");
        const JPH::Body &b = lock.GetBody();
        const auto *mp = static_cast<const JPH::SoftBodyMotionProperties *>(b.GetMotionProperties());
        xpbd_settings *s = Settings(*mp->GetSettings());
        if (!s)
            return false;
        const JPH::RVec3 p = b.GetPosition();
        const JPH::Quat q = b.GetRotation();
        xpbd_body_desc d = {};
        d.settings = s;
        d.position[0] = float(p.GetX()); d.position[1] = float(p.GetY()); d.position[2] = float(p.GetZ());
        d.rotation[0] = q.GetX(); d.rotation[1] = q.GetY(); d.rotation[2] = q.GetZ(); d.rotation[3] = q.GetW();
        d.num_iterations = mp->GetNumIterations();
        d.linear_damping = mp->GetLinearDamping();
        d.max_linear_velocity = mp->GetMaxLinearVelocity();
        d.restitution = b.GetRestitution();
        d.friction = b.GetFriction();
        d.pressure = mp->GetPressure();
        d.gravity_factor = mp->GetGravityFactor();
        d.vertex_radius = mp->GetVertexRadius();
        d.update_position = mp->GetUpdatePosition() ? 1 : 0;
        d.make_rotation_identity = 0;           // applied by Jolt when it created the body
        d.allow_sleeping = b.GetAllowSleeping() ? 1 : 0;
        if (xpbd_body_create(mContext, &d, &m.body) != XPBD_OK)
            return Fail(xpbd_last_error());
        // the body's whole state (vertices, velocities, bounds, sleep timer, rod and skin states)
        JPH::StateRecorderImpl rec;
        mp->SaveState(rec);
        const std::string data = rec.GetData();
        if (xpbd_body_restore_state(mContext, m.body, data.data(), data.size(), d.position, d.rotation) != XPBD_OK)
            return Fail(xpbd_last_error());
        if (!mp->GetSettings()->mSkinnedConstraints.empty()
            && xpbd_body_set_skin_options(mContext, m.body, mp->GetEnableSkinConstraints() ? 1 : 0,
                                          mp->GetSkinnedMaxDistanceMultiplier()) != XPBD_OK)
            return Fail(xpbd_last_error());
        m.layer = b.GetObjectLayer();
    }
    mSystem.GetBodyInterface().DeactivateBody(id);
    mBodies.push_back(m);
    return true;
}

// DetermineCollidingShapes (SoftBodyMotionProperties.cpp:120-255): the rigid bodies the
// broadphase reports for a soft body's bounds, in query order, that its layer and collision
// group accept; sensors only with a contact listener.  The device repeats the final bounds
// test every collision step, so the query box here only has to contain every box a step of
// this update can ask for: the bodies' bounds grown by the distance a vertex can travel.
bool SoftBodyUpdater::GatherColliders(float dt, int collision_steps)
{
    JPH::AABox query;
    for (const Managed &m : mBodies)
    {
        JPH::BodyLockRead lock(mSystem.GetBodyLockInterface(), m.id);
        if (!lock.Succeeded())
            return Fail("a soft body was removed from the PhysicsSystem");
        const JPH::Body &b = lock.GetBody();
        const auto *mp = static_cast<const JPH::SoftBodyMotionProperties *>(b.GetMotionProperties());
        JPH::AABox box = b.GetWorldSpaceBounds();
        const float reach = mp->GetMaxLinearVelocity() * dt + mp->GetVertexRadius()
                            + mSystem.GetGravity().Length() * mp->GetGravityFactor() * dt * dt;
        box.ExpandBy(JPH::Vec3::sReplicate(reach));
        query.Encapsulate(box);
    }
    BodyIDCollector hits;
    mSystem.GetBroadPhaseQuery().CollideAABox(query, hits);

    const bool listener = mSystem.GetSoftBodyContactListener() != nullptr;
    std::vector<xpbd_collider_desc> desc;
    std::vector<JPH::BodyID> ids;
    std::vector<std::unique_ptr<ShapeData>> shapes;
    std::vector<std::vector<float>> accept(mBodies.size());
    bool any_rejected = false;
    for (JPH::BodyID hit : hits.mHits)
    {
        JPH::BodyLockRead lock(mSystem.GetBodyLockInterface(), hit);
        if (!lock.Succeeded())
            continue;
        const JPH::Body &b = lock.GetBody();
        if (!b.IsRigidBody() || (b.IsSensor() && !listener))
            continue;
        xpbd_collider_desc d = {};
        std::unique_ptr<ShapeData> tables(new ShapeData);
        std::string shape_err;
        if (!DescribeShape(b.GetShape(), d, *tables, shape_err))
            return Fail(shape_err);
        shapes.push_back(std::move(tables));
        const JPH::RVec3 com = b.GetCenterOfMassPosition();
        const JPH::Quat q = b.GetRotation();
        d.position[0] = float(com.GetX()); d.position[1] = float(com.GetY()); d.position[2] = float(com.GetZ());
        d.rotation[0] = q.GetX(); d.rotation[1] = q.GetY(); d.rotation[2] = q.GetZ(); d.rotation[3] = q.GetW();
        d.friction = b.GetFriction();
        d.restitution = b.GetRestitution();
        const JPH::AABox &wb = b.GetWorldSpaceBounds();
        d.world_bounds[0] = wb.mMin.GetX(); d.world_bounds[1] = wb.mMin.GetY(); d.world_bounds[2] = wb.mMin.GetZ();
        d.world_bounds[3] = wb.mMax.GetX(); d.world_bounds[4] = wb.mMax.GetY(); d.world_bounds[5] = wb.mMax.GetZ();
        d.motion_type = (int32_t)b.GetMotionType();
        if (b.IsDynamic())
        {
            const JPH::MotionProperties *mp = b.GetMotionProperties();
            const JPH::Vec3 inv_i = mp->GetInverseInertiaDiagonal();
            const JPH::Quat iq = mp->GetInertiaRotation();
            const JPH::Vec3 lv = b.GetLinearVelocity(), av = b.GetAngularVelocity();
            d.inv_mass = mp->GetInverseMass();
            d.inv_inertia_diagonal[0] = inv_i.GetX(); d.inv_inertia_diagonal[1] = inv_i.GetY(); d.inv_inertia_diagonal[2] = inv_i.GetZ();
            d.inertia_rotation[0] = iq.GetX(); d.inertia_rotation[1] = iq.GetY(); d.inertia_rotation[2] = iq.GetZ(); d.inertia_rotation[3] = iq.GetW();
            d.linear_velocity[0] = lv.GetX(); d.linear_velocity[1] = lv.GetY(); d.linear_velocity[2] = lv.GetZ();
            d.angular_velocity[0] = av.GetX(); d.angular_velocity[1] = av.GetY(); d.angular_velocity[2] = av.GetZ();
        }
        d.is_sensor = b.IsSensor() ? 1 : 0;
        d.user_id = hit.GetIndexAndSequenceNumber();
        // the soft bodies' layer and collision-group filters, per soft body
        for (size_t k = 0; k < mBodies.size(); ++k)
        {
            JPH::BodyLockRead soft(mSystem.GetBodyLockInterface(), mBodies[k].id);
            const bool ok = soft.Succeeded()
                            && mSystem.GetDefaultBroadPhaseLayerFilter(mBodies[k].layer).ShouldCollide(b.GetBroadPhaseLayer())
                            && mSystem.GetDefaultLayerFilter(mBodies[k].layer).ShouldCollide(b.GetObjectLayer())
                            && soft.GetBody().GetCollisionGroup().CanCollide(b.GetCollisionGroup());
            any_rejected |= !ok;
            accept[k].insert(accept[k].end(), { 1.0f, 1.0f, 1.0f, ok ? 1.0f : 0.0f });
        }
        desc.push_back(d);
        ids.push_back(hit);
    }
    if (xpbd_set_colliders(mContext, desc.data(), (uint32_t)desc.size()) != XPBD_OK)
        return Fail(xpbd_last_error());
    for (size_t k = 0; k < mBodies.size(); ++k)
        if (xpbd_body_set_contact_settings(mContext, mBodies[k].body, (uint32_t)desc.size(),
                                           any_rejected ? accept[k].data() : nullptr) != XPBD_OK)
            return Fail(xpbd_last_error());
    mColliders = ids;
    return true;
}

// JobSoftBodyFinalize (PhysicsSystem.cpp:2842-2897): the state back into the body, the new
// position into the body and the broadphase
bool SoftBodyUpdater::WriteBack(const Managed &m)
{
    size_t size = 0;
    if (xpbd_body_save_state(mContext, m.body, nullptr, 0, &size) != XPBD_OK)
        return Fail(xpbd_last_error());
    std::string data(size, '\0');
    if (xpbd_body_save_state(mContext, m.body, &data[0], size, &size) != XPBD_OK)
        return Fail(xpbd_last_error());
    float pos[3], rot[4];
    if (xpbd_body_read(mContext, m.body, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, pos, rot) != XPBD_OK)
        return Fail(xpbd_last_error());
    {
        JPH::BodyLockWrite lock(mSystem.GetBodyLockInterface(), m.id);
        if (!lock.Succeeded())
            return Fail("a soft body was removed from the PhysicsSystem");
        JPH::StateRecorderImpl rec;
        rec.WriteBytes(data.data(), data.size());
        rec.Rewind();
        static_cast<JPH::SoftBodyMotionProperties *>(lock.GetBody().GetMotionProperties())->RestoreState(rec);
        if (rec.IsFailed())
            return Fail("RestoreState failed");
    }
    mSystem.GetBodyInterface().SetPosition(m.id, JPH::RVec3(pos[0], pos[1], pos[2]), JPH::EActivation::DontActivate);
    return true;
}

bool SoftBodyUpdater::Update(float dt, int collision_steps)
{
    if (!mContext)
        return false;
    if (mBodies.empty())
        return true;
    const JPH::Vec3 g = mSystem.GetGravity();
    const float gravity[3] = { g.GetX(), g.GetY(), g.GetZ() };
    if (xpbd_set_gravity(mContext, gravity) != XPBD_OK)
        return Fail(xpbd_last_error());
    if (!GatherColliders(dt, collision_steps))
        return false;
    if (xpbd_update(mContext, dt, collision_steps, 1) != XPBD_OK || xpbd_synchronize(mContext) != XPBD_OK)
        return Fail(xpbd_last_error());
    // UpdateRigidBodyVelocities (SoftBodyMotionProperties.cpp:923-935), body by body
    const uint32_t n = (uint32_t)mColliders.size();
    std::vector<float> lin(3 * size_t(n)), ang(3 * size_t(n));
    std::vector<int32_t> updated(n);
    JPH::BodyInterface &bi = mSystem.GetBodyInterface();
    for (const Managed &m : mBodies)
    {
        if (n && xpbd_body_velocity_deltas(mContext, m.body, n, lin.data(), ang.data(), updated.data()) != XPBD_OK)
            return Fail(xpbd_last_error());
        for (uint32_t c = 0; c < n; ++c)
            if (updated[c])
                bi.AddLinearAndAngularVelocity(mColliders[c], JPH::Vec3(lin[3 * c], lin[3 * c + 1], lin[3 * c + 2]),
                                               JPH::Vec3(ang[3 * c], ang[3 * c + 1], ang[3 * c + 2]));
        if (!WriteBack(m))
            return false;
    }
    return true;
}

} // namespace JoltXpbd
