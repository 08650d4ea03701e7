// Jolt Physics side of the MI355X soft-body update: a C++ adapter an application compiles
// next to Jolt Physics (v5.5.1) to run its soft bodies through the C ABI of
// include/jolt_xpbd.h instead of PhysicsSystem's soft-body jobs
// (PhysicsSystem.cpp:2701-2897: JobSoftBodyPrepare / Collide / Simulate / Finalize).
//
// Usage, per PhysicsSystem:
//
//     JoltXpbd::SoftBodyUpdater gpu(physics_system);      // one device context
//     gpu.Add(soft_body_id);                               // for each soft body to move
//     ...
//     physics_system.Update(dt, collision_steps, ...);     // rigid bodies (the soft bodies
//                                                          // taken over are inactive in it)
//     gpu.Update(dt, collision_steps);                     // their soft-body update
//
// Add() uploads the body's SoftBodySharedSettings (once per settings object), carries its
// state over with SoftBodyMotionProperties::SaveState and deactivates it in the
// PhysicsSystem.  Update() gathers the colliders the way DetermineCollidingShapes does
// (broadphase order, layer and collision-group filters), runs one xpbd_update, hands the
// velocity changes of dynamic colliders back (UpdateRigidBodyVelocities), and writes every
// body's state back with RestoreState and its new position with BodyInterface::SetPosition
// (JobSoftBodyFinalize), so the rest of Jolt sees the soft bodies as if it had simulated
// them.
//
// Supported colliders: plane, sphere, box, capsule, cylinder, tapered capsule, tapered
// cylinder, convex hull, triangle, mesh and height field bodies, each possibly inside
// ScaledShapes (jolt_xpbd_shapes.h); Update() fails (false, LastError()) when a soft body's
// bounds reach any other shape (compound, rotated / translated, offset centre of mass), and
// the caller keeps the CPU path for it.
#pragma once

#include <Jolt/Jolt.h>
#include <Jolt/Physics/PhysicsSystem.h>
#include <Jolt/Physics/SoftBody/SoftBodySharedSettings.h>

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

struct xpbd_context;
struct xpbd_settings;

namespace JoltXpbd {

class SoftBodyUpdater
{
public:
    explicit SoftBodyUpdater(JPH::PhysicsSystem &system, int device = 0);
    ~SoftBodyUpdater();
    SoftBodyUpdater(const SoftBodyUpdater &) = delete;
    SoftBodyUpdater &operator=(const SoftBodyUpdater &) = delete;

    bool IsValid() const { return mContext != nullptr; }
    const std::string &LastError() const { return mError; }

    /// Moves soft body `id` onto the device.  False (and the body untouched) when its settings
    /// or state cannot run there.
    bool Add(JPH::BodyID id);

    /// The soft-body part of PhysicsSystem::Update(dt, collision_steps) for every body added.
    bool Update(float dt, int collision_steps);

private:
    struct Managed
    {
        JPH::BodyID id;
        int32_t body;               // handle in the device context
        JPH::ObjectLayer layer;
    };

    bool Fail(const std::string &msg);
    xpbd_settings *Settings(const JPH::SoftBodySharedSettings &settings);
    bool GatherColliders(float dt, int collision_steps);
    bool WriteBack(const Managed &m);

    JPH::PhysicsSystem &mSystem;
    xpbd_context *mContext = nullptr;
    std::unordered_map<const JPH::SoftBodySharedSettings *, xpbd_settings *> mSettings;
    std::vector<Managed> mBodies;
    std::vector<JPH::BodyID> mColliders;        // the last collider list, in broadphase order
    std::string mError;
};

} // namespace JoltXpbd
