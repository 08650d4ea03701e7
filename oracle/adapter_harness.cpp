// Test harness of the Jolt-side adapter (joltphysics_amd/adapter/jolt_xpbd_adapter.cpp): built
// into one library with the reference (FMA off) and jolt_oracle.cpp, so a test builds two
// identical worlds with the jo_* functions, lets Jolt simulate the soft bodies of one and
// hands those of the other to the adapter.  Test infrastructure only.

#include "../joltphysics_amd/adapter/jolt_xpbd_adapter.h"

using namespace JPH;

extern "C" void *jo_world_physics_system(void *wp);
extern "C" uint32 jo_world_soft_body_id(void *wp, int body);

extern "C" {

void *ja_create(void *wp, int device)
{
    auto *u = new JoltXpbd::SoftBodyUpdater(*(PhysicsSystem *)jo_world_physics_system(wp), device);
    if (!u->IsValid())
    {
        delete u;
        return nullptr;
    }
    return u;
}

void ja_destroy(void *u) { delete (JoltXpbd::SoftBodyUpdater *)u; }

int ja_add(void *u, void *wp, int body)
{
    return ((JoltXpbd::SoftBodyUpdater *)u)->Add(BodyID(jo_world_soft_body_id(wp, body))) ? 0 : -1;
}

int ja_update(void *u, float dt, int collision_steps)
{
    return ((JoltXpbd::SoftBodyUpdater *)u)->Update(dt, collision_steps) ? 0 : -1;
}

const char *ja_last_error(void *u) { return ((JoltXpbd::SoftBodyUpdater *)u)->LastError().c_str(); }

} // extern "C"
